// internal.h -- host helpers shared by the translation units of libufvm.
#pragma once
#include <hip/hip_runtime.h>

#include "state.h"

namespace ufvm {
namespace detail {

constexpr int kScal = 32;   // device scalar slots per context

// run f, translating exceptions into a status code and ctx->err
template <typename F>
int guard(ufvm_ctx *ctx, F &&f) {
  if (!ctx) return UFVM_ERR_ARG;
  try {
    UFVM_HIP(hipSetDevice(ctx->device));
    f();
    ctx->err.clear();
    return UFVM_OK;
  } catch (const Error &e) {
    ctx->err = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    ctx->err = "host out of memory";
    return UFVM_ERR_ARG;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return UFVM_ERR_ARG;
  }
}

// event from the context's pool (created on first use)
inline hipEvent_t pool_event(ufvm_ctx *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  UFVM_HIP(hipEventCreate(&e));
  return e;
}

// HIP events around one launch of a kernel class on stream s while profiling
// is enabled on the context (every prof_stride-th launch of each class)
struct ProfScope {
  ufvm_ctx *ctx;
  hipStream_t s;
  bool on;
  ProfRec rec;
  ProfScope(ufvm_ctx *c, hipStream_t st, const char *kind, double bytes)
      : ctx(c), s(st), on(c->prof_on && kind && *kind) {
    if (!on) return;
    // sampled launches only: fewer event records keep the host's launch queue
    // ahead of the GPU, so an interval holds the kernel and not a launch wait
    on = (ctx->prof_seen[kind]++ % ctx->prof_stride) == 0;
    if (!on) return;
    rec.kind = kind;
    rec.bytes = bytes;
    rec.a = pool_event(ctx);
    rec.b = pool_event(ctx);
    (void)hipEventRecord(rec.a, s);
  }
  // a scope spanning several back-to-back launches of one kernel class:
  // begin with bytes 0, then add(bytes) once per launch
  void add(double b) {
    if (!on) return;
    if (!added) rec.launches = 0;
    added = true;
    rec.bytes += b;
    ++rec.launches;
  }
  bool added = false;
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.b, s);
    ctx->prof.push_back(rec);
  }
};

System &sys_of(ufvm_ctx *ctx, int32_t id);
RefLevel &level_of(System &S, int32_t level);
void noalloc_setup(int device);
void ensure_schedule(ufvm_ctx *ctx, RefLevel &L);
void ensure_partial(ufvm_ctx *ctx, int64_t n, int width);
void read_scal(ufvm_ctx *ctx, int first, int count);
int32_t agglomerate(ufvm_ctx *ctx, System &S, int32_t max_coarse);
void solve_alg(ufvm_ctx *ctx, RefLevel &L, int smoother, int n_iter, double rrf, double *r0,
               double *r1, bool device_test = false);
void apply_amg(ufvm_ctx *ctx, System &S, const ufvm_amg_cfg &cfg, double *r0, double *r1,
               int32_t *ncycles);
void pcg_setup(ufvm_ctx *ctx, System &S, const ufvm_pcg_cfg &cfg);
void pcg_refresh(ufvm_ctx *ctx, System &S);
void pcg_solve(ufvm_ctx *ctx, System &S, int32_t *iters, double *relres);
void galerkin_update(ufvm_ctx *ctx, System &S);
int32_t create_system(ufvm_ctx *ctx, int32_t n, int32_t nnz, const int32_t *ptr,
                      const int32_t *col);

}  // namespace detail
}  // namespace ufvm
