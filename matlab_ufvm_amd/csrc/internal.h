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

System &sys_of(ufvm_ctx *ctx, int32_t id);
void ensure_partial(ufvm_ctx *ctx, int64_t n, int width);
void read_scal(ufvm_ctx *ctx, int first, int count);
int32_t agglomerate(ufvm_ctx *ctx, System &S, int32_t max_coarse);
void solve_alg(ufvm_ctx *ctx, RefLevel &L, int smoother, int n_iter, double rrf, double *r0,
               double *r1);
void apply_amg(ufvm_ctx *ctx, System &S, const ufvm_amg_cfg &cfg, double *r0, double *r1,
               int32_t *ncycles);
void pcg_setup(ufvm_ctx *ctx, System &S, const ufvm_pcg_cfg &cfg);
void pcg_solve(ufvm_ctx *ctx, System &S, int32_t *iters, double *relres);
void galerkin_update(ufvm_ctx *ctx, System &S);
int32_t create_system(ufvm_ctx *ctx, int32_t n, int32_t nnz, const int32_t *ptr,
                      const int32_t *col);

}  // namespace detail
}  // namespace ufvm
