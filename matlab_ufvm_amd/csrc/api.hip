// api.hip -- the extern "C" entry points of libufvm (declared in include/ufvm.h)
// and the host orchestration of the GPU kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <tuple>

#include "internal.h"
#include "kernels.h"
#include "pcg.h"
#include "state.h"
#include "ufvm.h"

using namespace ufvm;

namespace ufvm {
namespace detail {

System &sys_of(ufvm_ctx *ctx, int32_t id) {
  auto it = ctx->systems.find(id);
  if (it == ctx->systems.end()) throw Error(UFVM_ERR_ARG, "unknown system id");
  return *it->second;
}

RefLevel &level_of(System &S, int32_t level) {
  if (level < 0 || level >= S.nlevels) throw Error(UFVM_ERR_ARG, "level out of range");
  return *S.lev[level];
}

void ensure_partial(ufvm_ctx *ctx, int64_t n, int width) {
  size_t need = (size_t)k::nblocks(std::max<int64_t>(n, 1)) * width + 8;
  if (ctx->partial.n < need) {
    UFVM_HIP(hipStreamSynchronize(ctx->stream));  // the old buffer may still be in use
    ctx->partial.alloc(need);
  }
}

// read device scalar slots [0, count) to host (synchronises the stream)
void read_scal(ufvm_ctx *ctx, int first, int count) {
  UFVM_HIP(hipMemcpyAsync(ctx->hscal + first, ctx->scal.p + first, count * sizeof(double),
                          hipMemcpyDeviceToHost, ctx->stream));
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
}

// ------------------------------------------------------------ reference levels

void upload_pattern(RefLevel &L, hipStream_t s) {
  L.n = L.g.n;
  L.nnz = (int32_t)L.g.col.size();
  // L.g lives as long as the level: the copies need not be waited for here
  L.ptr.upload_async(L.g.ptr, s);
  L.col.upload_async(L.g.col, s);
  L.ac.alloc(L.n);
  L.anb.alloc(std::max(L.nnz, 1));
  L.bc.alloc(L.n);
  L.dphi.alloc(L.n);
  L.r.alloc(L.n);
  L.rc.alloc(L.n);
  L.dinv.alloc(L.n);
  L.sched = false;
  L.low_ok = false;
  L.cs = false;  // a new pattern: schedules and colour chunks are rebuilt on use
}

double residual_l1(ufvm_ctx *ctx, RefLevel &L, double *out_r) {
  ensure_partial(ctx, L.n, 1);
  k::residual_ref(ctx->stream, L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, out_r,
                  ctx->partial.p);
  k::reduce_final(ctx->stream, ctx->partial.p, k::nblocks(L.n), ctx->scal.p);
  read_scal(ctx, 0, 1);
  return ctx->hscal[0];
}

// A level-scheduled sweep runs in one workgroup (a barrier between levels)
// when no level holds more rows than this: two passes of the workgroup per
// level at most, cheaper than a launch per level.
constexpr int32_t kOneWgWidest = 2 * k::kSweepBlock;

static int32_t nlev(const Schedule &S) { return (int32_t)S.level_ptr.size() - 1; }

void ensure_schedule(ufvm_ctx *ctx, RefLevel &L) {
  if (L.sched) return;
  lex_schedule(L.g, true, L.fwd);
  lex_schedule(L.g, false, L.bwd);
  L.fwd_rows.upload(L.fwd.rows, ctx->stream);
  L.bwd_rows.upload(L.bwd.rows, ctx->stream);
  L.fwd_lp.upload(L.fwd.level_ptr, ctx->stream);
  L.bwd_lp.upload(L.bwd.level_ptr, ctx->stream);
  int32_t widest = 0;
  for (const Schedule *S : {&L.fwd, &L.bwd})
    for (size_t q = 0; q + 1 < S->level_ptr.size(); ++q)
      widest = std::max(widest, S->level_ptr[q + 1] - S->level_ptr[q]);
  L.one_wg = widest <= kOneWgWidest;
  L.widest = widest;
  L.max_row = 0;
  for (int32_t i = 0; i < L.n; ++i) L.max_row = std::max(L.max_row, L.g.ptr[i + 1] - L.g.ptr[i]);
  if (widest <= k::kPipeBlock) {
    const int32_t n = L.n, w = std::max(L.max_row, 1);
    L.pipe_w = w;
    std::vector<int32_t> ecol((size_t)n * w), elen(n);
    for (auto [S, C, N] : {std::tuple{&L.fwd, &L.fwd_ecol, &L.fwd_elen},
                           std::tuple{&L.bwd, &L.bwd_ecol, &L.bwd_elen}}) {
      for (int32_t s = 0; s < n; ++s) {
        const int32_t i = S->rows[s], k0 = L.g.ptr[i], len = L.g.ptr[i + 1] - k0;
        elen[s] = len;
        // past the row's length: column 0 for every lane (one cache line per
        // gather instruction; the value is never summed)
        for (int32_t u = 0; u < w; ++u) ecol[(size_t)u * n + s] = u < len ? L.g.col[k0 + u] : 0;
      }
      C->upload(ecol, ctx->stream);
      N->upload(elen, ctx->stream);
    }
    L.fwd_eanb.alloc((size_t)n * w);
    L.bwd_eanb.alloc((size_t)n * w);
    L.fwd_eac.alloc(n);
    L.bwd_eac.alloc(n);
  }
  L.low_ok = lower_lists(L.g, L.low);
  L.lptr.upload(L.low.ptr, ctx->stream);
  L.lnb.upload(L.low.nb, ctx->stream);
  L.lpji.upload(L.low.pos_ji, ctx->stream);
  L.lpij.upload(L.low.pos_ij, ctx->stream);
  L.max_tri = 0;
  if (L.low_ok)
    for (int32_t i = 0; i < L.n; ++i) {
      int32_t up = 0;
      for (int32_t k = L.g.ptr[i]; k < L.g.ptr[i + 1]; ++k) up += L.g.col[k] > i;
      L.max_tri = std::max({L.max_tri, up, L.low.ptr[i + 1] - L.low.ptr[i]});
    }
  if (widest <= k::kPipeBlock && L.low_ok) {
    // DILU ELL copies: forward over the lower lists in forward-schedule
    // order, backward over the entries with column above the row in
    // backward-schedule order
    const int32_t n = L.n;
    auto build = [&](const Schedule &S, bool upper, int32_t &w_out, DBuf<int32_t> &C,
                     DBuf<int32_t> &N, DBuf<int32_t> &E, DBuf<double> &V) {
      std::vector<int32_t> len(n);
      int32_t w = 1;
      for (int32_t s = 0; s < n; ++s) {
        const int32_t i = S.rows[s];
        int32_t m = 0;
        if (upper) {
          for (int32_t k = L.g.ptr[i]; k < L.g.ptr[i + 1]; ++k) m += L.g.col[k] > i;
        } else {
          m = L.low.ptr[i + 1] - L.low.ptr[i];
        }
        len[s] = m;
        w = std::max(w, m);
      }
      std::vector<int32_t> col((size_t)n * w, 0), src((size_t)n * w, -1);
      for (int32_t s = 0; s < n; ++s) {
        const int32_t i = S.rows[s];
        int32_t u = 0;
        if (upper) {
          for (int32_t k = L.g.ptr[i]; k < L.g.ptr[i + 1]; ++k)
            if (L.g.col[k] > i) {
              col[(size_t)u * n + s] = L.g.col[k];
              src[(size_t)u * n + s] = k;
              ++u;
            }
        } else {
          for (int32_t e = L.low.ptr[i]; e < L.low.ptr[i + 1]; ++e, ++u) {
            col[(size_t)u * n + s] = L.low.nb[e];
            src[(size_t)u * n + s] = L.low.pos_ji[e];
          }
        }
      }
      w_out = w;
      C.upload(col, ctx->stream);
      N.upload(len, ctx->stream);
      E.upload(src, ctx->stream);
      V.alloc((size_t)n * w);
    };
    build(L.fwd, false, L.dpipe_wf, L.dfwd_ecol, L.dfwd_elen, L.dfwd_esrc, L.dfwd_eval);
    build(L.bwd, true, L.dpipe_wb, L.dbwd_ecol, L.dbwd_elen, L.dbwd_esrc, L.dbwd_eval);
    L.dilu_y.alloc(n);
    // the factor's copy (cfdFactorizeILU.m:21: the neighbour n-2 is skipped)
    const int32_t skip = n - 2;
    std::vector<int32_t> len(n);
    int32_t w = 1;
    for (int32_t s = 0; s < n; ++s) {
      const int32_t j = L.fwd.rows[s];
      int32_t m = 0;
      for (int32_t e = L.low.ptr[j]; e < L.low.ptr[j + 1]; ++e) m += L.low.nb[e] != skip;
      len[s] = m;
      w = std::max(w, m);
    }
    std::vector<int32_t> col((size_t)n * w, 0), sji((size_t)n * w, -1), sij((size_t)n * w, -1);
    for (int32_t s = 0; s < n; ++s) {
      const int32_t j = L.fwd.rows[s];
      int32_t u = 0;
      for (int32_t e = L.low.ptr[j]; e < L.low.ptr[j + 1]; ++e)
        if (L.low.nb[e] != skip) {
          col[(size_t)u * n + s] = L.low.nb[e];
          sji[(size_t)u * n + s] = L.low.pos_ji[e];
          sij[(size_t)u * n + s] = L.low.pos_ij[e];
          ++u;
        }
    }
    L.fpipe_w = w;
    L.fac_ecol.upload(col, ctx->stream);
    L.fac_elen.upload(len, ctx->stream);
    L.fac_sji.upload(sji, ctx->stream);
    L.fac_sij.upload(sij, ctx->stream);
    L.fac_eji.alloc((size_t)n * w);
    L.fac_eij.alloc((size_t)n * w);
  }
  L.sched = true;
}

// LDS a workgroup may allocate
constexpr size_t kLdsBytes = 65536;

// the device's LDS limit per workgroup (the current device's, read once)
static size_t max_lds_bytes() {
  static const size_t v = [] {
    int dev = 0, m = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&m, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
      m = (int)kLdsBytes;
    return (size_t)m;
  }();
  return v;
}

// the widest schedule level the one-wave staged sweep takes (two passes)
constexpr int32_t kStagedWidest = 128;

void sor_sweep(ufvm_ctx *ctx, RefLevel &L, const double *stop = nullptr) {
  // levels of at most kPipeBlock rows whose unknowns do not fit in LDS: the
  // pipelined sweep (for those that fit, the LDS sweep below was as fast)
  if (L.widest <= k::kPipeBlock && !(L.one_wg && L.n <= k::kLdsRows) &&
      k::gs_sweep_pipe_lds(std::max(nlev(L.fwd), nlev(L.bwd))) <= kLdsBytes) {
    // the ELL values from the current coefficients (a few microseconds
    // against milliseconds of sweeps), then the two sweeps
    const int32_t n = L.n, w = L.pipe_w;
    k::pipe_values(ctx->stream, L.fwd_rows.p, n, w, L.ptr.p, L.anb.p, L.ac.p, L.fwd_eanb.p,
                   L.fwd_eac.p);
    k::pipe_values(ctx->stream, L.bwd_rows.p, n, w, L.ptr.p, L.anb.p, L.ac.p, L.bwd_eanb.p,
                   L.bwd_eac.p);
    k::gs_sweep_pipe(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), n, w, L.fwd_ecol.p,
                     L.fwd_elen.p, L.fwd_eanb.p, L.fwd_eac.p, L.bc.p, L.dphi.p, stop);
    k::gs_sweep_pipe(ctx->stream, L.bwd_rows.p, L.bwd_lp.p, nlev(L.bwd), n, w, L.bwd_ecol.p,
                     L.bwd_elen.p, L.bwd_eanb.p, L.bwd_eac.p, L.bc.p, L.dphi.p, stop);
    return;
  }
  // small levels with narrow schedule levels and short rows: the whole level
  // in LDS, one wave sweeping (falls through if the launch is refused)
  if (L.one_wg && L.widest <= kStagedWidest && L.max_row <= 16 && L.n < 65536 &&
      k::sor_staged_lds(L.n, L.max_row, nlev(L.fwd), nlev(L.bwd)) <= max_lds_bytes() &&
      k::sor_staged(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.bwd_rows.p,
                    L.bwd_lp.p, nlev(L.bwd), L.n, L.max_row, L.ptr.p, L.col.p, L.ac.p, L.anb.p,
                    L.bc.p, L.dphi.p, stop))
    return;
  if (L.one_wg && L.n <= k::kLdsRows) {
    k::sor_lds(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.bwd_rows.p, L.bwd_lp.p,
               nlev(L.bwd), L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, stop);
    return;
  }
  if (L.one_wg) {
    k::gs_sweep_block(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.ptr.p, L.col.p,
                      L.ac.p, L.anb.p, L.bc.p, L.dphi.p, stop);
    k::gs_sweep_block(ctx->stream, L.bwd_rows.p, L.bwd_lp.p, nlev(L.bwd), L.ptr.p, L.col.p,
                      L.ac.p, L.anb.p, L.bc.p, L.dphi.p, stop);
    return;
  }
  const auto &F = L.fwd.level_ptr;
  for (size_t l = 0; l + 1 < F.size(); ++l)
    k::gs_rows(ctx->stream, L.fwd_rows.p + F[l], F[l + 1] - F[l], L.ptr.p, L.col.p, L.ac.p,
               L.anb.p, L.bc.p, L.dphi.p, stop);
  const auto &B = L.bwd.level_ptr;
  for (size_t l = 0; l + 1 < B.size(); ++l)
    k::gs_rows(ctx->stream, L.bwd_rows.p + B[l], B[l + 1] - B[l], L.ptr.p, L.col.p, L.ac.p,
               L.anb.p, L.bc.p, L.dphi.p, stop);
}

// Colour classes of a greedy first-fit colouring (index order) and their
// sliced-ELL chunks (built on first use; the pattern is fixed)
static void ensure_colour_chunks(ufvm_ctx *ctx, RefLevel &L) {
  if (L.cs) return;
  hipStream_t s = ctx->stream;
  std::vector<int32_t> colour;
  const int32_t nc = greedy_colour(L.g, colour);
  std::vector<int32_t> mc_ptr(nc + 1, 0), mc_rows(std::max(L.n, 1), 0);
  for (int32_t c : colour) ++mc_ptr[c + 1];
  for (int32_t c = 0; c < nc; ++c) mc_ptr[c + 1] += mc_ptr[c];
  std::vector<int32_t> at(mc_ptr.begin(), mc_ptr.end() - 1);
  for (int32_t i = 0; i < L.n; ++i) mc_rows[at[colour[i]]++] = i;
  L.cs_cptr.assign(nc + 1, 0);
  std::vector<int32_t> row;
  for (int32_t c = 0; c < nc; ++c) {
    for (int32_t t = mc_ptr[c]; t < mc_ptr[c + 1]; ++t) row.push_back(mc_rows[t]);
    while (row.size() % 64) row.push_back(-1);
    L.cs_cptr[c + 1] = (int32_t)(row.size() / 64);
  }
  const int32_t nq = (int32_t)(row.size() / 64);
  std::vector<int32_t> off(nq + 1, 0);
  for (int32_t q = 0; q < nq; ++q) {
    int32_t w = 0;
    for (int32_t u = 0; u < 64; ++u) {
      const int32_t i = row[(size_t)q * 64 + u];
      if (i >= 0) w = std::max(w, L.g.ptr[i + 1] - L.g.ptr[i]);
    }
    off[q + 1] = off[q] + 64 * w;
  }
  const size_t slots = std::max(off[nq], 1);
  std::vector<int32_t> col(slots, -1), src(slots, -1);
  for (int32_t q = 0; q < nq; ++q) {
    const int32_t w = (off[q + 1] - off[q]) / 64;
    for (int32_t u = 0; u < 64; ++u) {
      const int32_t i = row[(size_t)q * 64 + u];
      if (i < 0) continue;
      for (int32_t e = 0; e < w && L.g.ptr[i] + e < L.g.ptr[i + 1]; ++e) {
        col[off[q] + (size_t)e * 64 + u] = L.g.col[L.g.ptr[i] + e];
        src[off[q] + (size_t)e * 64 + u] = L.g.ptr[i] + e;
      }
    }
  }
  if (row.empty()) row.push_back(-1);
  L.cs_row.upload(row, s);
  L.cs_off.upload(off, s);
  L.cs_col.upload(col, s);
  L.cs_src.upload(src, s);
  L.cs_val.alloc(slots);
  L.cs_slots = (int64_t)nq * 64;
  L.cs = true;
}

// sum of |r|, r = b - A x, over the colour chunks (each row's arithmetic is
// k_residual_ref's; the per-block sums group the rows by colour)
static double residual_l1_chunks(ufvm_ctx *ctx, RefLevel &L) {
  ensure_partial(ctx, (int32_t)std::max<int64_t>(L.cs_slots, 1), 1);
  k::residual_csell(ctx->stream, L.cs_slots, L.cs_row.p, L.cs_off.p, L.cs_col.p, L.cs_val.p,
                    L.ac.p, L.bc.p, L.dphi.p, L.r.p, ctx->partial.p);
  k::reduce_final(ctx->stream, ctx->partial.p, k::nblocks(L.cs_slots), ctx->scal.p);
  read_scal(ctx, 0, 1);
  return ctx->hscal[0];
}

// Symmetric multicolour Gauss-Seidel sweep in the reference numbering: the
// classes of a greedy first-fit colouring (index order) upward, then
// downward without the last class (re-solving it repeats its arithmetic).
// Each row is the reference's Gauss-Seidel row update (k_gs_rows' arithmetic,
// on the colour chunks; cs_val must hold the current anb).
void mcgs_sweep(ufvm_ctx *ctx, RefLevel &L, const double *stop = nullptr) {
  ensure_colour_chunks(ctx, L);
  const int32_t nc = (int32_t)L.cs_cptr.size() - 1;
  auto pass = [&](int32_t c) {
    k::gs_csell(ctx->stream, L.cs_cptr[c], L.cs_cptr[c + 1], L.cs_row.p, L.cs_off.p, L.cs_col.p,
                L.cs_val.p, L.ac.p, L.bc.p, L.dphi.p, stop);
  };
  for (int32_t c = 0; c < nc; ++c) pass(c);
  for (int32_t c = nc - 2; c >= 0; --c) pass(c);
}

void dilu_factor(ufvm_ctx *ctx, RefLevel &L) {
  if (!L.low_ok) throw Error(UFVM_ERR_NUMERIC, "ILU: the index for i in j is not found");
  const auto &F = L.fwd.level_ptr;
  const int32_t skip = L.n - 2;  // cfdFactorizeILU.m:21 skips 1-based row N-1
  if (L.one_wg && L.widest <= kStagedWidest && L.max_tri <= 16 && L.n < 65536 &&
      k::dilu_factor_staged_lds(L.n, L.max_tri, nlev(L.fwd)) <= max_lds_bytes() &&
      k::dilu_factor_staged(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.n, L.max_tri,
                            L.ac.p, L.anb.p, L.lptr.p, L.lnb.p, L.lpji.p, L.lpij.p, skip,
                            L.dinv.p))
    return;
  if (L.widest <= k::kPipeBlock && !(L.one_wg && L.n <= k::kLdsRows) &&
      k::gs_sweep_pipe_lds(nlev(L.fwd)) <= kLdsBytes) {
    const int64_t m = (int64_t)L.n * L.fpipe_w;
    k::pipe_gather(ctx->stream, m, L.fac_sji.p, L.anb.p, L.fac_eji.p);
    k::pipe_gather(ctx->stream, m, L.fac_sij.p, L.anb.p, L.fac_eij.p);
    k::dilu_factor_pipe(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.n, L.fpipe_w,
                        L.fac_ecol.p, L.fac_elen.p, L.fac_eji.p, L.fac_eij.p, L.ac.p, L.dinv.p);
    return;
  }
  if (L.one_wg && L.n <= k::kLdsRows) {
    k::dilu_factor_lds(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.n, L.ac.p, L.anb.p,
                       L.lptr.p, L.lnb.p, L.lpji.p, L.lpij.p, skip, L.dinv.p);
    return;
  }
  if (L.one_wg) {
    k::dilu_factor_block(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.ac.p, L.anb.p,
                         L.lptr.p, L.lnb.p, L.lpji.p, L.lpij.p, skip, L.dinv.p);
    return;
  }
  for (size_t l = 0; l + 1 < F.size(); ++l)
    k::dilu_factor_rows(ctx->stream, L.fwd_rows.p + F[l], F[l + 1] - F[l], L.ac.p, L.anb.p,
                        L.lptr.p, L.lnb.p, L.lpji.p, L.lpij.p, skip, L.dinv.p);
}

void dilu_sweep(ufvm_ctx *ctx, RefLevel &L, const double *stop = nullptr) {
  ensure_partial(ctx, L.n, 1);
  k::residual_ref(ctx->stream, L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, L.rc.p,
                  ctx->partial.p, stop);
  // small levels with narrow schedule levels: the whole level in LDS, one
  // wave substituting (falls through if the launch is refused)
  if (L.low_ok && L.one_wg && L.widest <= kStagedWidest && L.max_tri <= 16 && L.n < 65536 &&
      k::dilu_staged_lds(L.n, L.max_tri, nlev(L.fwd), nlev(L.bwd)) <= max_lds_bytes() &&
      k::dilu_staged(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.bwd_rows.p,
                     L.bwd_lp.p, nlev(L.bwd), L.n, L.max_tri, L.ptr.p, L.col.p, L.anb.p,
                     L.lptr.p, L.lnb.p, L.lpji.p, L.dinv.p, L.rc.p, L.dphi.p, stop))
    return;
  // levels too large for LDS whose schedule levels hold at most kPipeBlock
  // rows: the pipelined substitutions over the ELL copies (values from the
  // current coefficients first)
  if (L.widest <= k::kPipeBlock && L.low_ok && !(L.one_wg && L.n <= k::kLdsRows) &&
      k::gs_sweep_pipe_lds(std::max(nlev(L.fwd), nlev(L.bwd))) <= kLdsBytes) {
    const int32_t n = L.n;
    k::pipe_gather(ctx->stream, (int64_t)n * L.dpipe_wf, L.dfwd_esrc.p, L.anb.p, L.dfwd_eval.p);
    k::pipe_gather(ctx->stream, (int64_t)n * L.dpipe_wb, L.dbwd_esrc.p, L.anb.p, L.dbwd_eval.p);
    k::dilu_pipe(ctx->stream, false, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), n, L.dpipe_wf,
                 L.dfwd_ecol.p, L.dfwd_elen.p, L.dfwd_eval.p, L.dinv.p, L.rc.p, L.dilu_y.p,
                 L.dphi.p, stop);
    k::dilu_pipe(ctx->stream, true, L.bwd_rows.p, L.bwd_lp.p, nlev(L.bwd), n, L.dpipe_wb,
                 L.dbwd_ecol.p, L.dbwd_elen.p, L.dbwd_eval.p, L.dinv.p, L.rc.p, L.dilu_y.p,
                 L.dphi.p, stop);
    return;
  }
  if (L.one_wg && L.n <= k::kLdsRows) {
    k::dilu_lds(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.bwd_rows.p, L.bwd_lp.p,
                nlev(L.bwd), L.n, L.ptr.p, L.col.p, L.anb.p, L.lptr.p, L.lnb.p, L.lpji.p,
                L.dinv.p, L.rc.p, L.dphi.p, stop);
    return;
  }
  if (L.one_wg) {
    k::dilu_fwd_block<double>(ctx->stream, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.anb.p,
                              L.lptr.p, L.lnb.p, L.lpji.p, L.dinv.p, L.rc.p, stop);
    k::dilu_bwd_block<double>(ctx->stream, L.bwd_rows.p, L.bwd_lp.p, nlev(L.bwd), L.n, L.ptr.p,
                              L.col.p, L.anb.p, L.dinv.p, L.rc.p, L.dphi.p, stop);
    return;
  }
  const auto &F = L.fwd.level_ptr;
  for (size_t l = 0; l + 1 < F.size(); ++l)
    k::dilu_fwd_rows(ctx->stream, L.fwd_rows.p + F[l], F[l + 1] - F[l], L.anb.p, L.lptr.p,
                     L.lnb.p, L.lpji.p, L.dinv.p, L.rc.p, stop);
  const auto &B = L.bwd.level_ptr;
  for (size_t l = 0; l + 1 < B.size(); ++l)
    k::dilu_bwd_rows(ctx->stream, L.bwd_rows.p + B[l], B[l + 1] - B[l], L.n, L.ptr.p, L.col.p,
                     L.anb.p, L.dinv.p, L.rc.p, L.dphi.p, stop);
}

// cfdSolveAlgebraicSystem.m:11-66 with the loop test on the device: the
// same sweeps and residuals, each launched n_iter times, and every kernel
// after the stop flag is set returns at once.  A NaN initial residual is
// recorded (alg[3]) and reported by the cycle loop (check_alg_nan).
static void solve_alg_device(ufvm_ctx *ctx, RefLevel &L, int smoother, int n_iter, double rrf,
                             bool chunks) {
  hipStream_t s = ctx->stream;
  double *st = ctx->alg.p;  // allocated and cleared by apply_amg, the only caller
  if (!st) throw Error(UFVM_ERR_STATE, "device-side smoothing test outside an AMG solve");
  // SOR on a small level: the whole call in one launch with the level in LDS
  // (falls through if the launch is refused)
  if (smoother == UFVM_SMOOTHER_SOR && L.one_wg && L.widest <= kStagedWidest &&
      L.max_row <= 16 && L.n < 65536 &&
      k::sor_smooth_staged_lds(L.n, L.max_row, nlev(L.fwd), nlev(L.bwd)) <= max_lds_bytes() &&
      k::sor_smooth_staged(s, L.fwd_rows.p, L.fwd_lp.p, nlev(L.fwd), L.bwd_rows.p, L.bwd_lp.p,
                           nlev(L.bwd), L.n, L.max_row, L.ptr.p, L.col.p, L.ac.p, L.anb.p,
                           L.bc.p, L.dphi.p, L.r.p, n_iter, rrf, st))
    return;
  auto residual = [&](const double *stop) -> int32_t {
    if (chunks) {
      ensure_partial(ctx, (int32_t)std::max<int64_t>(L.cs_slots, 1), 1);
      k::residual_csell(s, L.cs_slots, L.cs_row.p, L.cs_off.p, L.cs_col.p, L.cs_val.p, L.ac.p,
                        L.bc.p, L.dphi.p, L.r.p, ctx->partial.p, stop);
      return k::nblocks(L.cs_slots);
    }
    ensure_partial(ctx, L.n, 1);
    k::residual_ref(s, L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, L.r.p,
                    ctx->partial.p, stop);
    return k::nblocks(L.n);
  };
  k::alg_test(s, ctx->partial.p, residual(nullptr), L.n, rrf, 0, st);
  if (smoother == UFVM_SMOOTHER_ILU) dilu_factor(ctx, L);
  const double *stop = st + 2;
  for (int iter = 1; iter <= n_iter; ++iter) {
    if (smoother == UFVM_SMOOTHER_ILU) dilu_sweep(ctx, L, stop);
    else if (smoother == UFVM_SMOOTHER_SOR) sor_sweep(ctx, L, stop);
    else mcgs_sweep(ctx, L, stop);
    k::alg_test(s, ctx->partial.p, residual(stop), L.n, rrf, 1, st);
  }
}

// cfdSolveAlgebraicSystem.m:11-66
void solve_alg(ufvm_ctx *ctx, RefLevel &L, int smoother, int n_iter, double rrf, double *r0,
               double *r1, bool device_test) {
  if (smoother != UFVM_SMOOTHER_ILU && smoother != UFVM_SMOOTHER_SOR &&
      smoother != UFVM_SMOOTHER_MCGS)
    throw Error(UFVM_ERR_ARG, "smoother must be ILU, SOR or MCGS");
  if (smoother != UFVM_SMOOTHER_MCGS) ensure_schedule(ctx, L);
  const bool chunks = smoother == UFVM_SMOOTHER_MCGS;
  if (chunks) {
    ensure_colour_chunks(ctx, L);
    k::gather_vals<double>(ctx->stream, (int64_t)L.cs_val.n, L.cs_src.p, L.anb.p, L.cs_val.p);
  }
  if (device_test) {
    // a smoothing call inside an AMG cycle: nobody reads its residuals, so
    // the stopping test runs on the device and the host does not wait
    solve_alg_device(ctx, L, smoother, n_iter, rrf, chunks);
    return;
  }
  auto l1 = [&]() { return chunks ? residual_l1_chunks(ctx, L) : residual_l1(ctx, L, L.r.p); };
  const double initial = l1() / L.n;
  double fin = initial;
  if (std::isnan(initial)) throw Error(UFVM_ERR_NUMERIC, "NaN residual");
  if (smoother == UFVM_SMOOTHER_ILU) dilu_factor(ctx, L);
  int iter = 1;
  while (iter <= n_iter && fin > rrf * initial) {
    if (smoother == UFVM_SMOOTHER_ILU) dilu_sweep(ctx, L);
    else if (smoother == UFVM_SMOOTHER_SOR) sor_sweep(ctx, L);
    else mcgs_sweep(ctx, L);
    fin = l1() / L.n;
    ++iter;
  }
  if (r0) *r0 = initial;
  if (r1) *r1 = fin;
}

// cfdAgglomerate.m:7-17
int32_t agglomerate(ufvm_ctx *ctx, System &S, int32_t max_coarse) {
  hipStream_t s = ctx->stream;
  S.pcg.ready = false;
  int32_t iLevel = 1;
  std::vector<double> h_anb;
  std::vector<std::unique_ptr<AggResult>> keep;
  while (iLevel <= max_coarse) {
    iLevel = iLevel + 1;
    RefLevel &F = *S.lev[iLevel - 2];
    h_anb.resize(std::max(F.nnz, 1));
    F.anb.download(h_anb.data(), F.nnz, s);
    UFVM_HIP(hipStreamSynchronize(s));
    AggResult &ag = *keep.emplace_back(new AggResult());
    agglomerate_level(F.g, h_anb.data(), ag);
    F.h_parents = ag.parents;
    // the host arrays stay alive (keep) until the stream is synchronised:
    // the next level's download or the end of this function
    F.parents.upload_async(F.h_parents, s);
    F.child_ptr.upload_async(ag.child_ptr, s);
    F.child.upload_async(ag.child, s);
    {
      std::vector<int32_t> &gptr = ag.gal_ptr, &gsrc = ag.gal_src;
      galerkin_lists(F.g, ag, gptr, gsrc);
      F.gal_out = (int32_t)gptr.size() - 1;
      F.gal_ptr.upload_async(gptr, s);
      F.gal_src.upload_async(gsrc, s);
    }
    if ((int32_t)S.lev.size() < iLevel) S.lev.emplace_back(new RefLevel());
    RefLevel &C = *S.lev[iLevel - 1];
    C.g = std::move(ag.coarse);
    upload_pattern(C, s);
    k::coarse_lhs(s, F.gal_out, C.n, F.gal_ptr.p, F.gal_src.p, F.ac.p, F.anb.p, C.ac.p, C.anb.p);
    UFVM_HIP(hipMemsetAsync(C.bc.p, 0, C.n * sizeof(double), s));
    UFVM_HIP(hipMemsetAsync(C.dphi.p, 0, C.n * sizeof(double), s));
    S.nlevels = iLevel;
    if (ag.nc <= 5) break;
  }
  S.nlevels = iLevel;
  S.hier_valid = true;
  UFVM_HIP(hipStreamSynchronize(s));
  return iLevel;
}

// ---------------------------------------------------------------- AMG cycles

struct Cyc {
  ufvm_ctx *ctx;
  System *S;
  const ufvm_amg_cfg *cfg;
};

RefLevel &L1(Cyc &c, int level) { return *c.S->lev[level - 1]; }  // 1-based

// cfdRestrict.m:5-12
void cyc_restrict(Cyc &c, int level, bool presmooth) {
  RefLevel &F = L1(c, level);
  if (presmooth)
    solve_alg(c.ctx, F, c.cfg->smoother, c.cfg->pre_sweep, 0.1, nullptr, nullptr, true);
  ensure_partial(c.ctx, F.n, 1);
  k::residual_ref(c.ctx->stream, F.n, F.ptr.p, F.col.p, F.ac.p, F.anb.p, F.bc.p, F.dphi.p, F.r.p,
                  c.ctx->partial.p);
  RefLevel &C = L1(c, level + 1);
  k::restrict_ref(c.ctx->stream, C.n, F.child_ptr.p, F.child.p, F.r.p, C.bc.p);
}

// cfdProlongate.m:5-6
void cyc_prolong(Cyc &c, int level) {
  RefLevel &F = L1(c, level - 1);
  RefLevel &C = L1(c, level);
  k::prolong_ref(c.ctx->stream, F.n, F.parents.p, C.dphi.p, F.dphi.p);
  solve_alg(c.ctx, F, c.cfg->smoother, c.cfg->post_sweep, c.cfg->termination, nullptr, nullptr,
            true);
}

void cyc_coarsest(Cyc &c, int level) {
  solve_alg(c.ctx, L1(c, level), c.cfg->smoother, c.cfg->post_sweep, 0.1, nullptr, nullptr,
            true);
}

double cyc_final(Cyc &c) { return residual_l1(c.ctx, L1(c, 1), L1(c, 1).r.p); }

double vcycle(Cyc &c, int G) {  // cfdApplyVCycle.m:5-25
  int g = 1;
  while (g < G) { cyc_restrict(c, g, true); ++g; }
  cyc_coarsest(c, g);
  while (g > 1) { cyc_prolong(c, g); --g; }
  return cyc_final(c);
}

double wcycle(Cyc &c, int G) {  // cfdApplyWCycle.m:3-83
  int g = 1;
  while (g < G) { cyc_restrict(c, g, true); ++g; }
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_restrict(c, g - 1, false);
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_prolong(c, g - 1);
  cyc_restrict(c, g - 2, false);
  cyc_restrict(c, g - 1, true);
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_prolong(c, g - 1);
  cyc_prolong(c, g - 2);
  cyc_restrict(c, g - 3, false);
  cyc_restrict(c, g - 2, true);
  cyc_restrict(c, g - 1, true);
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_prolong(c, g - 1);
  cyc_restrict(c, g - 2, false);
  cyc_restrict(c, g - 1, true);
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_restrict(c, g - 1, false);
  cyc_coarsest(c, g);
  while (g > 1) { cyc_prolong(c, g); --g; }
  return cyc_final(c);
}

double fcycle(Cyc &c, int G) {  // cfdApplyFCycle.m:5-63
  int g = 1;
  while (g < G) { cyc_restrict(c, g, true); ++g; }
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_restrict(c, g - 1, false);
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_prolong(c, g - 1);
  cyc_restrict(c, g - 2, false);
  cyc_restrict(c, g - 1, true);
  cyc_coarsest(c, g);
  cyc_prolong(c, g);
  cyc_prolong(c, g - 1);
  cyc_prolong(c, g - 2);
  cyc_restrict(c, g - 3, false);
  cyc_restrict(c, g - 2, true);
  cyc_restrict(c, g - 1, true);
  cyc_coarsest(c, g);
  while (g > 1) { cyc_prolong(c, g); --g; }
  return cyc_final(c);
}

// cfdApplyAMG.m:9-60
void apply_amg(ufvm_ctx *ctx, System &S, const ufvm_amg_cfg &cfg, double *r0, double *r1,
               int32_t *ncycles) {
  if (cfg.cycle < 0 || cfg.cycle > 2) throw Error(UFVM_ERR_ARG, "bad cycle type");
  if (cfg.smoother != UFVM_SMOOTHER_ILU && cfg.smoother != UFVM_SMOOTHER_SOR)
    throw Error(UFVM_ERR_ARG, "smoother must be ILU or SOR");
  if (cfg.rebuild || !S.hier_valid) agglomerate(ctx, S, cfg.max_coarse_levels);
  // a freshly agglomerated hierarchy has zero coarse dphi (cfdSetupCoefficients.m:17)
  for (int32_t l = 1; l < S.nlevels; ++l)
    UFVM_HIP(hipMemsetAsync(S.lev[l]->dphi.p, 0, S.lev[l]->n * sizeof(double), ctx->stream));
  const int32_t G = S.nlevels;
  if (cfg.cycle != UFVM_CYCLE_V && G < 4)
    throw Error(UFVM_ERR_UNSUPPORTED, "W- and F-cycles need at least 4 levels");
  Cyc c{ctx, &S, &cfg};
  // the cycle's smoothing calls keep their stopping test on the device
  // (solve_alg_device); a NaN initial residual there is reported here
  if (!ctx->alg.p) ctx->alg.alloc(4);
  UFVM_HIP(hipMemsetAsync(ctx->alg.p, 0, 4 * sizeof(double), ctx->stream));
  auto check_nan = [&]() {
    double seen = 0.0;  // on the context's stream: never the null stream
    UFVM_HIP(hipMemcpyAsync(&seen, ctx->alg.p + 3, sizeof(double), hipMemcpyDeviceToHost,
                            ctx->stream));
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    if (seen != 0.0) throw Error(UFVM_ERR_NUMERIC, "NaN residual");
  };
  const double initial = cyc_final(c);
  double fin = initial;
  int32_t nCycle = 1;
  while (nCycle <= cfg.max_cycles && fin > cfg.termination * initial) {
    if (cfg.cycle == UFVM_CYCLE_V) fin = vcycle(c, G);
    else if (cfg.cycle == UFVM_CYCLE_W) fin = wcycle(c, G);
    else fin = fcycle(c, G);
    check_nan();
    ++nCycle;
  }
  if (r0) *r0 = initial;
  if (r1) *r1 = fin;
  if (ncycles) *ncycles = nCycle;
}

// Galerkin coarse operators of an existing hierarchy recomputed from the
// current level-1 coefficients (parents and coarse patterns unchanged)
void galerkin_update(ufvm_ctx *ctx, System &S) {
  if (S.nlevels < 2) throw Error(UFVM_ERR_STATE, "no AMG hierarchy to update");
  for (int32_t l = 0; l + 1 < S.nlevels; ++l) {
    RefLevel &F = *S.lev[l];
    RefLevel &C = *S.lev[l + 1];
    k::coarse_lhs(ctx->stream, F.gal_out, C.n, F.gal_ptr.p, F.gal_src.p, F.ac.p, F.anb.p, C.ac.p,
                  C.anb.p);
  }
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
  S.hier_valid = true;
}

int32_t create_system(ufvm_ctx *ctx, int32_t n, int32_t nnz, const int32_t *ptr,
                      const int32_t *col) {
  if (n <= 0 || nnz < 0 || !ptr || (nnz > 0 && !col))
    throw Error(UFVM_ERR_ARG, "bad system size or null pointer");
  if (ptr[0] != 0 || ptr[n] != nnz) throw Error(UFVM_ERR_ARG, "ptr must run from 0 to nnz");
  for (int32_t i = 0; i < n; ++i)
    if (ptr[i + 1] < ptr[i]) throw Error(UFVM_ERR_ARG, "ptr not monotone");
  for (int32_t e = 0; e < nnz; ++e)
    if (col[e] < 0 || col[e] >= n) throw Error(UFVM_ERR_ARG, "column index out of range");
  std::unique_ptr<System> S(new System());
  S->n = n;
  S->nnz = nnz;
  S->lev.emplace_back(new RefLevel());
  RefLevel &L = *S->lev[0];
  L.g.n = n;
  L.g.ptr.assign(ptr, ptr + n + 1);
  L.g.col.assign(col, col + nnz);
  upload_pattern(L, ctx->stream);
  UFVM_HIP(hipMemsetAsync(L.ac.p, 0, n * sizeof(double), ctx->stream));
  UFVM_HIP(hipMemsetAsync(L.anb.p, 0, std::max(nnz, 1) * sizeof(double), ctx->stream));
  UFVM_HIP(hipMemsetAsync(L.bc.p, 0, n * sizeof(double), ctx->stream));
  UFVM_HIP(hipMemsetAsync(L.dphi.p, 0, n * sizeof(double), ctx->stream));
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
  const int32_t id = ctx->next_id++;
  ctx->systems[id] = std::move(S);
  return id;
}

}  // namespace detail
}  // namespace ufvm

using namespace ufvm::detail;

// ======================================================================= C ABI

extern "C" {

int ufvm_ctx_create(int32_t device, ufvm_ctx **out) {
  if (!out) return UFVM_ERR_ARG;
  *out = nullptr;
  ufvm_ctx *ctx = new (std::nothrow) ufvm_ctx();
  if (!ctx) return UFVM_ERR_ARG;
  ctx->device = device;
  int rc = guard(ctx, [&] {
    int ndev = 0;
    UFVM_HIP(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw Error(UFVM_ERR_ARG, "no such device");
    UFVM_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->scal.alloc(kScal);
    UFVM_HIP(hipMemsetAsync(ctx->scal.p, 0, kScal * sizeof(double), ctx->stream));
    ctx->counter.alloc(4);
    UFVM_HIP(hipMemsetAsync(ctx->counter.p, 0, 4 * sizeof(unsigned), ctx->stream));
    {
      std::lock_guard<std::mutex> g(device_mgmt_mutex());
      UFVM_HIP(hipHostMalloc((void **)&ctx->hscal, kScal * sizeof(double), hipHostMallocDefault));
    }
    ctx->partial.alloc(1024);
    noalloc_setup(device);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
  if (rc != UFVM_OK) {
    ufvm_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return UFVM_OK;
}

int ufvm_ctx_destroy(ufvm_ctx *ctx) {
  if (!ctx) return UFVM_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  ctx->cases.clear();
  ctx->systems.clear();
  for (auto &r : ctx->prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : ctx->event_pool) (void)hipEventDestroy(e);
  ctx->partial.release();
  ctx->scal.release();
  ctx->counter.release();
  {
    std::lock_guard<std::mutex> g(device_mgmt_mutex());
    if (ctx->hscal) (void)hipHostFree(ctx->hscal);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
  return UFVM_OK;
}

const char *ufvm_last_error(const ufvm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void *ufvm_ctx_stream(ufvm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int ufvm_ctx_synchronize(ufvm_ctx *ctx) {
  return guard(ctx, [&] {
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    std::lock_guard<std::mutex> g(device_mgmt_mutex());
    UFVM_HIP(hipDeviceSynchronize());
  });
}

int ufvm_system_create(ufvm_ctx *ctx, int32_t n, int32_t nnz, const int32_t *ptr,
                       const int32_t *col, int32_t *sys_out) {
  return guard(ctx, [&] {
    if (!sys_out) throw Error(UFVM_ERR_ARG, "null sys_out");
    *sys_out = create_system(ctx, n, nnz, ptr, col);
  });
}

int ufvm_system_destroy(ufvm_ctx *ctx, int32_t sys) {
  return guard(ctx, [&] {
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    if (!ctx->systems.erase(sys)) throw Error(UFVM_ERR_ARG, "unknown system id");
  });
}

static void set_coeffs(ufvm_ctx *ctx, int32_t sys, const double *ac, const double *anb,
                       const double *bc, hipMemcpyKind kind) {
  System &S = sys_of(ctx, sys);
  RefLevel &L = *S.lev[0];
  if (ac) UFVM_HIP(hipMemcpyAsync(L.ac.p, ac, L.n * sizeof(double), kind, ctx->stream));
  if (anb && L.nnz)
    UFVM_HIP(hipMemcpyAsync(L.anb.p, anb, L.nnz * sizeof(double), kind, ctx->stream));
  if (bc) UFVM_HIP(hipMemcpyAsync(L.bc.p, bc, L.n * sizeof(double), kind, ctx->stream));
  UFVM_HIP(hipMemsetAsync(L.dphi.p, 0, L.n * sizeof(double), ctx->stream));
  if (ac || anb) {
    S.hier_valid = false;
    S.pcg.ready = false;
    S.nlevels = 1;
  }
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
}

int ufvm_system_set_coefficients(ufvm_ctx *ctx, int32_t sys, const double *ac, const double *anb,
                                 const double *bc) {
  return guard(ctx, [&] { set_coeffs(ctx, sys, ac, anb, bc, hipMemcpyHostToDevice); });
}

int ufvm_system_set_coefficients_dev(ufvm_ctx *ctx, int32_t sys, const double *ac,
                                     const double *anb, const double *bc) {
  return guard(ctx, [&] { set_coeffs(ctx, sys, ac, anb, bc, hipMemcpyDeviceToDevice); });
}

int ufvm_system_update_coefficients(ufvm_ctx *ctx, int32_t sys, const double *ac,
                                    const double *anb, const double *bc) {
  return guard(ctx, [&] {
    System &S = sys_of(ctx, sys);
    if (!S.hier_valid || S.nlevels < 2)
      throw Error(UFVM_ERR_STATE, "no AMG hierarchy: call ufvm_amg_agglomerate");
    RefLevel &L = *S.lev[0];
    hipStream_t s = ctx->stream;
    if (ac) UFVM_HIP(hipMemcpyAsync(L.ac.p, ac, L.n * sizeof(double), hipMemcpyHostToDevice, s));
    if (anb && L.nnz)
      UFVM_HIP(hipMemcpyAsync(L.anb.p, anb, L.nnz * sizeof(double), hipMemcpyHostToDevice, s));
    if (bc) UFVM_HIP(hipMemcpyAsync(L.bc.p, bc, L.n * sizeof(double), hipMemcpyHostToDevice, s));
    UFVM_HIP(hipMemsetAsync(L.dphi.p, 0, L.n * sizeof(double), s));
    galerkin_update(ctx, S);
    if (S.pcg.ready) pcg_refresh(ctx, S);
  });
}

int ufvm_system_set_dphi(ufvm_ctx *ctx, int32_t sys, int32_t level, const double *dphi) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    if (!dphi) throw Error(UFVM_ERR_ARG, "null dphi");
    UFVM_HIP(hipMemcpyAsync(L.dphi.p, dphi, L.n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_system_get_dphi(ufvm_ctx *ctx, int32_t sys, int32_t level, double *dphi) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    if (!dphi) throw Error(UFVM_ERR_ARG, "null dphi");
    L.dphi.download(dphi, L.n, ctx->stream);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_system_set_bc(ufvm_ctx *ctx, int32_t sys, int32_t level, const double *bc) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    if (!bc) throw Error(UFVM_ERR_ARG, "null bc");
    UFVM_HIP(hipMemcpyAsync(L.bc.p, bc, L.n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_system_get_bc(ufvm_ctx *ctx, int32_t sys, int32_t level, double *bc) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    if (!bc) throw Error(UFVM_ERR_ARG, "null bc");
    L.bc.download(bc, L.n, ctx->stream);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_compute_residuals(ufvm_ctx *ctx, int32_t sys, int32_t level, double *r_out,
                           double *l1_out) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    const double l1 = residual_l1(ctx, L, L.r.p);
    if (r_out) {
      L.r.download(r_out, L.n, ctx->stream);
      UFVM_HIP(hipStreamSynchronize(ctx->stream));
    }
    if (l1_out) *l1_out = l1;
  });
}

int ufvm_amg_agglomerate(ufvm_ctx *ctx, int32_t sys, int32_t max_coarse_levels, int32_t *nlevels) {
  return guard(ctx, [&] {
    if (max_coarse_levels < 1 || max_coarse_levels > 30)
      throw Error(UFVM_ERR_ARG, "maxCoarseLevels must be in [1, 30]");
    System &S = sys_of(ctx, sys);
    const int32_t nl = agglomerate(ctx, S, max_coarse_levels);
    if (nlevels) *nlevels = nl;
  });
}

int ufvm_amg_num_levels(ufvm_ctx *ctx, int32_t sys, int32_t *nlevels) {
  return guard(ctx, [&] {
    System &S = sys_of(ctx, sys);
    if (nlevels) *nlevels = S.nlevels;
  });
}

int ufvm_amg_level_info(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *n, int32_t *nnz) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    if (n) *n = L.n;
    if (nnz) *nnz = L.nnz;
  });
}

int ufvm_amg_level_get(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *parents, int32_t *ptr,
                       int32_t *col, double *ac, double *anb) {
  return guard(ctx, [&] {
    System &S = sys_of(ctx, sys);
    RefLevel &L = level_of(S, level);
    if (parents) {
      if (level >= S.nlevels - 1) throw Error(UFVM_ERR_ARG, "the coarsest level has no parents");
      std::copy(L.h_parents.begin(), L.h_parents.end(), parents);
    }
    if (ptr) std::copy(L.g.ptr.begin(), L.g.ptr.end(), ptr);
    if (col) std::copy(L.g.col.begin(), L.g.col.end(), col);
    if (ac) L.ac.download(ac, L.n, ctx->stream);
    if (anb) L.anb.download(anb, L.nnz, ctx->stream);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_solve_algebraic_system(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t smoother,
                                int32_t n_iter, double rrf, double *r0, double *r1) {
  return guard(ctx, [&] {
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    solve_alg(ctx, L, smoother, n_iter, rrf, r0, r1);
  });
}

int ufvm_apply_amg(ufvm_ctx *ctx, int32_t sys, const ufvm_amg_cfg *cfg, double *r0, double *r1,
                   int32_t *ncycles) {
  return guard(ctx, [&] {
    if (!cfg) throw Error(UFVM_ERR_ARG, "null cfg");
    apply_amg(ctx, sys_of(ctx, sys), *cfg, r0, r1, ncycles);
  });
}

int ufvm_pcg_setup(ufvm_ctx *ctx, int32_t sys, const ufvm_pcg_cfg *cfg) {
  return guard(ctx, [&] {
    if (!cfg) throw Error(UFVM_ERR_ARG, "null cfg");
    pcg_setup(ctx, sys_of(ctx, sys), *cfg);
  });
}

int ufvm_pcg_solve(ufvm_ctx *ctx, int32_t sys, int32_t *iters, double *relres) {
  return guard(ctx, [&] { pcg_solve(ctx, sys_of(ctx, sys), iters, relres); });
}

int ufvm_pcg_solve_batch(ufvm_ctx *ctx, int32_t sys, int32_t k, const double *b, double *x,
                         int32_t *iters, double *relres) {
  return guard(ctx, [&] { pcg_solve_batch(ctx, sys_of(ctx, sys), k, b, x, iters, relres); });
}

int ufvm_pcg_solve_group(ufvm_ctx *ctx, int32_t g, const int32_t *sys, const double *const *b,
                         double *const *x, int32_t *iters, double *relres) {
  return guard(ctx, [&] {
    if (g < 1 || !sys) throw Error(UFVM_ERR_ARG, "a group needs g >= 1 systems");
    std::vector<System *> v;
    for (int32_t q = 0; q < g; ++q) v.push_back(&sys_of(ctx, sys[q]));
    pcg_solve_group(ctx, v, b, x, iters, relres);
  });
}

int ufvm_pcg_apply_vcycle(ufvm_ctx *ctx, int32_t sys, const double *r, double *z) {
  return guard(ctx, [&] { pcg_apply_vcycle(ctx, sys_of(ctx, sys), r, z); });
}

int ufvm_pcg_level_colours(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *ncolours,
                           int32_t *colour) {
  return guard(ctx, [&] {
    System &S = sys_of(ctx, sys);
    if (!S.pcg.ready) throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first");
    if (level < 0 || level >= S.pcg.nlev) throw Error(UFVM_ERR_ARG, "level out of range");
    FastLevel &L = *S.pcg.lev[level];
    if (ncolours) *ncolours = L.ncol;
    if (colour) std::copy(L.colour.begin(), L.colour.end(), colour);
  });
}

int ufvm_pcg_level_layout(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *index_bytes,
                          int64_t *slots) {
  return guard(ctx, [&] {
    System &S = sys_of(ctx, sys);
    if (!S.pcg.ready) throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first");
    if (level < 0 || level >= S.pcg.nlev) throw Error(UFVM_ERR_ARG, "level out of range");
    const FastLevel &L = *S.pcg.lev[level];
    if (index_bytes) *index_bytes = L.sd16.p ? 2 : (L.scol.p ? 4 : 0);
    if (slots) *slots = L.h_soff.empty() ? 0 : L.h_soff.back();
  });
}

int ufvm_prof_enable(ufvm_ctx *ctx, int32_t on) {
  return guard(ctx, [&] {
    if (on < 0) throw Error(UFVM_ERR_ARG, "sampling stride must be >= 0");
    ctx->prof_on = on != 0;
    if (on) ctx->prof_stride = on;
  });
}

int ufvm_prof_reset(ufvm_ctx *ctx) {
  return guard(ctx, [&] {
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    for (auto &r : ctx->prof) {
      ctx->event_pool.push_back(r.a);
      ctx->event_pool.push_back(r.b);
    }
    ctx->prof.clear();
    ctx->prof_seen.clear();
  });
}

int ufvm_prof_read(ufvm_ctx *ctx, const char *kind, double *total_ms, int64_t *launches,
                   double *bytes) {
  return guard(ctx, [&] {
    if (!kind) throw Error(UFVM_ERR_ARG, "null kind");
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    double ms = 0.0, by = 0.0;
    int64_t cnt = 0;
    for (auto &r : ctx->prof) {
      if (r.kind != kind) continue;
      float t = 0.f;
      UFVM_HIP(hipEventElapsedTime(&t, r.a, r.b));
      ms += t;
      by += r.bytes;
      cnt += r.launches;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = cnt;
    if (bytes) *bytes = by;
  });
}

}  // extern "C"
