// api.hip -- the extern "C" entry points of libufvm (declared in include/ufvm.h)
// and the host orchestration of the GPU kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>

#include "internal.h"
#include "kernels.h"
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

hipEvent_t pool_event(ufvm_ctx *ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  UFVM_HIP(hipEventCreate(&e));
  return e;
}

struct ProfScope {
  ufvm_ctx *ctx;
  bool on;
  ProfRec rec;
  ProfScope(ufvm_ctx *c, const char *kind, double bytes) : ctx(c), on(c->prof_on) {
    if (!on) return;
    // sampled launches only: fewer event records keep the host's launch queue
    // ahead of the GPU, so an interval holds the kernel and not a launch wait
    on = (ctx->prof_seen[kind]++ % ctx->prof_stride) == 0;
    if (!on) return;
    rec.kind = kind;
    rec.bytes = bytes;
    rec.a = pool_event(ctx);
    rec.b = pool_event(ctx);
    (void)hipEventRecord(rec.a, ctx->stream);
  }
  ~ProfScope() {
    if (!on) return;
    (void)hipEventRecord(rec.b, ctx->stream);
    ctx->prof.push_back(rec);
  }
};

// ------------------------------------------------------------ reference levels

void upload_pattern(RefLevel &L, hipStream_t s) {
  L.n = L.g.n;
  L.nnz = (int32_t)L.g.col.size();
  L.ptr.upload(L.g.ptr, s);
  L.col.upload(L.g.col, s);
  L.ac.alloc(L.n);
  L.anb.alloc(std::max(L.nnz, 1));
  L.bc.alloc(L.n);
  L.dphi.alloc(L.n);
  L.r.alloc(L.n);
  L.rc.alloc(L.n);
  L.dinv.alloc(L.n);
  L.sched = false;
  L.low_ok = false;
}

double residual_l1(ufvm_ctx *ctx, RefLevel &L, double *out_r) {
  ensure_partial(ctx, L.n, 1);
  k::residual_ref(ctx->stream, L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, out_r,
                  ctx->partial.p);
  k::reduce_final(ctx->stream, ctx->partial.p, k::nblocks(L.n), ctx->scal.p);
  read_scal(ctx, 0, 1);
  return ctx->hscal[0];
}

void ensure_schedule(ufvm_ctx *ctx, RefLevel &L) {
  if (L.sched) return;
  lex_schedule(L.g, true, L.fwd);
  lex_schedule(L.g, false, L.bwd);
  L.fwd_rows.upload(L.fwd.rows, ctx->stream);
  L.bwd_rows.upload(L.bwd.rows, ctx->stream);
  L.low_ok = lower_lists(L.g, L.low);
  L.lptr.upload(L.low.ptr, ctx->stream);
  L.lnb.upload(L.low.nb, ctx->stream);
  L.lpji.upload(L.low.pos_ji, ctx->stream);
  L.lpij.upload(L.low.pos_ij, ctx->stream);
  L.sched = true;
}

void sor_sweep(ufvm_ctx *ctx, RefLevel &L) {
  const auto &F = L.fwd.level_ptr;
  for (size_t l = 0; l + 1 < F.size(); ++l)
    k::gs_rows(ctx->stream, L.fwd_rows.p + F[l], F[l + 1] - F[l], L.ptr.p, L.col.p, L.ac.p,
               L.anb.p, L.bc.p, L.dphi.p);
  const auto &B = L.bwd.level_ptr;
  for (size_t l = 0; l + 1 < B.size(); ++l)
    k::gs_rows(ctx->stream, L.bwd_rows.p + B[l], B[l + 1] - B[l], L.ptr.p, L.col.p, L.ac.p,
               L.anb.p, L.bc.p, L.dphi.p);
}

void dilu_factor(ufvm_ctx *ctx, RefLevel &L) {
  if (!L.low_ok) throw Error(UFVM_ERR_NUMERIC, "ILU: the index for i in j is not found");
  const auto &F = L.fwd.level_ptr;
  const int32_t skip = L.n - 2;  // cfdFactorizeILU.m:21 skips 1-based row N-1
  for (size_t l = 0; l + 1 < F.size(); ++l)
    k::dilu_factor_rows(ctx->stream, L.fwd_rows.p + F[l], F[l + 1] - F[l], L.ac.p, L.anb.p,
                        L.lptr.p, L.lnb.p, L.lpji.p, L.lpij.p, skip, L.dinv.p);
}

void dilu_sweep(ufvm_ctx *ctx, RefLevel &L) {
  ensure_partial(ctx, L.n, 1);
  k::residual_ref(ctx->stream, L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, L.rc.p,
                  ctx->partial.p);
  const auto &F = L.fwd.level_ptr;
  for (size_t l = 0; l + 1 < F.size(); ++l)
    k::dilu_fwd_rows(ctx->stream, L.fwd_rows.p + F[l], F[l + 1] - F[l], L.anb.p, L.lptr.p,
                     L.lnb.p, L.lpji.p, L.dinv.p, L.rc.p);
  const auto &B = L.bwd.level_ptr;
  for (size_t l = 0; l + 1 < B.size(); ++l)
    k::dilu_bwd_rows(ctx->stream, L.bwd_rows.p + B[l], B[l + 1] - B[l], L.n, L.ptr.p, L.col.p,
                     L.anb.p, L.dinv.p, L.rc.p, L.dphi.p);
}

// cfdSolveAlgebraicSystem.m:11-66
void solve_alg(ufvm_ctx *ctx, RefLevel &L, int smoother, int n_iter, double rrf, double *r0,
               double *r1) {
  if (smoother != UFVM_SMOOTHER_ILU && smoother != UFVM_SMOOTHER_SOR)
    throw Error(UFVM_ERR_ARG, "smoother must be ILU or SOR");
  ensure_schedule(ctx, L);
  const double initial = residual_l1(ctx, L, L.r.p) / L.n;
  double fin = initial;
  if (std::isnan(initial)) throw Error(UFVM_ERR_NUMERIC, "NaN residual");
  if (smoother == UFVM_SMOOTHER_ILU) dilu_factor(ctx, L);
  int iter = 1;
  while (iter <= n_iter && fin > rrf * initial) {
    if (smoother == UFVM_SMOOTHER_ILU) dilu_sweep(ctx, L);
    else sor_sweep(ctx, L);
    fin = residual_l1(ctx, L, L.r.p) / L.n;
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
  while (iLevel <= max_coarse) {
    iLevel = iLevel + 1;
    RefLevel &F = *S.lev[iLevel - 2];
    h_anb.resize(std::max(F.nnz, 1));
    F.anb.download(h_anb.data(), F.nnz, s);
    UFVM_HIP(hipStreamSynchronize(s));
    AggResult ag;
    agglomerate_level(F.g, h_anb.data(), ag);
    F.h_parents = ag.parents;
    F.parents.upload(ag.parents, s);
    F.child_ptr.upload(ag.child_ptr, s);
    F.child.upload(ag.child, s);
    F.slot.upload(ag.entry_slot.empty() ? std::vector<int32_t>{-1} : ag.entry_slot, s);
    if ((int32_t)S.lev.size() < iLevel) S.lev.emplace_back(new RefLevel());
    RefLevel &C = *S.lev[iLevel - 1];
    C.g = std::move(ag.coarse);
    upload_pattern(C, s);
    k::coarse_lhs(s, C.n, F.child_ptr.p, F.child.p, F.ptr.p, F.ac.p, F.anb.p, F.slot.p, C.ptr.p,
                  C.ac.p, C.anb.p);
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
  if (presmooth) solve_alg(c.ctx, F, c.cfg->smoother, c.cfg->pre_sweep, 0.1, nullptr, nullptr);
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
  solve_alg(c.ctx, F, c.cfg->smoother, c.cfg->post_sweep, c.cfg->termination, nullptr, nullptr);
}

void cyc_coarsest(Cyc &c, int level) {
  solve_alg(c.ctx, L1(c, level), c.cfg->smoother, c.cfg->post_sweep, 0.1, nullptr, nullptr);
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

// ------------------------------------------------------- precision-form AMG

// Precision-form copy of hierarchy level l.  kind 0 (level 1): unknowns
// renumbered so every colour class of a global greedy colouring is one
// contiguous row range.  kind 1 (coarse levels): blocks of kBRows consecutive
// reference rows, each sorted by a colouring of the block's own subgraph.
// kind 2: the coarsest level, kept in reference order for the dense solve.
// kind 3 (coarse levels smoothed by damped Jacobi): rows sorted by length
// inside windows.  Rows keep their cconn order in every kind.
void build_fast_level(ufvm_ctx *ctx, System &S, int l, int kind) {
  hipStream_t s = ctx->stream;
  RefLevel &R = *S.lev[l];
  auto &FLp = S.pcg.lev[l];
  FLp.reset(new FastLevel());
  FastLevel &FL = *FLp;
  const int32_t n = R.n;
  FL.n = n;
  FL.nnz = R.nnz;
  FL.kind = kind;
  std::vector<double> ac(n), anb(std::max(R.nnz, 1));
  R.ac.download(ac.data(), n, s);
  R.anb.download(anb.data(), R.nnz, s);
  UFVM_HIP(hipStreamSynchronize(s));
  FL.perm.resize(n);
  if (kind == 0) {
    FL.ncol = greedy_colour(R.g, FL.colour);
    FL.col_off.assign(FL.ncol + 1, 0);
    for (int32_t i = 0; i < n; ++i) ++FL.col_off[FL.colour[i] + 1];
    for (int c = 0; c < FL.ncol; ++c) FL.col_off[c + 1] += FL.col_off[c];
    std::vector<int32_t> fillp(FL.col_off.begin(), FL.col_off.end() - 1);
    for (int32_t i = 0; i < n; ++i) FL.perm[fillp[FL.colour[i]]++] = i;
  } else if (kind == 1) {
    const int32_t B = k::kBRows;
    FL.ncol = greedy_colour_blocks(R.g, B, FL.colour);
    FL.bncol = FL.ncol;
    FL.nblk = (n + B - 1) / B;
    FL.h_bptr.assign(FL.nblk + 1, 0);
    FL.h_boff.assign((size_t)FL.nblk * (FL.bncol + 1), 0);
    for (int32_t q = 0; q < FL.nblk; ++q) {
      const int32_t lo = q * B, hi = std::min(n, lo + B);
      FL.h_bptr[q] = lo;
      int32_t *off = &FL.h_boff[(size_t)q * (FL.bncol + 1)];
      std::vector<int32_t> cnt(FL.bncol + 1, 0);
      for (int32_t i = lo; i < hi; ++i) ++cnt[FL.colour[i] + 1];
      off[0] = lo;
      for (int c = 0; c < FL.bncol; ++c) off[c + 1] = off[c] + cnt[c + 1];
      std::vector<int32_t> fillp(off, off + FL.bncol);
      for (int32_t i = lo; i < hi; ++i) FL.perm[fillp[FL.colour[i]]++] = i;
    }
    FL.h_bptr[FL.nblk] = n;
    FL.bptr.upload(FL.h_bptr, s);
    FL.boff.upload(FL.h_boff, s);
  } else if (kind == 3) {
    // Jacobi-smoothed: row order is free, so rows are sorted by decreasing
    // length inside windows of kSigma rows; the 64-row sliced-ELL chunks then
    // carry little padding while neighbours stay close in memory
    constexpr int32_t kSigma = 256;
    FL.ncol = 1;
    FL.colour.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) FL.perm[i] = i;
    auto len = [&](int32_t i) { return R.g.ptr[i + 1] - R.g.ptr[i]; };
    for (int32_t w0 = 0; w0 < n; w0 += kSigma)
      std::stable_sort(FL.perm.begin() + w0, FL.perm.begin() + std::min(n, w0 + kSigma),
                       [&](int32_t a, int32_t b) { return len(a) > len(b); });
  } else {
    // kind 2 (dense coarsest) and kind 4 (DILU finest level): reference order
    FL.ncol = 1;
    FL.colour.assign(n, 0);
    for (int32_t i = 0; i < n; ++i) FL.perm[i] = i;
  }
  std::vector<int32_t> iperm(n);
  for (int32_t t = 0; t < n; ++t) iperm[FL.perm[t]] = t;
  FL.h_ptr.assign(n + 1, 0);
  std::vector<int32_t> col(std::max(R.nnz, 1));
  std::vector<double> a64(std::max(R.nnz, 1)), d64(n);
  for (int32_t t = 0; t < n; ++t) {
    const int32_t i = FL.perm[t];
    FL.h_ptr[t + 1] = FL.h_ptr[t] + (R.g.ptr[i + 1] - R.g.ptr[i]);
    int32_t o = FL.h_ptr[t];
    for (int32_t kk = R.g.ptr[i]; kk < R.g.ptr[i + 1]; ++kk, ++o) {
      col[o] = iperm[R.g.col[kk]];
      a64[o] = anb[kk];
    }
    d64[t] = ac[i];
  }
  std::vector<float> a32(a64.begin(), a64.end()), d32(d64.begin(), d64.end());
  FL.ptr.upload(FL.h_ptr, s);
  FL.col.upload(col, s);
  FL.perm_d.upload(FL.perm, s);
  FL.ac64.upload(d64, s);
  FL.anb64.upload(a64, s);
  FL.ac32.upload(d32, s);
  FL.anb32.upload(a32, s);
  FL.b32.alloc(n); FL.x32.alloc(n); FL.r32.alloc(n); FL.y32.alloc(n);
  FL.b64.alloc(n); FL.x64.alloc(n); FL.r64.alloc(n); FL.y64.alloc(n);
  if (kind == 4) {
    // the reference's DILU factor (cfdFactorizeILU.m:10-47, row N-1 quirk
    // kept) in FP64 over the lexicographic schedule, stored in both precisions
    ensure_schedule(ctx, R);
    if (!R.low_ok) throw Error(UFVM_ERR_NUMERIC, "ILU: the index for i in j is not found");
    const auto &F = R.fwd.level_ptr;
    for (size_t q = 0; q + 1 < F.size(); ++q)
      k::dilu_factor_rows(s, R.fwd_rows.p + F[q], F[q + 1] - F[q], R.ac.p, R.anb.p, R.lptr.p,
                          R.lnb.p, R.lpji.p, R.lpij.p, R.n - 2, R.dinv.p);
    FL.dinv64.alloc(n);
    FL.dinv32.alloc(n);
    k::convert<double, double>(s, n, R.dinv.p, FL.dinv64.p);
    k::convert<double, float>(s, n, R.dinv.p, FL.dinv32.p);
  }
  // iterates start finite: the sliced-ELL padding multiplies x[row] by 0
  for (DBuf<float> *v : {&FL.b32, &FL.x32, &FL.r32, &FL.y32})
    UFVM_HIP(hipMemsetAsync(v->p, 0, n * sizeof(float), s));
  for (DBuf<double> *v : {&FL.b64, &FL.x64, &FL.r64, &FL.y64})
    UFVM_HIP(hipMemsetAsync(v->p, 0, n * sizeof(double), s));
  // the FPCG's operator product runs on level 0's sliced-ELL copy whatever its
  // kind (a single-level hierarchy makes level 0 the dense coarsest level)
  if (kind == 0 || kind == 3 || l == 0) {
    const int32_t nq = (n + 63) / 64;
    FL.h_soff.assign(nq + 1, 0);
    for (int32_t q = 0; q < nq; ++q) {
      int32_t w = 0;
      for (int32_t t = q * 64; t < std::min(n, q * 64 + 64); ++t)
        w = std::max(w, FL.h_ptr[t + 1] - FL.h_ptr[t]);
      FL.h_soff[q + 1] = FL.h_soff[q] + 64 * w;
    }
    const size_t tot = std::max(FL.h_soff[nq], 1);
    // padding: col = the row itself, value 0 (rows past n in the last chunk
    // are never processed)
    std::vector<int32_t> scol(tot, 0);
    std::vector<double> sv64(tot, 0.0);
    for (int32_t t = 0; t < n; ++t) {
      const int32_t q = t / 64, base = FL.h_soff[q] + (t & 63);
      const int32_t w = (FL.h_soff[q + 1] - FL.h_soff[q]) / 64;
      const int32_t len = FL.h_ptr[t + 1] - FL.h_ptr[t];
      for (int32_t e = 0; e < w; ++e) {
        scol[base + e * 64] = e < len ? col[FL.h_ptr[t] + e] : t;
        sv64[base + e * 64] = e < len ? a64[FL.h_ptr[t] + e] : 0.0;
      }
    }
    std::vector<float> sv32(sv64.begin(), sv64.end());
    FL.soff.upload(FL.h_soff, s);
    FL.scol.upload(scol, s);
    FL.sval64.upload(sv64, s);
    FL.sval32.upload(sv32, s);
  }
  UFVM_HIP(hipStreamSynchronize(s));
}

void link_levels(ufvm_ctx *ctx, System &S, int l) {
  hipStream_t s = ctx->stream;
  RefLevel &R = *S.lev[l];
  FastLevel &F = *S.pcg.lev[l];
  FastLevel &C = *S.pcg.lev[l + 1];
  std::vector<int32_t> ipf(F.n), ipc(C.n);
  for (int32_t t = 0; t < F.n; ++t) ipf[F.perm[t]] = t;
  for (int32_t t = 0; t < C.n; ++t) ipc[C.perm[t]] = t;
  std::vector<int32_t> parent(F.n);
  for (int32_t t = 0; t < F.n; ++t) parent[t] = ipc[R.h_parents[F.perm[t]]];
  // children of coarse (new) row q: the reference children of perm_c[q], which
  // are listed in increasing reference fine index (the reference summation order)
  std::vector<int32_t> cnt(C.n + 1, 0);
  for (int32_t i = 0; i < F.n; ++i) ++cnt[ipc[R.h_parents[i]] + 1];
  for (int32_t q = 0; q < C.n; ++q) cnt[q + 1] += cnt[q];
  std::vector<int32_t> child(F.n);
  std::vector<int32_t> fillc(cnt.begin(), cnt.end() - 1);
  for (int32_t i = 0; i < F.n; ++i) child[fillc[ipc[R.h_parents[i]]]++] = ipf[i];
  // the same lists in 64-row chunks, column-major, padded with -1
  const int32_t nq = (C.n + 63) / 64;
  std::vector<int32_t> soff(nq + 1, 0);
  for (int32_t q = 0; q < nq; ++q) {
    int32_t w = 0;
    for (int32_t p = q * 64; p < std::min(C.n, q * 64 + 64); ++p) w = std::max(w, cnt[p + 1] - cnt[p]);
    soff[q + 1] = soff[q] + 64 * w;
  }
  std::vector<int32_t> schild(std::max(soff[nq], 1), -1);
  for (int32_t p = 0; p < C.n; ++p)
    for (int32_t e = cnt[p]; e < cnt[p + 1]; ++e)
      schild[soff[p / 64] + (p & 63) + (e - cnt[p]) * 64] = child[e];
  F.parent.upload(parent, s);
  F.csoff.upload(soff, s);
  F.cschild.upload(schild, s);
  UFVM_HIP(hipStreamSynchronize(s));
}

template <typename V>
struct VecSel;
template <>
struct VecSel<float> {
  static const float *sval(FastLevel &L) { return L.sval32.p; }
  static const float *dinv(FastLevel &L) { return L.dinv32.p; }
  static float *b(FastLevel &L) { return L.b32.p; }
  static float *x(FastLevel &L) { return L.x32.p; }
  static float *y(FastLevel &L) { return L.y32.p; }
  static float *r(FastLevel &L) { return L.r32.p; }
  static const float *ac(FastLevel &L) { return L.ac32.p; }
  static const float *anb(FastLevel &L) { return L.anb32.p; }
};
template <>
struct VecSel<double> {
  static const double *sval(FastLevel &L) { return L.sval64.p; }
  static const double *dinv(FastLevel &L) { return L.dinv64.p; }
  static double *b(FastLevel &L) { return L.b64.p; }
  static double *x(FastLevel &L) { return L.x64.p; }
  static double *y(FastLevel &L) { return L.y64.p; }
  static double *r(FastLevel &L) { return L.r64.p; }
  static const double *ac(FastLevel &L) { return L.ac64.p; }
  static const double *anb(FastLevel &L) { return L.anb64.p; }
};

// Symmetric multicolour Gauss-Seidel sweeps on the finest level (kind 0):
// colours upward then downward; the downward pass skips the last colour and
// every sweep after the first skips colour 0, because re-solving the colour
// just solved repeats identical arithmetic.  from_zero: x holds no data yet;
// first_done: the first pass (colour 0 from zero) was already made by the
// kernel that produced b.
template <typename V>
void mcgs_sweeps(ufvm_ctx *ctx, FastLevel &L, int nsweep, bool from_zero, bool first_done) {
  using Sel = VecSel<V>;
  const double w = sizeof(V);
  auto pass = [&](int c, int32_t zero_above) {
    const int32_t r0 = L.col_off[c], r1 = L.col_off[c + 1];
    double bytes = 0.0;
    if (ctx->prof_on) {
      // algorithmic bytes: the colour's CSR entries (col + value), its ac, b,
      // x write, and one read of every other unknown
      const double R = r1 - r0, Z = L.h_ptr[r1] - L.h_ptr[r0];
      bytes = (4.0 + w) * Z + 3.0 * w * R + w * (L.n - R);
    }
    ProfScope ps(ctx, "mcgs0", bytes);
    k::mcgs_sell<V>(ctx->stream, r0, r1, L.soff.p, L.scol.p, Sel::sval(L), Sel::ac(L), Sel::b(L),
                    Sel::x(L), zero_above);
  };
  for (int sw = 0; sw < nsweep; ++sw) {
    for (int c = (sw == 0 ? (first_done ? 1 : 0) : 1); c < L.ncol; ++c)
      pass(c, (sw == 0 && from_zero) ? L.col_off[c + 1] : L.n);
    for (int c = L.ncol - 2; c >= 0; --c) pass(c, L.n);
  }
}

// One sweep of the reference's DILU smoother (cfdSolveILU.m:14-93) on a
// kind-4 level: rc = b - A x (rc = b from zero), forward substitution, then
// backward substitution with x += z, each over the lexicographic level
// schedule of the reference level R.
template <typename V>
void dilu_sweep_fast(ufvm_ctx *ctx, RefLevel &R, FastLevel &L, V *x, bool from_zero) {
  using Sel = VecSel<V>;
  hipStream_t s = ctx->stream;
  V *rc = Sel::r(L);
  if (from_zero) {
    k::fill<V>(s, L.n, x, V(0));
    k::convert<V, V>(s, L.n, Sel::b(L), rc);
  } else {
    k::residual_sell<V>(s, L.n, 0, L.soff.p, L.scol.p, Sel::sval(L), Sel::ac(L), Sel::b(L), x,
                        rc);
  }
  const auto &F = R.fwd.level_ptr;
  for (size_t q = 0; q + 1 < F.size(); ++q)
    k::dilu_fwd_rows<V>(s, R.fwd_rows.p + F[q], F[q + 1] - F[q], Sel::anb(L), R.lptr.p, R.lnb.p,
                        R.lpji.p, Sel::dinv(L), rc);
  const auto &B = R.bwd.level_ptr;
  for (size_t q = 0; q + 1 < B.size(); ++q)
    k::dilu_bwd_rows<V>(s, R.bwd_rows.p + B[q], B[q + 1] - B[q], L.n, L.ptr.p, L.col.p,
                        Sel::anb(L), Sel::dinv(L), rc, x);
}

// Jacobi levels with at most this many rows fuse the prolongation into the
// first post-smoothing sweep.
constexpr int32_t kFusedProlongRows = 100000;

// One V-cycle from level l down; returns the level's solution vector.
// first_done (finest level only): x already holds the first colour pass.
template <typename V>
V *vcycle_fast(ufvm_ctx *ctx, System &S, int l, bool first_done = false) {
  using Sel = VecSel<V>;
  PcgState &P = S.pcg;
  FastLevel &L = *P.lev[l];
  hipStream_t s = ctx->stream;
  const int pre = P.cfg.pre_sweep, post = P.cfg.post_sweep;
  if (l == P.nlev - 1) {
    k::dense_solve<V>(s, L.n, P.ainv.p, Sel::b(L), Sel::x(L));
    return Sel::x(L);
  }
  FastLevel &C = *P.lev[l + 1];
  V *x = Sel::x(L);
  const V omega = V(2.0 / 3.0);
  if (L.kind == 3) {
    // x already holds one Jacobi sweep from zero (fused into the restriction
    // that produced b); further sweeps ping-pong between x and y
    V *y = Sel::y(L);
    for (int it = 1; it < pre; ++it) {
      k::jacobi_sell<V>(s, L.n, L.soff.p, L.scol.p, Sel::sval(L), Sel::ac(L), Sel::b(L), x,
                        nullptr, nullptr, y, omega);
      std::swap(x, y);
    }
    if (pre == 0) k::fill<V>(s, L.n, x, V(0));
    k::residual_sell<V>(s, L.n, 0, L.soff.p, L.scol.p, Sel::sval(L), Sel::ac(L), Sel::b(L), x,
                        Sel::r(L));
    if (C.kind == 3 && pre > 0)
      k::restrict_jacobi<V>(s, C.n, L.csoff.p, L.cschild.p, Sel::r(L), Sel::ac(C), Sel::b(C),
                            Sel::x(C), omega);
    else
      k::restrict_fast<V>(s, C.n, L.csoff.p, L.cschild.p, Sel::r(L), Sel::b(C));
    V *xc = vcycle_fast<V>(ctx, S, l + 1);
    // small levels are latency-bound: the first post-sweep adds the coarse
    // correction on the fly instead of a separate prolongation launch
    const bool fuse = post > 0 && L.n <= kFusedProlongRows;
    if (!fuse) k::prolong_fast<V>(s, 0, L.n, L.parent.p, xc, x);
    for (int it = 0; it < post; ++it) {
      const bool pro = fuse && it == 0;
      k::jacobi_sell<V>(s, L.n, L.soff.p, L.scol.p, Sel::sval(L), Sel::ac(L), Sel::b(L), x,
                        pro ? L.parent.p : nullptr, pro ? xc : nullptr, y, omega);
      std::swap(x, y);
    }
    return x;
  }
  if (L.kind == 4) {
    RefLevel &R = *S.lev[l];
    if (pre == 0) k::fill<V>(s, L.n, x, V(0));
    for (int it = 0; it < pre; ++it) dilu_sweep_fast<V>(ctx, R, L, x, it == 0);
    k::residual_sell<V>(s, L.n, 0, L.soff.p, L.scol.p, Sel::sval(L), Sel::ac(L), Sel::b(L), x,
                        Sel::r(L));
    if (C.kind == 3 && pre > 0)
      k::restrict_jacobi<V>(s, C.n, L.csoff.p, L.cschild.p, Sel::r(L), Sel::ac(C), Sel::b(C),
                            Sel::x(C), omega);
    else
      k::restrict_fast<V>(s, C.n, L.csoff.p, L.cschild.p, Sel::r(L), Sel::b(C));
    V *xc = vcycle_fast<V>(ctx, S, l + 1);
    k::prolong_fast<V>(s, 0, L.n, L.parent.p, xc, x);
    for (int it = 0; it < post; ++it) dilu_sweep_fast<V>(ctx, R, L, x, false);
    return x;
  }
  if (L.kind == 0) {
    if (pre > 0) mcgs_sweeps<V>(ctx, L, pre, true, first_done);
    else k::fill<V>(s, L.n, x, V(0));
  } else {
    if (pre > 0)
      k::bmcgs<V>(s, L.nblk, L.bptr.p, L.boff.p, L.bncol, L.ptr.p, L.col.p, Sel::ac(L),
                  Sel::anb(L), Sel::b(L), x, nullptr, nullptr, x, 0, pre);
    else k::fill<V>(s, L.n, x, V(0));
  }
  // after a symmetric Gauss-Seidel sweep the rows of colour 0 (solved last)
  // have a residual of rounding noise only: it is taken as 0
  if (L.kind == 0)
    k::residual_sell<V>(s, L.n, pre > 0 ? L.col_off[1] : 0, L.soff.p, L.scol.p, Sel::sval(L),
                        Sel::ac(L), Sel::b(L), x, Sel::r(L));
  else
    k::residual_fast<V>(s, L.n, L.ptr.p, L.col.p, Sel::ac(L), Sel::anb(L), Sel::b(L), x,
                        Sel::r(L));
  if (C.kind == 3 && pre > 0)
    k::restrict_jacobi<V>(s, C.n, L.csoff.p, L.cschild.p, Sel::r(L), Sel::ac(C), Sel::b(C),
                          Sel::x(C), omega);
  else
    k::restrict_fast<V>(s, C.n, L.csoff.p, L.cschild.p, Sel::r(L), Sel::b(C));
  V *xc = vcycle_fast<V>(ctx, S, l + 1);
  if (L.kind == 1 && post > 0) {
    V *y = Sel::y(L);
    k::bmcgs<V>(s, L.nblk, L.bptr.p, L.boff.p, L.bncol, L.ptr.p, L.col.p, Sel::ac(L),
                Sel::anb(L), Sel::b(L), x, L.parent.p, xc, y, 2, post);
    return y;
  }
  // the post-sweep's first pass recomputes colour 0 from its neighbours
  // alone, so only the other colours need the coarse correction
  k::prolong_fast<V>(s, (L.kind == 0 && post > 0) ? L.col_off[1] : 0, L.n, L.parent.p, xc, x);
  if (post > 0) mcgs_sweeps<V>(ctx, L, post, false, false);
  return x;
}

void pcg_setup(ufvm_ctx *ctx, System &S, const ufvm_pcg_cfg &cfg) {
  if (!S.hier_valid) throw Error(UFVM_ERR_STATE, "no AMG hierarchy: call ufvm_amg_agglomerate");
  if (cfg.precision != 32 && cfg.precision != 64) throw Error(UFVM_ERR_ARG, "precision must be 32 or 64");
  if (cfg.max_coarse_cells < 1 || cfg.max_coarse_cells > 4096)
    throw Error(UFVM_ERR_ARG, "max_coarse_cells must be in [1, 4096]");
  PcgState &P = S.pcg;
  P.cfg = cfg;
  int32_t nlev = S.nlevels;
  for (int32_t l = 0; l < S.nlevels; ++l)
    if (S.lev[l]->n <= cfg.max_coarse_cells) { nlev = l + 1; break; }
  if (S.lev[nlev - 1]->n > 4096)
    throw Error(UFVM_ERR_UNSUPPORTED, "coarsest level too large for the dense solve");
  P.nlev = nlev;
  // constant nullspace: all row sums vanish (pure-Neumann pressure correction)
  RefLevel &R0 = *S.lev[0];
  if (cfg.nullspace < 0) {
    std::vector<double> ac(R0.n), anb(std::max(R0.nnz, 1));
    R0.ac.download(ac.data(), R0.n, ctx->stream);
    R0.anb.download(anb.data(), R0.nnz, ctx->stream);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    bool sing = true;
    for (int32_t i = 0; i < R0.n && sing; ++i) {
      double rs = ac[i], mag = std::fabs(ac[i]);
      for (int32_t kk = R0.g.ptr[i]; kk < R0.g.ptr[i + 1]; ++kk) {
        rs += anb[kk];
        mag += std::fabs(anb[kk]);
      }
      if (std::fabs(rs) > 1e-9 * mag) sing = false;
    }
    P.singular = sing ? 1 : 0;
  } else {
    P.singular = cfg.nullspace ? 1 : 0;
  }
  P.lev.clear();
  P.lev.resize(nlev);
  if (cfg.coarse_smoother != 0 && cfg.coarse_smoother != 1)
    throw Error(UFVM_ERR_ARG, "coarse_smoother must be 0 or 1");
  if (cfg.smoother != 0 && cfg.smoother != 1) throw Error(UFVM_ERR_ARG, "smoother must be 0 or 1");
  for (int32_t l = 0; l < nlev; ++l)
    build_fast_level(ctx, S, l,
                     l == nlev - 1 ? 2
                                   : (l == 0 ? (cfg.smoother == 1 ? 4 : 0)
                                             : (cfg.coarse_smoother == 1 ? 3 : 1)));
  for (int32_t l = 0; l + 1 < nlev; ++l) link_levels(ctx, S, l);
  // dense inverse of the coarsest level (+ mean(diag)/n * 11^T when singular)
  RefLevel &RC = *S.lev[nlev - 1];
  const int32_t nc = RC.n;
  double shift = 0.0;
  if (P.singular) {
    std::vector<double> ac(nc);
    RC.ac.download(ac.data(), nc, ctx->stream);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    for (double v : ac) shift += v;
    shift /= (double)nc * nc;
  }
  P.ainv.alloc((size_t)nc * nc);
  P.colbuf.alloc(nc);
  // dense matrix in the multicolour numbering of the coarsest level (identity
  // permutation there, since it has a single colour class)
  k::csr_to_dense(ctx->stream, nc, RC.ptr.p, RC.col.p, RC.ac.p, RC.anb.p, shift, P.ainv.p);
  k::gj_inverse(ctx->stream, nc, P.ainv.p, P.colbuf.p);
  const int32_t n0 = R0.n;
  for (DBuf<double> *v : {&P.x, &P.r, &P.p, &P.w}) {
    v->alloc(n0);
    UFVM_HIP(hipMemsetAsync(v->p, 0, n0 * sizeof(double), ctx->stream));
  }
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
  P.ready = true;
}

template <typename V>
void pcg_solve_t(ufvm_ctx *ctx, System &S, int32_t *iters, double *relres) {
  using Sel = VecSel<V>;
  PcgState &P = S.pcg;
  RefLevel &R0 = *S.lev[0];
  FastLevel &L0 = *P.lev[0];
  hipStream_t s = ctx->stream;
  const int32_t n = R0.n;
  if (!L0.soff.p || !L0.scol.p) throw Error(UFVM_ERR_STATE, "level 0 has no sliced-ELL copy");
  ensure_partial(ctx, n, 5);
  double *scal = ctx->scal.p;
  // r = P b (multicolour numbering), projected when singular
  k::permute_gather(s, n, L0.perm_d.p, R0.bc.p, P.r.p);
  if (P.singular) {
    k::sum_partial(s, n, P.r.p, ctx->partial.p);
    k::reduce_final(s, ctx->partial.p, k::nblocks(n), scal + 12);
    k::axpy_const(s, n, scal + 12, -1.0 / (double)n, P.r.p);
  }
  k::fill<double>(s, n, P.x.p, 0.0);
  k::fill<double>(s, n, P.w.p, 0.0);
  k::dot_partial(s, n, P.r.p, P.r.p, ctx->partial.p);
  k::reduce_final(s, ctx->partial.p, k::nblocks(n), scal + 10);
  if (sizeof(V) == 4) k::convert<double, float>(s, n, P.r.p, (float *)Sel::b(L0));
  else k::convert<double, double>(s, n, P.r.p, (double *)Sel::b(L0));
  read_scal(ctx, 10, 1);
  const double nb = std::sqrt(ctx->hscal[10]);
  double rr = nb;
  int it = 0;
  if (nb > 0.0) {
    // The host keeps one iteration queued beyond the one whose residual it
    // waits for, so the GPU never idles on the convergence test.  The queued
    // iteration after the converged one finds the device flag set and leaves
    // x and r unchanged.
    const double thr = (P.cfg.tol * nb) * (P.cfg.tol * nb);
    k::fill<double>(s, 1, scal + 11, 0.0);
    // iterations after the first find their first colour pass made by the
    // previous iteration's x/r update (pre-smoothing by the colour-blocked
    // Gauss-Seidel only)
    const bool fuse0 = L0.kind == 0 && P.cfg.pre_sweep > 0;
    auto issue = [&](int k) {
      const V *z = vcycle_fast<V>(ctx, S, 0, fuse0 && k > 0);
      k::fcg_dots<V>(s, n, z, P.w.p, P.r.p, ctx->partial.p);
      k::fcg_update_p<V>(s, n, z, ctx->partial.p, P.singular, k == 0, scal, P.p.p);
      {
        ProfScope ps(ctx, "spmv0",
                     ctx->prof_on ? 4.0 * (n / 64 + 1) + 12.0 * L0.nnz + 8.0 * 3 * n : 0.0);
        k::spmv_dot(s, n, L0.soff.p, L0.scol.p, L0.ac64.p, L0.sval64.p, P.p.p, P.w.p,
                    ctx->partial.p);
      }
      k::fcg_update_xr<V>(s, n, scal, P.p.p, P.w.p, P.x.p, P.r.p, Sel::b(L0), ctx->partial.p,
                          thr, ctx->dring + 2 * (k & 1), k, fuse0 ? L0.col_off[1] : 0,
                          Sel::ac(L0), Sel::x(L0));
      UFVM_HIP(hipEventRecord(ctx->iter_ev[k & 1], s));
    };
    // Run one iteration ahead unless the residual history predicts that the
    // next one converges (then the one after it would be issued for nothing).
    int issued = 0;
    bool ahead = true;
    double prev2 = -1.0;
    while (it < P.cfg.max_iter) {
      while (issued < P.cfg.max_iter && issued < it + (ahead ? 2 : 1)) issue(issued++);
      UFVM_HIP(hipEventSynchronize(ctx->iter_ev[it & 1]));
      const volatile double *slot = ctx->hring + 2 * (it & 1);
      if (slot[1] != (double)it) throw Error(UFVM_ERR_STATE, "FPCG residual ring out of step");
      const double rr2 = slot[0];   // the value the device compared with thr
      rr = std::sqrt(rr2);
      ++it;
      if (std::isnan(rr)) throw Error(UFVM_ERR_NUMERIC, "FPCG breakdown (NaN residual)");
      if (rr2 <= thr) break;
      // at most iteration `it` is in flight now: predict its residual from the
      // last reduction factor
      if (prev2 > 0.0) ahead = rr2 * (rr2 / prev2) > 4.0 * thr;
      prev2 = rr2;
    }
  }
  k::permute_scatter(s, n, L0.perm_d.p, P.x.p, R0.dphi.p);
  UFVM_HIP(hipStreamSynchronize(s));
  if (iters) *iters = it;
  if (relres) *relres = nb > 0 ? rr / nb : 0.0;
}

void pcg_solve(ufvm_ctx *ctx, System &S, int32_t *iters, double *relres) {
  if (!S.pcg.ready) throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first");
  if (S.pcg.cfg.precision == 32) pcg_solve_t<float>(ctx, S, iters, relres);
  else pcg_solve_t<double>(ctx, S, iters, relres);
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
  const double initial = cyc_final(c);
  double fin = initial;
  int32_t nCycle = 1;
  while (nCycle <= cfg.max_cycles && fin > cfg.termination * initial) {
    if (cfg.cycle == UFVM_CYCLE_V) fin = vcycle(c, G);
    else if (cfg.cycle == UFVM_CYCLE_W) fin = wcycle(c, G);
    else fin = fcycle(c, G);
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
    k::coarse_lhs(ctx->stream, C.n, F.child_ptr.p, F.child.p, F.ptr.p, F.ac.p, F.anb.p, F.slot.p,
                  C.ptr.p, C.ac.p, C.anb.p);
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
    UFVM_HIP(hipHostMalloc((void **)&ctx->hscal, kScal * sizeof(double), hipHostMallocDefault));
    UFVM_HIP(hipHostMalloc((void **)&ctx->hring, 8 * sizeof(double),
                           hipHostMallocMapped | hipHostMallocCoherent));
    UFVM_HIP(hipHostGetDevicePointer((void **)&ctx->dring, ctx->hring, 0));
    for (hipEvent_t &e : ctx->iter_ev) UFVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ctx->partial.alloc(1024);
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
  if (ctx->hscal) (void)hipHostFree(ctx->hscal);
  if (ctx->hring) (void)hipHostFree(ctx->hring);
  for (hipEvent_t e : ctx->iter_ev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return UFVM_OK;
}

const char *ufvm_last_error(const ufvm_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void *ufvm_ctx_stream(ufvm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int ufvm_ctx_synchronize(ufvm_ctx *ctx) {
  return guard(ctx, [&] {
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
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

int ufvm_pcg_apply_vcycle(ufvm_ctx *ctx, int32_t sys, const double *r, double *z) {
  return guard(ctx, [&] {
    System &S = sys_of(ctx, sys);
    if (!S.pcg.ready) throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first");
    if (!r || !z) throw Error(UFVM_ERR_ARG, "null vector");
    PcgState &P = S.pcg;
    FastLevel &L0 = *P.lev[0];
    const int32_t n = L0.n;
    hipStream_t s = ctx->stream;
    UFVM_HIP(hipMemcpyAsync(P.w.p, r, n * sizeof(double), hipMemcpyHostToDevice, s));
    k::permute_gather(s, n, L0.perm_d.p, P.w.p, P.p.p);
    if (P.cfg.precision == 32) {
      k::convert<double, float>(s, n, P.p.p, L0.b32.p);
      const float *z32 = vcycle_fast<float>(ctx, S, 0);
      k::convert<float, double>(s, n, z32, P.p.p);
    } else {
      k::convert<double, double>(s, n, P.p.p, L0.b64.p);
      const double *z64 = vcycle_fast<double>(ctx, S, 0);
      k::convert<double, double>(s, n, z64, P.p.p);
    }
    k::permute_scatter(s, n, L0.perm_d.p, P.p.p, P.w.p);
    UFVM_HIP(hipMemcpyAsync(z, P.w.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
    UFVM_HIP(hipStreamSynchronize(s));
  });
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
      ++cnt;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = cnt;
    if (bytes) *bytes = by;
  });
}

}  // extern "C"
