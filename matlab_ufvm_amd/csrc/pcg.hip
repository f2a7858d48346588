// pcg.hip -- precision-form pressure solve: FP32 (or FP64) AMG V-cycle over
// the reference's agglomeration hierarchy inside FP64 flexible CG (Notay
// FCG(1)).  No reference counterpart; the hierarchy (parents, coarse
// patterns and Galerkin operators) is the reference's, bit-exact.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <map>

#include "kernels.h"
#include "pcg.h"

namespace ufvm {
namespace detail {

// the column indices of a level's sliced-ELL copy, in whichever form it has
static k::SellCols sell_cols(const FastLevel &L) {
  k::SellCols c;
  if (L.sd16.p) {
    c.d16 = L.sd16.p;
    c.base = L.scbase.p;
  } else {
    c.col = L.scol.p;
  }
  return c;
}
// bytes per sliced-ELL column index (2: the 16-bit form)
static double index_bytes(const FastLevel &L) { return L.sd16.p ? 2.0 : 4.0; }

// Precision-form copy of hierarchy level l.  kind 0 (level 1): unknowns
// renumbered so every colour class of a global greedy colouring is one
// contiguous row range.  kind 1 (coarse levels): blocks of kBRows consecutive
// reference rows, each sorted by a colouring of the block's own subgraph.
// kind 2: the coarsest level, kept in reference order for the dense solve.
// kind 3 (coarse levels smoothed by damped Jacobi): rows sorted by length
// inside windows.  kind 5 (finest level, DILU in colour order): kind 0's
// numbering; kind 6 (ILU(0) in colour order) the same.  Rows keep their cconn
// order in every kind.
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
  FL.wide = kind == 3 && (int64_t)R.nnz > (int64_t)k::kSellRun * n;
  FL.perm.resize(n);
  if (kind == 0 || kind == 5 || kind == 6) {
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
  // pattern on the host: permuted CSR and, per entry, its reference position
  FL.h_ptr.assign(n + 1, 0);
  std::vector<int32_t> col(std::max(R.nnz, 1)), csrc(std::max(R.nnz, 1), -1);
  for (int32_t t = 0; t < n; ++t) {
    const int32_t i = FL.perm[t];
    FL.h_ptr[t + 1] = FL.h_ptr[t] + (R.g.ptr[i + 1] - R.g.ptr[i]);
    int32_t o = FL.h_ptr[t];
    for (int32_t kk = R.g.ptr[i]; kk < R.g.ptr[i + 1]; ++kk, ++o) {
      col[o] = iperm[R.g.col[kk]];
      csrc[o] = kk;
    }
  }
  FL.ptr.upload(FL.h_ptr, s);
  FL.col.upload(col, s);
  FL.csr_src.upload(csrc, s);
  FL.perm_d.upload(FL.perm, s);
  FL.ac64.alloc(n);
  FL.ac32.alloc(n);
  FL.anb64.alloc(std::max(R.nnz, 1));
  FL.anb32.alloc(std::max(R.nnz, 1));
  if (kind == 4) {
    ensure_schedule(ctx, R);
    if (!R.low_ok) throw Error(UFVM_ERR_NUMERIC, "ILU: the index for i in j is not found");
    FL.dinv64.alloc(n);
    FL.dinv32.alloc(n);
  }
  if (kind == 5 || kind == 6) {
    // lower = earlier colour = smaller permuted index (a colour class has no
    // internal edges)
    HostCSR pg;
    pg.n = n;
    pg.ptr = FL.h_ptr;
    pg.col.assign(col.begin(), col.begin() + R.nnz);
    LowerLists low;
    if (!lower_lists(pg, low)) throw Error(UFVM_ERR_NUMERIC, "ILU: the index for i in j is not found");
    FL.lptr.upload(low.ptr, s);
    FL.lnb.upload(low.nb.empty() ? std::vector<int32_t>{0} : low.nb, s);
    FL.lpji.upload(low.pos_ji.empty() ? std::vector<int32_t>{0} : low.pos_ji, s);
    FL.lpij.upload(low.pos_ij.empty() ? std::vector<int32_t>{0} : low.pos_ij, s);
    std::vector<int32_t> iota(n);
    for (int32_t t = 0; t < n; ++t) iota[t] = t;
    FL.iota.upload(iota, s);
    FL.dinv64.alloc(n);
    FL.dinv32.alloc(n);
    if (kind == 6) {
      Ilu0Lists U;
      ilu0_lists(pg, low, U);
      FL.hptr.upload(U.ptr, s);
      FL.hsrc.upload(U.src.empty() ? std::vector<int32_t>{0} : U.src, s);
      FL.hdst.upload(U.dst.empty() ? std::vector<int32_t>{0} : U.dst, s);
      FL.f64.alloc(std::max(R.nnz, 1));
    }
  }
  // the FPCG's operator product runs on level 0's sliced-ELL copy whatever its
  // kind (a single-level hierarchy makes level 0 the dense coarsest level)
  if (kind == 0 || kind == 3 || kind == 5 || kind == 6 || l == 0) {
    const int32_t nq = (n + 63) / 64;
    FL.h_soff.assign(nq + 1, 0);
    for (int32_t q = 0; q < nq; ++q) {
      int32_t w = 0;
      for (int32_t t = q * 64; t < std::min(n, q * 64 + 64); ++t)
        w = std::max(w, FL.h_ptr[t + 1] - FL.h_ptr[t]);
      FL.h_soff[q + 1] = FL.h_soff[q] + 64 * w;
    }
    const size_t tot = std::max(FL.h_soff[nq], 1);
    // padding: col = the row itself, value 0 (source -1; rows past n in the
    // last chunk are never processed)
    std::vector<int32_t> scol(tot, 0), ssrc(tot, -1), fsrc(tot, -1);
    for (int32_t t = 0; t < n; ++t) {
      const int32_t q = t / 64, base = FL.h_soff[q] + (t & 63);
      const int32_t w = (FL.h_soff[q + 1] - FL.h_soff[q]) / 64;
      const int32_t len = FL.h_ptr[t + 1] - FL.h_ptr[t];
      for (int32_t e = 0; e < w; ++e) {
        scol[base + e * 64] = e < len ? col[FL.h_ptr[t] + e] : t;
        ssrc[base + e * 64] = e < len ? csrc[FL.h_ptr[t] + e] : -1;
        fsrc[base + e * 64] = e < len ? FL.h_ptr[t] + e : -1;
      }
    }
    FL.soff.upload(FL.h_soff, s);
    std::vector<uint16_t> d16;
    std::vector<int32_t> cbase;
    if (S.pcg.cfg.index_width == 0 && pack_sell_columns(FL.h_ptr, FL.h_soff, scol, n, d16, cbase)) {
      FL.sd16.upload(d16, s);
      FL.scbase.upload(cbase, s);
    } else {
      FL.scol.upload(scol, s);
    }
    FL.sell_src.upload(ssrc, s);
    FL.sval64.alloc(tot);
    FL.sval32.alloc(tot);
    if (kind == 0 && l == 0 && S.pcg.cfg.precision == 16) FL.sval16.alloc(tot);
    if (kind == 6) {
      FL.fsrc.upload(fsrc, s);
      FL.sf64.alloc(tot);
      FL.sf32.alloc(tot);
    }
  }
  UFVM_HIP(hipStreamSynchronize(s));
}

// Values of level copy l from the reference level's current coefficients
// (device gathers through the maps build_fast_level recorded), both
// precisions, and the DILU factor of a kind-4 level.
void refresh_level(ufvm_ctx *ctx, System &S, int l) {
  hipStream_t s = ctx->stream;
  RefLevel &R = *S.lev[l];
  FastLevel &FL = *S.pcg.lev[l];
  const int32_t n = FL.n;
  k::gather_vals<double>(s, n, FL.perm_d.p, R.ac.p, FL.ac64.p);
  k::gather_vals<float>(s, n, FL.perm_d.p, R.ac.p, FL.ac32.p);
  if (FL.nnz > 0) {
    k::gather_vals<double>(s, FL.nnz, FL.csr_src.p, R.anb.p, FL.anb64.p);
    k::gather_vals<float>(s, FL.nnz, FL.csr_src.p, R.anb.p, FL.anb32.p);
  }
  if (FL.sell_src.p) {
    const int64_t slots = FL.h_soff.back();
    k::gather_vals<double>(s, slots, FL.sell_src.p, R.anb.p, FL.sval64.p);
    k::gather_vals<float>(s, slots, FL.sell_src.p, R.anb.p, FL.sval32.p);
    if (FL.sval16.p) k::scale_half(s, n, FL.soff.p, FL.sval64.p, FL.ac64.p, FL.sval16.p);
  }
  if (FL.kind == 4) {
    // the reference's DILU factor (cfdFactorizeILU.m:10-47, row N-1 quirk
    // kept) in FP64 over the lexicographic schedule, stored in both precisions
    const auto &F = R.fwd.level_ptr;
    if (R.one_wg)
      k::dilu_factor_block(s, R.fwd_rows.p, R.fwd_lp.p, (int32_t)F.size() - 1, R.ac.p, R.anb.p,
                           R.lptr.p, R.lnb.p, R.lpji.p, R.lpij.p, R.n - 2, R.dinv.p);
    else
      for (size_t q = 0; q + 1 < F.size(); ++q)
        k::dilu_factor_rows(s, R.fwd_rows.p + F[q], F[q + 1] - F[q], R.ac.p, R.anb.p, R.lptr.p,
                            R.lnb.p, R.lpji.p, R.lpij.p, R.n - 2, R.dinv.p);
    k::convert<double, double>(s, n, R.dinv.p, FL.dinv64.p);
    k::convert<double, float>(s, n, R.dinv.p, FL.dinv32.p);
  }
  if (FL.kind == 5) {
    // DILU in colour order, FP64, one launch per colour:
    // d_j = a_jj - sum over earlier-colour neighbours i (increasing i) of
    // (a_ji / d_i) a_ij
    for (int c = 0; c < FL.ncol; ++c)
      k::dilu_factor_rows(s, FL.iota.p + FL.col_off[c], FL.col_off[c + 1] - FL.col_off[c],
                          FL.ac64.p, FL.anb64.p, FL.lptr.p, FL.lnb.p, FL.lpji.p, FL.lpij.p, -1,
                          FL.dinv64.p);
    k::convert<double, float>(s, n, FL.dinv64.p, FL.dinv32.p);
  }
  if (FL.kind == 6) {
    // ILU(0) in colour order, FP64, one launch per colour; then the factor's
    // off-diagonals into the sliced-ELL layout in both precisions
    if (FL.nnz > 0) k::convert<double, double>(s, FL.nnz, FL.anb64.p, FL.f64.p);
    for (int c = 0; c < FL.ncol; ++c)
      k::ilu0_factor_rows(s, FL.col_off[c], FL.col_off[c + 1], FL.ac64.p, FL.lptr.p, FL.lnb.p,
                          FL.lpji.p, FL.hptr.p, FL.hsrc.p, FL.hdst.p, FL.f64.p, FL.dinv64.p);
    k::convert<double, float>(s, n, FL.dinv64.p, FL.dinv32.p);
    const int64_t slots = FL.h_soff.back();
    k::gather_vals<double>(s, slots, FL.fsrc.p, FL.f64.p, FL.sf64.p);
    k::gather_vals<float>(s, slots, FL.fsrc.p, FL.f64.p, FL.sf32.p);
  }
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


// matrix arrays of a level copy and vectors of a level workspace, by precision
template <typename V>
struct Prec;
template <>
struct Prec<float> {
  static const float *ac(const FastLevel &L) { return L.ac32.p; }
  static const float *anb(const FastLevel &L) { return L.anb32.p; }
  static const float *sval(const FastLevel &L) { return L.sval32.p; }
  static const float *dinv(const FastLevel &L) { return L.dinv32.p; }
  static const float *sfac(const FastLevel &L) { return L.sf32.p; }
  static float *b(LevelWork &W) { return W.b32.p; }
  static float *x(LevelWork &W) { return W.x32.p; }
  static float *y(LevelWork &W) { return W.y32.p; }
  static float *r(LevelWork &W) { return W.r32.p; }
  static DBuf<float> &buf(LevelWork &W, int i) {
    return i == 0 ? W.b32 : i == 1 ? W.x32 : i == 2 ? W.r32 : W.y32;
  }
};
template <>
struct Prec<double> {
  static const double *ac(const FastLevel &L) { return L.ac64.p; }
  static const double *anb(const FastLevel &L) { return L.anb64.p; }
  static const double *sval(const FastLevel &L) { return L.sval64.p; }
  static const double *dinv(const FastLevel &L) { return L.dinv64.p; }
  static const double *sfac(const FastLevel &L) { return L.sf64.p; }
  static double *b(LevelWork &W) { return W.b64.p; }
  static double *x(LevelWork &W) { return W.x64.p; }
  static double *y(LevelWork &W) { return W.y64.p; }
  static double *r(LevelWork &W) { return W.r64.p; }
  static DBuf<double> &buf(LevelWork &W, int i) {
    return i == 0 ? W.b64 : i == 1 ? W.x64 : i == 2 ? W.r64 : W.y64;
  }
};

// Jacobi levels with at most this many rows fuse the prolongation into the
// first post-smoothing sweep.
constexpr int32_t kFusedProlongRows = 100000;
constexpr int32_t kTailRows = 40000;   // profiling: levels from here down are "tail"

// the FP16 row-scaled finest-level values of an FP32 V-cycle (precision 16),
// or null
template <typename V>
const _Float16 *half_of(const FastLevel &L) {
  return std::is_same<V, float>::value ? L.sval16.p : nullptr;
}

// One preconditioner V-cycle of system S on workspace W and stream s.
template <typename V>
struct Cycle {
  using P = Prec<V>;
  ufvm_ctx *ctx;
  System &S;
  SolveWork &W;
  hipStream_t s;

  // Symmetric multicolour Gauss-Seidel sweeps on the finest level (kind 0):
  // colours upward then downward; the downward pass skips the last colour
  // and every sweep after the first skips colour 0, because re-solving the
  // colour just solved repeats identical arithmetic.  from_zero: x holds no
  // data yet; first_done: the first pass (colour 0 from zero) was already
  // made by the kernel that produced b.
  void mcgs_sweeps(const FastLevel &L, LevelWork &LW, int nsweep, bool from_zero,
                   bool first_done) {
    const double w = sizeof(V);
    // one event pair around the back-to-back colour passes of this call (the
    // event overhead is spread over several launches)
    ProfScope ps(ctx, s, "mcgs0", 0.0);
    auto pass = [&](int c, int32_t zero_above) {
      const int32_t r0 = L.col_off[c], r1 = L.col_off[c + 1];
      const bool h = half_of<V>(L) != nullptr;
      if (ctx->prof_on) {
        // algorithmic bytes: the colour's entries (col + value), its ac, b,
        // x write, and one read of every other unknown
        const double R = r1 - r0, Z = L.h_ptr[r1] - L.h_ptr[r0];
        ps.add((index_bytes(L) + (h ? 2.0 : w)) * Z + 3.0 * w * R + w * (L.n - R));
      }
      if constexpr (std::is_same<V, float>::value) {
        if (h) {
          k::mcgs_sell_h(s, r0, r1, L.soff.p, sell_cols(L), L.sval16.p, P::ac(L), P::b(LW),
                         P::x(LW), zero_above);
          return;
        }
      }
      k::mcgs_sell<V>(s, r0, r1, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW), P::x(LW),
                      zero_above);
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
  void dilu_sweep(const RefLevel &R, const FastLevel &L, LevelWork &LW, V *x, bool from_zero) {
    V *rc = P::r(LW);
    if (from_zero) {
      k::fill<V>(s, L.n, x, V(0));
      k::convert<V, V>(s, L.n, P::b(LW), rc);
    } else {
      k::residual_sell<V>(s, L.n, 0, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW), x, rc);
    }
    const auto &F = R.fwd.level_ptr;
    const auto &B = R.bwd.level_ptr;
    if (R.one_wg) {
      k::dilu_fwd_block<V>(s, R.fwd_rows.p, R.fwd_lp.p, (int32_t)F.size() - 1, P::anb(L),
                           R.lptr.p, R.lnb.p, R.lpji.p, P::dinv(L), rc);
      k::dilu_bwd_block<V>(s, R.bwd_rows.p, R.bwd_lp.p, (int32_t)B.size() - 1, L.n, L.ptr.p,
                           L.col.p, P::anb(L), P::dinv(L), rc, x);
      return;
    }
    for (size_t q = 0; q + 1 < F.size(); ++q)
      k::dilu_fwd_rows<V>(s, R.fwd_rows.p + F[q], F[q + 1] - F[q], P::anb(L), R.lptr.p, R.lnb.p,
                          R.lpji.p, P::dinv(L), rc);
    for (size_t q = 0; q + 1 < B.size(); ++q)
      k::dilu_bwd_rows<V>(s, R.bwd_rows.p + B[q], B[q + 1] - B[q], L.n, L.ptr.p, L.col.p,
                          P::anb(L), P::dinv(L), rc, x);
  }

  // One sweep of DILU (kind 5) or ILU(0) (kind 6) in colour order: a forward
  // pass over the colours upward (the residual b - A x fused in; from zero:
  // x = 0) and a backward pass over the colours downward (x += z; from zero:
  // x = z), one launch per colour each way on the sliced-ELL copy.  ILU(0)'s
  // substitutions read the factor's off-diagonals, the residual A's.
  void mc_dilu_sweep(const FastLevel &L, LevelWork &LW, V *x, bool from_zero) {
    V *rc = P::r(LW);
    const V *fac = L.kind == 6 ? P::sfac(L) : P::sval(L);
    for (int c = 0; c < L.ncol; ++c)
      k::mcilu_fwd_sell<V>(s, L.col_off[c], L.col_off[c + 1], L.soff.p, sell_cols(L), P::sval(L),
                           fac, P::ac(L), P::b(LW), x, P::dinv(L), rc, from_zero);
    for (int c = L.ncol - 1; c >= 0; --c)
      k::mcilu_bwd_sell<V>(s, L.col_off[c], L.col_off[c + 1], L.soff.p, sell_cols(L), fac,
                           P::dinv(L), rc, x, from_zero);
  }

  // smoothing sweeps of level l (the coarse levels may have their own counts)
  int pre_of(int l) const {
    const ufvm_pcg_cfg &c = S.pcg.cfg;
    return (l == 0 || c.coarse_pre_sweep < 0) ? c.pre_sweep : c.coarse_pre_sweep;
  }
  int post_of(int l) const {
    const ufvm_pcg_cfg &c = S.pcg.cfg;
    return (l == 0 || c.coarse_post_sweep < 0) ? c.post_sweep : c.coarse_post_sweep;
  }

  // precision 16's finest level: its residual from the FP16 values (false on
  // every other level)
  bool residual_half(const FastLevel &L, LevelWork &LW, int32_t zero_rows, const V *x) {
    if constexpr (std::is_same<V, float>::value) {
      if (L.sval16.p) {
        k::residual_sell_h(s, L.n, zero_rows, L.soff.p, sell_cols(L), L.sval16.p, P::ac(L),
                           P::b(LW), x, P::r(LW));
        return true;
      }
    }
    return false;
  }

  // residual of level l into r, then restriction to level l+1 (fused with the
  // first Jacobi sweep there when that level is Jacobi-smoothed)
  void down(const FastLevel &L, LevelWork &LW, int l, const V *x, int32_t zero_rows) {
    const PcgState &Pc = S.pcg;
    const FastLevel &C = *Pc.lev[l + 1];
    LevelWork &CW = W.lev[l + 1];
    // the coarse correction is scaled by coarse_scale through the restricted
    // right-hand side (the coarse cycle is linear in it)
    const V scale = V(Pc.cfg.coarse_scale);
    const bool jac = C.kind == 3 && pre_of(l + 1) > 0;
    if (L.kind == 1)
      k::residual_fast<V>(s, L.n, L.ptr.p, L.col.p, P::ac(L), P::anb(L), P::b(LW), x, P::r(LW));
    else if (!residual_half(L, LW, zero_rows, x))
      k::residual_sell<V>(s, L.n, zero_rows, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW),
                          x, P::r(LW), L.wide);
    if (jac)
      k::restrict_jacobi<V>(s, C.n, L.csoff.p, L.cschild.p, P::r(LW), P::ac(C), P::b(CW),
                            P::x(CW), V(2.0 / 3.0), scale);
    else
      k::restrict_fast<V>(s, C.n, L.csoff.p, L.cschild.p, P::r(LW), P::b(CW), scale);
  }

  // One V-cycle from level l down; returns the level's solution vector.
  // first_done (finest level only): x already holds the first colour pass.
  V *vcycle(int l, bool first_done = false) {
    const PcgState &Pc = S.pcg;
    const FastLevel &L = *Pc.lev[l];
    LevelWork &LW = W.lev[l];
    const int pre = pre_of(l), post = post_of(l);
    if (l == Pc.nlev - 1) {
      k::dense_solve<V>(s, L.n, Pc.ainv.p, P::b(LW), P::x(LW));
      return P::x(LW);
    }
    V *x = P::x(LW);
    const V omega = V(2.0 / 3.0);
    if (L.kind == 3) {
      // x already holds one Jacobi sweep from zero (fused into the
      // restriction that produced b); further sweeps ping-pong x and y
      V *y = P::y(LW);
      for (int it = 1; it < pre; ++it) {
        k::jacobi_sell<V>(s, L.n, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW), x, nullptr,
                          nullptr, y, omega, L.wide);
        std::swap(x, y);
      }
      if (pre == 0) k::fill<V>(s, L.n, x, V(0));
      down(L, LW, l, x, 0);
      V *xc;
      {
        // the latency-bound tail: every level from the first of at most
        // kTailRows rows down (events only when profiling)
        const bool tail = Pc.lev[l + 1]->n <= kTailRows && L.n > kTailRows;
        ProfScope ps(ctx, s, tail ? "tail" : "", 0.0);
        xc = vcycle(l + 1);
      }
      // small levels are latency-bound: the first post-sweep adds the coarse
      // correction on the fly instead of a separate prolongation launch
      const bool fuse = post > 0 && L.n <= kFusedProlongRows;
      if (!fuse) k::prolong_fast<V>(s, 0, L.n, L.parent.p, xc, x);
      for (int it = 0; it < post; ++it) {
        const bool pro = fuse && it == 0;
        k::jacobi_sell<V>(s, L.n, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW), x,
                          pro ? L.parent.p : nullptr, pro ? xc : nullptr, y, omega, L.wide);
        std::swap(x, y);
      }
      return x;
    }
    if (L.kind == 4 || L.kind == 5 || L.kind == 6) {
      const RefLevel &R = *S.lev[l];
      auto sweep = [&](bool from_zero) {
        if (L.kind == 4) dilu_sweep(R, L, LW, x, from_zero);
        else mc_dilu_sweep(L, LW, x, from_zero);
      };
      if (pre == 0) k::fill<V>(s, L.n, x, V(0));
      for (int it = 0; it < pre; ++it) sweep(it == 0);
      down(L, LW, l, x, 0);
      V *xc = vcycle(l + 1);
      k::prolong_fast<V>(s, 0, L.n, L.parent.p, xc, x);
      for (int it = 0; it < post; ++it) sweep(false);
      return x;
    }
    if (pre == 0) {
      k::fill<V>(s, L.n, x, V(0));
    } else if (L.kind == 0) {
      mcgs_sweeps(L, LW, pre, true, first_done);
    } else {
      k::bmcgs<V>(s, L.nblk, L.bptr.p, L.boff.p, L.bncol, L.ptr.p, L.col.p, P::ac(L), P::anb(L),
                  P::b(LW), x, nullptr, nullptr, x, 0, pre);
    }
    // after a symmetric Gauss-Seidel sweep the rows of colour 0 (solved
    // last) have a residual of rounding noise only: it is taken as 0
    down(L, LW, l, x, (L.kind == 0 && pre > 0) ? L.col_off[1] : 0);
    V *xc = vcycle(l + 1);
    if (L.kind == 1 && post > 0) {
      V *y = P::y(LW);
      k::bmcgs<V>(s, L.nblk, L.bptr.p, L.boff.p, L.bncol, L.ptr.p, L.col.p, P::ac(L), P::anb(L),
                  P::b(LW), x, L.parent.p, xc, y, 2, post);
      return y;
    }
    // the post-sweep's first pass recomputes colour 0 from its neighbours
    // alone, so only the other colours need the coarse correction
    k::prolong_fast<V>(s, (L.kind == 0 && post > 0) ? L.col_off[1] : 0, L.n, L.parent.p, xc, x);
    if (post > 0) mcgs_sweeps(L, LW, post, false, false);
    return x;
  }
};

// --------------------------------------------------------------- workspaces

// stride between the right-hand sides of a k-solve batch in an n-row vector
// (one solve: n; a batch: n rounded up to 64 rows, so every vector starts on
// a 256-byte boundary)
static int64_t batch_ld(int64_t n, int k) { return k == 1 ? n : (n + 63) / 64 * 64; }
// partial sums per right-hand side: the start-up reductions use one partial
// per 256-row block, the FCG kernels at most 6 per 1024-row block
static int64_t partial_stride(int32_t n0) { return 5 * (int64_t)k::nblocks((int64_t)n0) + 16; }
// ring doubles per right-hand side: {r.r, iteration} of even and odd
// iterations, then the start-up norm
constexpr int kRingPer = 6;

size_t work_bytes(const System &S, int k) {
  const PcgState &P = S.pcg;
  const size_t w = P.cfg.precision != 64 ? 4 : 8;
  size_t bytes = 0;
  for (int32_t l = 0; l < P.nlev; ++l) bytes += 4 * w * k * (size_t)batch_ld(P.lev[l]->n, k);
  const size_t n0 = S.lev[0]->n;
  bytes += 4 * 8 * k * (size_t)batch_ld(n0, k);                   // x r p w
  bytes += 2 * 8 * k * n0;                                        // bref xref
  bytes += 8 * k * (size_t)partial_stride((int32_t)n0);           // partials
  return bytes + 8 * k * kScal;
}

template <typename V>
static void alloc_levels(const System &S, SolveWork &W, hipStream_t s) {
  const PcgState &P = S.pcg;
  const int k = W.nrhs;
  W.lev.clear();
  W.lev.resize(P.nlev);
  for (int32_t l = 0; l < P.nlev; ++l) {
    W.lev[l].ld = batch_ld(P.lev[l]->n, k);
    for (int i = 0; i < 4; ++i) {
      // iterates start finite: the sliced-ELL padding multiplies x[row] by 0
      DBuf<V> &d = Prec<V>::buf(W.lev[l], i);
      d.alloc((size_t)W.lev[l].ld * k);
      UFVM_HIP(hipMemsetAsync(d.p, 0, d.n * sizeof(V), s));
    }
  }
}

std::unique_ptr<SolveWork> make_work(const System &S, hipStream_t s, int k) {
  if (!S.pcg.ready) throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first");
  if (k != 1 && k != 2 && k != 4 && k != 8)
    throw Error(UFVM_ERR_ARG, "a batch holds 1, 2, 4 or 8 right-hand sides");
  std::unique_ptr<SolveWork> W(new SolveWork());
  const int32_t n0 = S.lev[0]->n;
  W->n0 = n0;
  W->nrhs = k;
  W->ld0 = batch_ld(n0, k);
  if (S.pcg.cfg.precision != 64) alloc_levels<float>(S, *W, s);
  else alloc_levels<double>(S, *W, s);
  for (DBuf<double> *v : {&W->x, &W->r, &W->p, &W->w}) {
    v->alloc((size_t)W->ld0 * k);
    UFVM_HIP(hipMemsetAsync(v->p, 0, v->n * sizeof(double), s));
  }
  for (DBuf<double> *v : {&W->bref, &W->xref}) {
    v->alloc((size_t)n0 * k);
    UFVM_HIP(hipMemsetAsync(v->p, 0, v->n * sizeof(double), s));
  }
  W->partial.alloc((size_t)partial_stride(n0) * k);
  W->scal.alloc((size_t)kScal * k);
  UFVM_HIP(hipMemsetAsync(W->scal.p, 0, W->scal.n * sizeof(double), s));
  UFVM_HIP(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lk(device_mgmt_mutex());
  const int nring = kRingPer * k + 2;
  UFVM_HIP(hipHostMalloc((void **)&W->hring, nring * sizeof(double),
                         hipHostMallocMapped | hipHostMallocCoherent));
  UFVM_HIP(hipHostGetDevicePointer((void **)&W->dring, W->hring, 0));
  for (int i = 0; i < nring; ++i) W->hring[i] = 0.0;
  for (hipEvent_t &e : W->ev) UFVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  W->bytes = work_bytes(S, k);
  return W;
}

// -------------------------------------------------------------------- set-up

// dense inverse of the coarsest level (+ mean(diag)/n * 11^T when singular),
// in its reference numbering (the identity permutation of a kind-2 level)
static void build_coarsest(ufvm_ctx *ctx, System &S) {
  PcgState &P = S.pcg;
  RefLevel &RC = *S.lev[P.nlev - 1];
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
  P.colbuf.alloc(k::gj_scratch(nc));
  k::csr_to_dense(ctx->stream, nc, RC.ptr.p, RC.col.p, RC.ac.p, RC.anb.p, shift, P.ainv.p);
  k::gj_inverse(ctx->stream, nc, P.ainv.p, P.colbuf.p);
}

void pcg_setup(ufvm_ctx *ctx, System &S, const ufvm_pcg_cfg &cfg) {
  if (!S.hier_valid) throw Error(UFVM_ERR_STATE, "no AMG hierarchy: call ufvm_amg_agglomerate");
  if (cfg.precision != 16 && cfg.precision != 32 && cfg.precision != 64)
    throw Error(UFVM_ERR_ARG, "precision must be 16, 32 or 64");
  if (cfg.precision == 16 && cfg.smoother != 0)
    throw Error(UFVM_ERR_ARG, "precision 16 needs the multicolour Gauss-Seidel finest smoother");
  if (cfg.max_coarse_cells < 1 || cfg.max_coarse_cells > 4096)
    throw Error(UFVM_ERR_ARG, "max_coarse_cells must be in [1, 4096]");
  if (cfg.coarse_smoother != 0 && cfg.coarse_smoother != 1)
    throw Error(UFVM_ERR_ARG, "coarse_smoother must be 0 or 1");
  if (cfg.smoother < 0 || cfg.smoother > 3)
    throw Error(UFVM_ERR_ARG, "smoother must be 0, 1, 2 or 3");
  const double cscale = cfg.coarse_scale == 0.0 ? 2.5 : cfg.coarse_scale;   // 0: default
  if (!(cscale > 0.0 && cscale <= 4.0))
    throw Error(UFVM_ERR_ARG, "coarse_scale must be in (0, 4] (0: the default 2.5)");
  if (cfg.launch_mode < 0 || cfg.launch_mode > 2)
    throw Error(UFVM_ERR_ARG, "launch_mode must be 0, 1 or 2");
  if (cfg.nullspace < -1 || cfg.nullspace > 1)
    throw Error(UFVM_ERR_ARG, "nullspace must be -1, 0 or 1");
  if (cfg.index_width != 0 && cfg.index_width != 4)
    throw Error(UFVM_ERR_ARG, "index_width must be 0 (auto) or 4");
  int32_t nlev = S.nlevels;
  for (int32_t l = 0; l < S.nlevels; ++l)
    if (S.lev[l]->n <= cfg.max_coarse_cells) { nlev = l + 1; break; }
  if (S.lev[nlev - 1]->n > 4096)
    throw Error(UFVM_ERR_UNSUPPORTED, "coarsest level too large for the dense solve");
  // the configuration is valid: from here the previous set-up is replaced
  PcgState &P = S.pcg;
  P.ready = false;
  P.work_batch.clear();
  P.cfg = cfg;
  P.cfg.coarse_scale = cscale;
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
  for (int32_t l = 0; l < nlev; ++l)
    build_fast_level(ctx, S, l,
                     l == nlev - 1 ? 2
                                   : (l == 0 ? (cfg.smoother == 1   ? 4
                                                : cfg.smoother == 2 ? 5
                                                : cfg.smoother == 3 ? 6
                                                                    : 0)
                                             : (cfg.coarse_smoother == 1 ? 3 : 1)));
  for (int32_t l = 0; l + 1 < nlev; ++l) link_levels(ctx, S, l);
  for (int32_t l = 0; l < nlev; ++l) refresh_level(ctx, S, l);
  build_coarsest(ctx, S);
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
  P.ready = true;
  P.stale = false;
  P.work0 = make_work(S, ctx->stream);
}

// New coefficients on an unchanged hierarchy (same parents and patterns, e.g.
// after galerkin_update): refresh every level copy's values and the coarsest
// inverse on the GPU, keeping colourings, layouts and workspaces.
void pcg_refresh(ufvm_ctx *ctx, System &S) {
  PcgState &P = S.pcg;
  if (!P.ready) throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first");
  for (int32_t l = 0; l < P.nlev; ++l) refresh_level(ctx, S, l);
  build_coarsest(ctx, S);
  UFVM_HIP(hipStreamSynchronize(ctx->stream));
  P.stale = false;
}

// ---------------------------------------------------------------- FCG solve

bool FcgRun::ready(hipEvent_t e, bool block) {
  if (block) {
    UFVM_HIP(hipEventSynchronize(e));
    return true;
  }
  const hipError_t q = hipEventQuery(e);
  if (q == hipSuccess) return true;
  if (q == hipErrorNotReady) return false;
  UFVM_HIP(q);
  return false;
}

// The start of one solve on a single-right-hand-side workspace: r = P b
// (V-cycle numbering), projected when singular, x = w = 0, |r|^2 to the ring
// slot 4 and r as the V-cycle's right-hand side (FcgRun and FcgGroup).
template <typename V>
static void fcg_setup(hipStream_t s, const System &S, SolveWork &W, const double *b) {
  const PcgState &P = S.pcg;
  const FastLevel &L0 = *P.lev[0];
  const int32_t n = W.n0;
  double *scal = W.scal.p;
  k::permute_gather(s, n, L0.perm_d.p, b, W.r.p);
  if (P.singular) {
    k::sum_partial(s, n, W.r.p, W.partial.p);
    k::reduce_final(s, W.partial.p, k::nblocks(n), scal + 12);
    k::axpy_const(s, n, scal + 12, -1.0 / (double)n, W.r.p);
  }
  k::fill<double>(s, n, W.x.p, 0.0);
  k::fill<double>(s, n, W.w.p, 0.0);
  k::dot_partial(s, n, W.r.p, W.r.p, W.partial.p);
  k::reduce_final(s, W.partial.p, k::nblocks(n), scal + 10);
  k::convert<double, V>(s, n, W.r.p, Prec<V>::b(W.lev[0]));
  k::ring_put(s, scal + 10, W.dring + 4);
}

template <typename V>
void FcgRun::start_t(const double *b) {
  fcg_setup<V>(s_, *S_, *W_, b);
  UFVM_HIP(hipEventRecord(W_->ev[2], s_));
}

void FcgRun::start(const double *b, double *x) {
  if (!S_->pcg.ready || S_->pcg.stale)
    throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
  if (!S_->pcg.lev[0]->soff.p) throw Error(UFVM_ERR_STATE, "level 0 has no sliced-ELL copy");
  x_out_ = x;
  it_ = issued_ = 0;
  ahead_ = true;
  prev2_ = -1.0;
  nb_ = thr_ = rr_ = 0.0;
  if (S_->pcg.cfg.precision != 64) start_t<float>(b);
  else start_t<double>(b);
  phase_ = 1;
}

// One FCG(1) iteration k: z = M^-1 r (V-cycle), dots, p update, w = A p,
// x / r update.  The last kernel writes {r.r, k} to ring slot k & 1 and sets
// the device convergence flag, after which a queued iteration changes nothing.
// Only k == 0 versus k > 0 changes what is launched (every per-solve value
// lives in device scalars), so iterations can be captured and replayed.
template <typename V>
void FcgRun::issue_t(int k) {
  const PcgState &P = S_->pcg;
  const FastLevel &L0 = *P.lev[0];
  SolveWork &W = *W_;
  const int32_t n = W.n0;
  double *scal = W.scal.p;
  // iterations after the first find their first colour pass made by the
  // previous iteration's x/r update (multicolour pre-smoothing only)
  const bool fuse0 = L0.kind == 0 && P.cfg.pre_sweep > 0;
  Cycle<V> cyc{ctx_, *S_, W, s_};
  const V *z;
  {
    ProfScope ps(ctx_, s_, "vcycle", 0.0);
    z = cyc.vcycle(0, fuse0 && k > 0);
  }
  k::fcg_dots<V>(s_, n, z, W.w.p, W.r.p, W.partial.p);
  k::fcg_update_p<V>(s_, n, z, W.partial.p, P.singular, k == 0, scal, W.p.p);
  {
    ProfScope ps(ctx_, s_, "spmv0",
                 ctx_->prof_on ? 4.0 * (n / 64 + 1) + (8.0 + index_bytes(L0)) * L0.nnz + 8.0 * 3 * n
                               : 0.0);
    k::spmv_dot(s_, n, L0.soff.p, sell_cols(L0), L0.ac64.p, L0.sval64.p, W.p.p, W.w.p, W.partial.p);
  }
  k::fcg_update_xr<V>(s_, n, scal, W.p.p, W.w.p, W.x.p, W.r.p, Prec<V>::b(W.lev[0]), W.partial.p,
                      W.dring, fuse0 ? L0.col_off[1] : 0, Prec<V>::ac(L0),
                      Prec<V>::x(W.lev[0]));
}

void FcgRun::issue(int k) {
  const bool f32 = S_->pcg.cfg.precision != 64;
  if (!graphs_) {
    if (f32) issue_t<float>(k);
    else issue_t<double>(k);
  } else {
    hipGraphExec_t &gx = W_->iter_graph[k == 0 ? 0 : 1];
    if (!gx) {
      // capture this iteration's launches once per workspace, then replay
      hipGraph_t g = nullptr;
      // no other host thread may allocate, free or synchronise the device
      // while this stream captures (HIP refuses those during any capture)
      std::lock_guard<std::mutex> lk(device_mgmt_mutex());
      UFVM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeRelaxed));
      try {
        if (f32) issue_t<float>(k);
        else issue_t<double>(k);
      } catch (...) {
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      UFVM_HIP(hipStreamEndCapture(s_, &g));
      const hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      UFVM_HIP(e);
    }
    UFVM_HIP(hipGraphLaunch(gx, s_));
  }
  UFVM_HIP(hipEventRecord(W_->ev[k & 1], s_));
}

void FcgRun::finish() {
  k::permute_scatter(s_, W_->n0, S_->pcg.lev[0]->perm_d.p, W_->x.p, x_out_);
  UFVM_HIP(hipEventRecord(W_->ev[2], s_));
  phase_ = 3;
}

bool FcgRun::poll(bool block) {
  const ufvm_pcg_cfg &cfg = S_->pcg.cfg;
  for (;;) {
    if (phase_ == 4) return true;
    if (phase_ == 3) {
      if (!ready(W_->ev[2], block)) return false;
      phase_ = 4;
      continue;
    }
    if (phase_ == 1) {
      if (!ready(W_->ev[2], block)) return false;
      nb_ = std::sqrt(W_->hring[4]);
      if (!(nb_ > 0.0) || cfg.max_iter <= 0) {
        if (std::isnan(nb_)) throw Error(UFVM_ERR_NUMERIC, "NaN right-hand side");
        finish();
        continue;
      }
      thr_ = (cfg.tol * nb_) * (cfg.tol * nb_);
      rr_ = nb_;
      k::fcg_arm(s_, W_->scal.p, thr_);
      phase_ = 2;
    }
    // phase 2: keep one iteration queued beyond the one awaited unless the
    // residual history predicts that the awaited one converges
    while (issued_ < cfg.max_iter && issued_ < it_ + (ahead_ ? 2 : 1)) issue(issued_++);
    if (!ready(W_->ev[it_ & 1], block)) return false;
    const volatile double *slot = W_->hring + 2 * (it_ & 1);
    if (slot[1] != (double)it_) throw Error(UFVM_ERR_STATE, "FPCG residual ring out of step");
    const double rr2 = slot[0];   // the value the device compared with thr
    rr_ = std::sqrt(rr2);
    ++it_;
    if (std::isnan(rr_)) throw Error(UFVM_ERR_NUMERIC, "FPCG breakdown (NaN residual)");
    if (rr2 <= thr_ || it_ >= cfg.max_iter) {
      finish();
      continue;
    }
    // at most iteration it_ is in flight: predict its residual from the last
    // reduction factor
    if (prev2_ > 0.0) ahead_ = rr2 * (rr2 / prev2_) > 4.0 * thr_;
    prev2_ = rr2;
  }
}

// ------------------------------------------------------------ batched solves

// The V-cycle of Cycle<V> for the k right-hand sides of a batch workspace:
// finest level multicolour GS, damped-Jacobi coarse levels, dense coarsest
// level (the configurations a batch accepts), every launch covering all k.
template <typename V>
struct CycleB {
  using P = Prec<V>;
  System &S;
  SolveWork &W;
  hipStream_t s;
  int k() const { return W.nrhs; }

  int pre_of(int l) const {
    const ufvm_pcg_cfg &c = S.pcg.cfg;
    return (l == 0 || c.coarse_pre_sweep < 0) ? c.pre_sweep : c.coarse_pre_sweep;
  }
  int post_of(int l) const {
    const ufvm_pcg_cfg &c = S.pcg.cfg;
    return (l == 0 || c.coarse_post_sweep < 0) ? c.post_sweep : c.coarse_post_sweep;
  }

  void mcgs_sweeps(const FastLevel &L, LevelWork &LW, int nsweep, bool from_zero,
                   bool first_done) {
    auto pass = [&](int c, int32_t zero_above) {
      k::mcgs_sell_b<V>(s, k(), L.col_off[c], L.col_off[c + 1], L.soff.p, sell_cols(L), P::sval(L),
                        P::ac(L), P::b(LW), P::x(LW), LW.ld, zero_above);
    };
    for (int sw = 0; sw < nsweep; ++sw) {
      for (int c = (sw == 0 ? (first_done ? 1 : 0) : 1); c < L.ncol; ++c)
        pass(c, (sw == 0 && from_zero) ? L.col_off[c + 1] : L.n);
      for (int c = L.ncol - 2; c >= 0; --c) pass(c, L.n);
    }
  }

  void down(const FastLevel &L, LevelWork &LW, int l, const V *x, int32_t zero_rows) {
    const PcgState &Pc = S.pcg;
    const FastLevel &C = *Pc.lev[l + 1];
    LevelWork &CW = W.lev[l + 1];
    k::residual_sell_b<V>(s, k(), L.n, zero_rows, L.soff.p, sell_cols(L), P::sval(L), P::ac(L),
                          P::b(LW), x, P::r(LW), LW.ld);
    const bool jac = C.kind == 3 && pre_of(l + 1) > 0;
    k::restrict_b<V>(s, k(), C.n, L.csoff.p, L.cschild.p, P::r(LW), LW.ld,
                     jac ? P::ac(C) : nullptr, P::b(CW), P::x(CW), CW.ld, V(2.0 / 3.0),
                     V(Pc.cfg.coarse_scale));
  }

  V *vcycle(int l, bool first_done = false) {
    const PcgState &Pc = S.pcg;
    const FastLevel &L = *Pc.lev[l];
    LevelWork &LW = W.lev[l];
    const int pre = pre_of(l), post = post_of(l);
    if (l == Pc.nlev - 1) {
      k::dense_solve_b<V>(s, k(), L.n, Pc.ainv.p, P::b(LW), LW.ld, P::x(LW));
      return P::x(LW);
    }
    V *x = P::x(LW);
    const V omega = V(2.0 / 3.0);
    if (L.kind == 3) {
      V *y = P::y(LW);
      for (int it = 1; it < pre; ++it) {
        k::jacobi_sell_b<V>(s, k(), L.n, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW), x,
                            LW.ld, nullptr, nullptr, 0, y, omega);
        std::swap(x, y);
      }
      if (pre == 0) k::fill<V>(s, LW.ld * k(), x, V(0));
      down(L, LW, l, x, 0);
      V *xc = vcycle(l + 1);
      const int64_t ldc = W.lev[l + 1].ld;
      const bool fuse = post > 0 && L.n <= kFusedProlongRows;
      if (!fuse) k::prolong_b<V>(s, k(), 0, L.n, L.parent.p, xc, ldc, x, LW.ld);
      for (int it = 0; it < post; ++it) {
        const bool pro = fuse && it == 0;
        k::jacobi_sell_b<V>(s, k(), L.n, L.soff.p, sell_cols(L), P::sval(L), P::ac(L), P::b(LW), x,
                            LW.ld, pro ? L.parent.p : nullptr, pro ? xc : nullptr, ldc, y,
                            omega);
        std::swap(x, y);
      }
      return x;
    }
    // kind 0: the finest level
    if (pre == 0) k::fill<V>(s, LW.ld * k(), x, V(0));
    else mcgs_sweeps(L, LW, pre, true, first_done);
    down(L, LW, l, x, pre > 0 ? L.col_off[1] : 0);
    V *xc = vcycle(l + 1);
    k::prolong_b<V>(s, k(), post > 0 ? L.col_off[1] : 0, L.n, L.parent.p, xc, W.lev[l + 1].ld, x,
                    LW.ld);
    if (post > 0) mcgs_sweeps(L, LW, post, false, false);
    return x;
  }
};

bool batch_supported(const System &S) {
  const PcgState &P = S.pcg;
  if (!P.ready || P.cfg.precision == 16) return false;
  if (P.nlev > 1 && P.lev[0]->kind != 0) return false;
  for (int32_t l = 1; l + 1 < P.nlev; ++l)
    if (P.lev[l]->kind != 3) return false;
  return true;
}

FcgBatch::FcgBatch(ufvm_ctx *ctx, System &S, SolveWork &W, hipStream_t s)
    : ctx_(ctx), S_(&S), W_(&W), s_(s), k_(W.nrhs), nb_(k_), thr_(k_), rr_(k_), its_(k_),
      done_(k_) {
  const PcgState &P = S.pcg;
  if (P.cfg.precision == 16)
    throw Error(UFVM_ERR_UNSUPPORTED, "batched solves need precision 32 or 64");
  for (int32_t l = 1; l + 1 < P.nlev; ++l)
    if (P.lev[l]->kind != 3)
      throw Error(UFVM_ERR_UNSUPPORTED, "batched solves need coarse_smoother 1 (Jacobi)");
  if (P.nlev > 1 && P.lev[0]->kind != 0)
    throw Error(UFVM_ERR_UNSUPPORTED, "batched solves need the multicolour GS finest smoother");
}

k::BatchStrides FcgBatch::strides() const {
  k::BatchStrides bs;
  bs.ld = W_->ld0;
  bs.ldv = W_->lev[0].ld;
  bs.ps = partial_stride(W_->n0);
  bs.ss = kScal;
  bs.rs = kRingPer;
  return bs;
}

template <typename V>
void FcgBatch::start_t() {
  const PcgState &P = S_->pcg;
  const FastLevel &L0 = *P.lev[0];
  SolveWork &W = *W_;
  const int32_t n = W.n0;
  const int64_t ps = partial_stride(n);
  // per right-hand side, the start-up of a single solve (same kernels and
  // partial layout, so the same bits)
  for (int q = 0; q < k_; ++q) {
    double *r = W.r.p + q * W.ld0, *part = W.partial.p + q * ps, *scal = W.scal.p + q * kScal;
    k::permute_gather(s_, n, L0.perm_d.p, W.bref.p + (int64_t)q * n, r);
    if (P.singular) {
      k::sum_partial(s_, n, r, part);
      k::reduce_final(s_, part, k::nblocks(n), scal + 12);
      k::axpy_const(s_, n, scal + 12, -1.0 / (double)n, r);
    }
    k::dot_partial(s_, n, r, r, part);
    k::reduce_final(s_, part, k::nblocks(n), scal + 10);
    k::convert<double, V>(s_, n, r, Prec<V>::b(W.lev[0]) + q * W.lev[0].ld);
    k::ring_put(s_, scal + 10, W.dring + q * kRingPer + 4);
  }
  k::fill<double>(s_, (int64_t)W.ld0 * k_, W.x.p, 0.0);
  k::fill<double>(s_, (int64_t)W.ld0 * k_, W.w.p, 0.0);
  UFVM_HIP(hipEventRecord(W.ev[2], s_));
}

void FcgBatch::start(const double *B, double *X) {
  const PcgState &P = S_->pcg;
  if (!P.ready || P.stale)
    throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
  if (!B || !X) throw Error(UFVM_ERR_ARG, "null right-hand sides or solutions");
  X_ = X;
  it_ = issued_ = 0;
  if (B != W_->bref.p)
    UFVM_HIP(hipMemcpyAsync(W_->bref.p, B, sizeof(double) * W_->n0 * k_, hipMemcpyDefault, s_));
  if (P.cfg.precision != 64) start_t<float>();
  else start_t<double>();
  phase_ = 1;
}

template <typename V>
void FcgBatch::issue_t(int it) {
  const PcgState &P = S_->pcg;
  const FastLevel &L0 = *P.lev[0];
  SolveWork &W = *W_;
  const int32_t n = W.n0;
  const k::BatchStrides bs = strides();
  const bool fuse0 = L0.kind == 0 && P.cfg.pre_sweep > 0;
  CycleB<V> cyc{*S_, W, s_};
  const V *z = cyc.vcycle(0, fuse0 && it > 0);
  k::fcg_dots<V>(s_, n, z, W.w.p, W.r.p, W.partial.p, k_, bs);
  k::fcg_update_p<V>(s_, n, z, W.partial.p, P.singular, it == 0, W.scal.p, W.p.p, k_, bs);
  k::spmv_dot_b(s_, k_, n, L0.soff.p, sell_cols(L0), L0.ac64.p, L0.sval64.p, W.p.p, W.w.p, W.ld0,
                W.partial.p, bs.ps);
  k::fcg_update_xr<V>(s_, n, W.scal.p, W.p.p, W.w.p, W.x.p, W.r.p, Prec<V>::b(W.lev[0]),
                      W.partial.p, W.dring, fuse0 ? L0.col_off[1] : 0, Prec<V>::ac(L0),
                      Prec<V>::x(W.lev[0]), k_, bs);
  UFVM_HIP(hipEventRecord(W.ev[it & 1], s_));
}

void FcgBatch::finish() {
  const FastLevel &L0 = *S_->pcg.lev[0];
  SolveWork &W = *W_;
  for (int q = 0; q < k_; ++q)
    k::permute_scatter(s_, W.n0, L0.perm_d.p, W.x.p + q * W.ld0, W.xref.p + (int64_t)q * W.n0);
  if (X_ != W.xref.p)
    UFVM_HIP(hipMemcpyAsync(X_, W.xref.p, sizeof(double) * W.n0 * k_, hipMemcpyDefault, s_));
  UFVM_HIP(hipEventRecord(W.ev[2], s_));
  phase_ = 3;
}

bool FcgBatch::poll(bool block) {
  const ufvm_pcg_cfg &cfg = S_->pcg.cfg;
  auto ready = [&](hipEvent_t e) {
    if (block) {
      UFVM_HIP(hipEventSynchronize(e));
      return true;
    }
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return true;
    if (q != hipErrorNotReady) UFVM_HIP(q);
    return false;
  };
  for (;;) {
    if (phase_ == 4) return true;
    if (phase_ == 3) {
      if (!ready(W_->ev[2])) return false;
      phase_ = 4;
      continue;
    }
    if (phase_ == 1) {
      if (!ready(W_->ev[2])) return false;
      int left = 0;
      for (int q = 0; q < k_; ++q) {
        nb_[q] = std::sqrt(W_->hring[q * kRingPer + 4]);
        if (std::isnan(nb_[q])) throw Error(UFVM_ERR_NUMERIC, "NaN right-hand side");
        thr_[q] = (cfg.tol * nb_[q]) * (cfg.tol * nb_[q]);
        rr_[q] = nb_[q];
        its_[q] = 0;
        // a zero right-hand side is solved by x = 0: its updates stay off
        done_[q] = !(nb_[q] > 0.0) || cfg.max_iter <= 0;
        k::fcg_arm(s_, W_->scal.p + q * kScal, thr_[q], done_[q] ? 1 : 0);
        left += !done_[q];
      }
      if (left == 0) {
        finish();
        continue;
      }
      phase_ = 2;
    }
    // phase 2: one iteration queued beyond the awaited one
    while (issued_ < cfg.max_iter && issued_ < it_ + 2) {
      if (S_->pcg.cfg.precision != 64) issue_t<float>(issued_);
      else issue_t<double>(issued_);
      ++issued_;
    }
    if (!ready(W_->ev[it_ & 1])) return false;
    int left = 0;
    for (int q = 0; q < k_; ++q) {
      const volatile double *slot = W_->hring + q * kRingPer + 2 * (it_ & 1);
      if (slot[1] != (double)it_) throw Error(UFVM_ERR_STATE, "FPCG residual ring out of step");
      if (done_[q]) continue;
      const double rr2 = slot[0];
      rr_[q] = std::sqrt(rr2);
      its_[q] = it_ + 1;
      if (std::isnan(rr_[q])) throw Error(UFVM_ERR_NUMERIC, "FPCG breakdown (NaN residual)");
      if (rr2 <= thr_[q]) done_[q] = true;
      left += !done_[q];
    }
    ++it_;
    if (left == 0 || it_ >= cfg.max_iter) {
      finish();
      continue;
    }
  }
}

// ufvm_pcg_solve_batch: k right-hand sides (RHS-major, reference order, host
// or device memory) in, k solutions out, on the context stream
void pcg_solve_batch(ufvm_ctx *ctx, System &S, int k, const double *B, double *X,
                     int32_t *iters, double *relres) {
  PcgState &P = S.pcg;
  if (!P.ready || P.stale)
    throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
  if (k != 1 && k != 2 && k != 4 && k != 8)
    throw Error(UFVM_ERR_ARG, "a batch holds 1, 2, 4 or 8 right-hand sides");
  std::unique_ptr<SolveWork> &W = P.work_batch[k];
  if (!W) W = make_work(S, ctx->stream, k);
  FcgBatch run(ctx, S, *W, ctx->stream);
  run.start(B, X);
  while (!run.poll(true)) {
  }
  for (int q = 0; q < k; ++q) {
    if (iters) iters[q] = run.iterations(q);
    if (relres) relres[q] = run.relres(q);
  }
}

// ufvm_pcg_solve: the level-1 system's bc in, dphi out, on the context stream
void pcg_solve(ufvm_ctx *ctx, System &S, int32_t *iters, double *relres) {
  if (!S.pcg.ready || S.pcg.stale || !S.pcg.work0)
    throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
  RefLevel &R0 = *S.lev[0];
  FcgRun run(ctx, S, *S.pcg.work0, ctx->stream);
  run.start(R0.bc.p, R0.dphi.p);
  while (!run.poll(true)) {
  }
  if (iters) *iters = run.iterations();
  if (relres) *relres = run.relres();
}

template <typename V>
static void apply_vcycle_t(ufvm_ctx *ctx, System &S, SolveWork &W) {
  hipStream_t s = ctx->stream;
  const int32_t n = W.n0;
  const FastLevel &L0 = *S.pcg.lev[0];
  k::permute_gather(s, n, L0.perm_d.p, W.bref.p, W.p.p);
  k::convert<double, V>(s, n, W.p.p, Prec<V>::b(W.lev[0]));
  Cycle<V> cyc{ctx, S, W, s};
  const V *z = cyc.vcycle(0);
  k::convert<V, double>(s, n, z, W.p.p);
  k::permute_scatter(s, n, L0.perm_d.p, W.p.p, W.xref.p);
}

void pcg_apply_vcycle(ufvm_ctx *ctx, System &S, const double *r, double *z) {
  if (!S.pcg.ready || S.pcg.stale || !S.pcg.work0)
    throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
  if (!r || !z) throw Error(UFVM_ERR_ARG, "null vector");
  SolveWork &W = *S.pcg.work0;
  hipStream_t s = ctx->stream;
  const int32_t n = W.n0;
  UFVM_HIP(hipMemcpyAsync(W.bref.p, r, n * sizeof(double), hipMemcpyHostToDevice, s));
  if (S.pcg.cfg.precision != 64) apply_vcycle_t<float>(ctx, S, W);
  else apply_vcycle_t<double>(ctx, S, W);
  UFVM_HIP(hipMemcpyAsync(z, W.xref.p, n * sizeof(double), hipMemcpyDeviceToHost, s));
  UFVM_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------- group solves

bool group_supported(const System &S) {
  if (!batch_supported(S)) return false;
  const PcgState &P = S.pcg;
  for (int32_t l = 0; l + 1 < P.nlev; ++l) {
    const int pre = (l == 0 || P.cfg.coarse_pre_sweep < 0) ? P.cfg.pre_sweep
                                                          : P.cfg.coarse_pre_sweep;
    if (pre < 1) return false;
  }
  return true;
}

namespace {

// Phase key of a V-cycle launch: launches of different members with equal
// keys become one launch.  Keys order the descent by level, then every dense
// coarsest solve, then the ascent by level from the coarsest; inside a level
// they follow Cycle<V>'s order, so each member's launches keep their order.
struct GKey {
  int32_t stage, level, a, b, c;
  bool operator<(const GKey &o) const {
    if (stage != o.stage) return stage < o.stage;
    if (level != o.level) return level < o.level;
    if (a != o.a) return a < o.a;
    if (b != o.b) return b < o.b;
    return c < o.c;
  }
};

// Records the launches of one member's V-cycle exactly as Cycle<V> makes them
// for the kinds a group accepts (multicolour GS finest level, Jacobi coarse
// levels, dense coarsest level), with each launch's arguments.
template <typename V>
struct GroupRec {
  using P = Prec<V>;
  using A = k::GArgs<V>;
  System &S;
  SolveWork &W;
  const double *done;
  std::map<GKey, std::pair<int32_t, std::vector<std::pair<int, A>>>> &out;
  int member;

  int pre_of(int l) const {
    const ufvm_pcg_cfg &c = S.pcg.cfg;
    return (l == 0 || c.coarse_pre_sweep < 0) ? c.pre_sweep : c.coarse_pre_sweep;
  }
  int post_of(int l) const {
    const ufvm_pcg_cfg &c = S.pcg.cfg;
    return (l == 0 || c.coarse_post_sweep < 0) ? c.post_sweep : c.coarse_post_sweep;
  }
  A blank() const {
    A a{};
    a.done = done;
    return a;
  }
  void add(const GKey &key, k::GOp op, const A &a) {
    auto &e = out[key];
    if (!e.second.empty() && e.first != (int32_t)op)
      throw Error(UFVM_ERR_STATE, "group schedule: two operations share a phase");
    e.first = (int32_t)op;
    e.second.emplace_back(member, a);
  }
  static int32_t up(int l) { return 1000 - l; }

  void pass(const FastLevel &L, LevelWork &LW, const GKey &key, int c, int32_t zero_above) {
    const int32_t r0 = L.col_off[c], r1 = L.col_off[c + 1];
    if (r1 <= r0) return;
    A a = blank();
    a.n = r1;
    a.r0 = r0;
    a.aux = zero_above;
    a.nb = k::nblocks(r1 - (r0 & ~63));
    a.off = L.soff.p;
    a.cols = sell_cols(L);
    a.val = P::sval(L);
    a.ac = P::ac(L);
    a.b = P::b(LW);
    a.out = P::x(LW);
    add(key, k::G_MCGS, a);
  }
  void mcgs_sweeps(const FastLevel &L, LevelWork &LW, int nsweep, bool from_zero,
                   bool first_done, int32_t stage, int32_t level) {
    for (int sw = 0; sw < nsweep; ++sw) {
      for (int c = (sw == 0 ? (first_done ? 1 : 0) : 1); c < L.ncol; ++c)
        pass(L, LW, {stage, level, 1, 2 * sw, c}, c,
             (sw == 0 && from_zero) ? L.col_off[c + 1] : L.n);
      for (int c = L.ncol - 2; c >= 0; --c) pass(L, LW, {stage, level, 1, 2 * sw + 1, 1000 - c}, c, L.n);
    }
  }
  void down(const FastLevel &L, LevelWork &LW, int l, const V *x, int32_t zero_rows) {
    const PcgState &Pc = S.pcg;
    const FastLevel &C = *Pc.lev[l + 1];
    LevelWork &CW = W.lev[l + 1];
    A r = blank();
    r.n = L.n;
    r.aux = zero_rows;
    r.nb = k::nblocks(L.n);
    r.off = L.soff.p;
    r.cols = sell_cols(L);
    r.val = P::sval(L);
    r.ac = P::ac(L);
    r.b = P::b(LW);
    r.x = x;
    r.out = P::r(LW);
    r.wide = L.wide;
    add({0, l, 2, 0, 0}, k::G_RESID, r);
    A t = blank();
    t.n = C.n;
    t.nb = k::nblocks(C.n);
    t.off = L.csoff.p;
    t.child = L.cschild.p;
    t.x = P::r(LW);
    t.out = P::b(CW);
    t.scale = V(Pc.cfg.coarse_scale);
    if (C.kind == 3 && pre_of(l + 1) > 0) {
      t.ac = P::ac(C);
      t.out2 = P::x(CW);
      t.omega = V(2.0 / 3.0);
    }
    add({0, l, 3, 0, 0}, k::G_RESTRICT, t);
  }
  const V *vcycle(int l, bool first_done) {
    const PcgState &Pc = S.pcg;
    const FastLevel &L = *Pc.lev[l];
    LevelWork &LW = W.lev[l];
    const int pre = pre_of(l), post = post_of(l);
    if (l == Pc.nlev - 1) {
      A a = blank();
      a.n = L.n;
      a.nb = k::dense_blocks(L.n);
      a.ainv = Pc.ainv.p;
      a.b = P::b(LW);
      a.out = P::x(LW);
      add({1, 0, 0, 0, 0}, k::G_DENSE, a);
      return P::x(LW);
    }
    V *x = P::x(LW);
    const V omega = V(2.0 / 3.0);
    auto jac = [&](const GKey &key, const V *in, const int32_t *parent, const V *xc, V *outp) {
      A a = blank();
      a.n = L.n;
      a.nb = k::nblocks(L.n);
      a.off = L.soff.p;
      a.cols = sell_cols(L);
      a.val = P::sval(L);
      a.ac = P::ac(L);
      a.b = P::b(LW);
      a.x = in;
      a.parent = parent;
      a.xc = xc;
      a.out = outp;
      a.omega = omega;
      a.wide = L.wide;
      add(key, k::G_JACOBI, a);
    };
    auto prolong = [&](int32_t r0, const V *xc, V *xp) {
      if (L.n <= r0) return;
      A a = blank();
      a.n = L.n;
      a.r0 = r0;
      a.nb = k::nblocks(L.n - r0);
      a.parent = L.parent.p;
      a.xc = xc;
      a.out = xp;
      add({2, up(l), 0, 0, 0}, k::G_PROLONG, a);
    };
    if (L.kind == 3) {
      V *y = P::y(LW);
      for (int it = 1; it < pre; ++it) {
        jac({0, l, 1, it, 0}, x, nullptr, nullptr, y);
        std::swap(x, y);
      }
      down(L, LW, l, x, 0);
      const V *xc = vcycle(l + 1, false);
      const bool fuse = post > 0 && L.n <= kFusedProlongRows;
      if (!fuse) prolong(0, xc, x);
      for (int it = 0; it < post; ++it) {
        const bool pro = fuse && it == 0;
        jac({2, up(l), 1, it, 0}, x, pro ? L.parent.p : nullptr, pro ? xc : nullptr, y);
        std::swap(x, y);
      }
      return x;
    }
    // kind 0: the finest level, multicolour Gauss-Seidel
    mcgs_sweeps(L, LW, pre, true, first_done, 0, l);
    down(L, LW, l, x, pre > 0 ? L.col_off[1] : 0);
    const V *xc = vcycle(l + 1, false);
    prolong(post > 0 ? L.col_off[1] : 0, xc, x);
    if (post > 0) mcgs_sweeps(L, LW, post, false, false, 2, up(l));
    return x;
  }
};

}  // namespace

FcgGroup::FcgGroup(ufvm_ctx *ctx, std::vector<Member> members, hipStream_t s)
    : ctx_(ctx), m_(std::move(members)), st_(m_.size()), s_(s) {
  if (m_.empty()) throw Error(UFVM_ERR_ARG, "an empty group");
  for (size_t q = 0; q < m_.size(); ++q) {
    const Member &mb = m_[q];
    if (!mb.S || !mb.W || !mb.b || !mb.x) throw Error(UFVM_ERR_ARG, "incomplete group member");
    const PcgState &P = mb.S->pcg;
    if (!P.ready || P.stale)
      throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
    if (!group_supported(*mb.S))
      throw Error(UFVM_ERR_UNSUPPORTED,
                  "group solves need the multicolour GS finest smoother, Jacobi coarse levels "
                  "and pre-smoothing on every level");
    if (mb.W->nrhs != 1 || mb.W->n0 != mb.S->lev[0]->n)
      throw Error(UFVM_ERR_ARG, "a group member needs a single-solve workspace of its system");
    if (q == 0) {
      precision_ = P.cfg.precision;
      max_iter_ = P.cfg.max_iter;
    } else if (P.cfg.precision != precision_ || P.cfg.max_iter != max_iter_) {
      throw Error(UFVM_ERR_UNSUPPORTED, "group members need one V-cycle precision and max_iter");
    }
    for (size_t o = 0; o < q; ++o)
      if (m_[o].W == mb.W) throw Error(UFVM_ERR_ARG, "two group members share a workspace");
  }
  // a group lives for one solve each, so a captured iteration would be
  // replayed only some tens of times: graphs only when asked for (launch_mode 2)
  graphs_ = !ctx->prof_on && m_[0].S->pcg.cfg.launch_mode == 2;
  for (hipEvent_t &e : ev_) UFVM_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  for (MemberState &ms : st_) UFVM_HIP(hipEventCreateWithFlags(&ms.ev, hipEventDisableTiming));
  if (precision_ == 32) build<float>();
  else build<double>();
}

FcgGroup::~FcgGroup() {
  for (hipGraphExec_t &g : graph_)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipEvent_t &e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (MemberState &ms : st_)
    if (ms.ev) (void)hipEventDestroy(ms.ev);
}

// Both iteration schedules (the first, whose pre-smoother starts from colour
// 0, and the later ones, whose colour-0 pass the previous x/r update made)
// and the FCG table, in one device buffer of argument tables.
template <typename V>
void FcgGroup::build() {
  std::vector<char> host;
  auto put = [&](const void *p, size_t bytes) {
    const size_t at = (host.size() + 255) & ~(size_t)255;
    host.resize(at + bytes);
    std::memcpy(host.data() + at, p, bytes);
    return at;
  };
  for (int kind = 0; kind < 2; ++kind) {
    std::map<GKey, std::pair<int32_t, std::vector<std::pair<int, k::GArgs<V>>>>> sched;
    for (int q = 0; q < (int)m_.size(); ++q) {
      System &S = *m_[q].S;
      SolveWork &W = *m_[q].W;
      const FastLevel &L0 = *S.pcg.lev[0];
      const bool fuse0 = L0.kind == 0 && S.pcg.cfg.pre_sweep > 0;
      GroupRec<V> rec{S, W, W.scal.p + 11, sched, q};
      const V *z = rec.vcycle(0, fuse0 && kind == 1);
      // the V-cycle's result is the finest level's x on every path
      if (z != Prec<V>::x(W.lev[0]))
        throw Error(UFVM_ERR_STATE, "group schedule: unexpected V-cycle result vector");
    }
    for (auto &e : sched) {
      std::vector<k::GArgs<V>> tab;
      int32_t gx = 0;
      for (auto &pm : e.second.second) {
        tab.push_back(pm.second);
        gx = std::max(gx, pm.second.nb);
      }
      const size_t at = put(tab.data(), tab.size() * sizeof(k::GArgs<V>));
      steps_[kind].push_back({e.second.first, gx, (int32_t)tab.size(), at});
    }
  }
  std::vector<k::GFcgArgs<V>> f(m_.size());
  for (size_t q = 0; q < m_.size(); ++q) {
    System &S = *m_[q].S;
    SolveWork &W = *m_[q].W;
    const FastLevel &L0 = *S.pcg.lev[0];
    const bool fuse0 = L0.kind == 0 && S.pcg.cfg.pre_sweep > 0;
    k::GFcgArgs<V> &a = f[q];
    a.n = W.n0;
    a.nb = k::fold_blocks(W.n0);
    a.n0 = fuse0 ? L0.col_off[1] : 0;
    a.singular = S.pcg.singular;
    a.z = Prec<V>::x(W.lev[0]);
    a.acv = Prec<V>::ac(L0);
    a.rv = Prec<V>::b(W.lev[0]);
    a.z0 = Prec<V>::x(W.lev[0]);
    a.off = L0.soff.p;
    a.cols = sell_cols(L0);
    a.ac = L0.ac64.p;
    a.val = L0.sval64.p;
    a.p = W.p.p;
    a.w = W.w.p;
    a.x = W.x.p;
    a.r = W.r.p;
    a.partial = W.partial.p;
    a.scal = W.scal.p;
    a.ring = W.dring;
    fcg_gx_ = std::max(fcg_gx_, a.nb);
  }
  fcg_table_ = put(f.data(), f.size() * sizeof(k::GFcgArgs<V>));
  tables_.upload(host.data(), host.size(), s_);
}

template <typename V>
void FcgGroup::issue_t(int k) {
  const int kind = k == 0 ? 0 : 1;
  for (const Step &st : steps_[kind]) {
    k::group_row<V>(s_, (k::GOp)st.op, st.gx, st.g,
                    reinterpret_cast<const k::GArgs<V> *>(tables_.p + st.table));
  }
  const auto *ft = reinterpret_cast<const k::GFcgArgs<V> *>(tables_.p + fcg_table_);
  const int32_t g = (int32_t)m_.size();
  k::group_fcg<V>(s_, k::G_DOTS, fcg_gx_, g, ft);
  k::group_fcg<V>(s_, k == 0 ? k::G_UPDATE_P_FIRST : k::G_UPDATE_P, fcg_gx_, g, ft);
  k::group_fcg<V>(s_, k::G_SPMV, fcg_gx_, g, ft);
  k::group_fcg<V>(s_, k::G_UPDATE_XR, fcg_gx_, g, ft);
  k::group_fcg<V>(s_, k::G_RR, fcg_gx_, g, ft);
}

void FcgGroup::issue(int k) {
  const bool f32 = precision_ == 32;
  if (!graphs_) {
    if (f32) issue_t<float>(k);
    else issue_t<double>(k);
  } else {
    hipGraphExec_t &gx = graph_[k == 0 ? 0 : 1];
    if (!gx) {
      hipGraph_t g = nullptr;
      std::lock_guard<std::mutex> lk(device_mgmt_mutex());
      UFVM_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeRelaxed));
      try {
        if (f32) issue_t<float>(k);
        else issue_t<double>(k);
      } catch (...) {
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      UFVM_HIP(hipStreamEndCapture(s_, &g));
      const hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      UFVM_HIP(e);
    }
    UFVM_HIP(hipGraphLaunch(gx, s_));
  }
  launches_ += (int64_t)steps_[k == 0 ? 0 : 1].size() + 5;
  UFVM_HIP(hipEventRecord(ev_[k & 1], s_));
}

void FcgGroup::start() {
  for (size_t q = 0; q < m_.size(); ++q) {
    if (precision_ == 32) fcg_setup<float>(s_, *m_[q].S, *m_[q].W, m_[q].b);
    else fcg_setup<double>(s_, *m_[q].S, *m_[q].W, m_[q].b);
  }
  UFVM_HIP(hipEventRecord(ev_[2], s_));
  it_ = issued_ = 0;
  phase_ = 1;
}

// the member's solution into its x (and host copy); its launches skip it from
// now on (its flag is set, or armed so for a member that never iterates)
void FcgGroup::finish_member(int q) {
  MemberState &ms = st_[q];
  const Member &mb = m_[q];
  k::permute_scatter(s_, mb.W->n0, mb.S->pcg.lev[0]->perm_d.p, mb.W->x.p, mb.x);
  if (mb.x_host)
    UFVM_HIP(hipMemcpyAsync(mb.x_host, mb.x, (size_t)mb.W->n0 * sizeof(double),
                            hipMemcpyDeviceToHost, s_));
  UFVM_HIP(hipEventRecord(ms.ev, s_));
  ms.active = false;
  ms.finished = true;
}

bool FcgGroup::member_done(int q) {
  MemberState &ms = st_[q];
  if (!ms.finished) return false;
  const hipError_t e = hipEventQuery(ms.ev);
  if (e == hipSuccess) return true;
  if (e != hipErrorNotReady) UFVM_HIP(e);
  return false;
}

bool FcgGroup::poll(bool block) {
  auto ready = [&](hipEvent_t e) {
    if (block) {
      UFVM_HIP(hipEventSynchronize(e));
      return true;
    }
    const hipError_t q = hipEventQuery(e);
    if (q == hipSuccess) return true;
    if (q != hipErrorNotReady) UFVM_HIP(q);
    return false;
  };
  for (;;) {
    if (phase_ == 4) return true;
    if (phase_ == 3) {
      for (size_t q = 0; q < m_.size(); ++q)
        if (!ready(st_[q].ev)) return false;
      phase_ = 4;
      continue;
    }
    if (phase_ == 1) {
      if (!ready(ev_[2])) return false;
      int left = 0;
      for (size_t q = 0; q < m_.size(); ++q) {
        MemberState &ms = st_[q];
        const ufvm_pcg_cfg &cfg = m_[q].S->pcg.cfg;
        ms.nb = std::sqrt(m_[q].W->hring[4]);
        ms.rr = ms.nb;
        ms.thr = (cfg.tol * ms.nb) * (cfg.tol * ms.nb);
        const bool nan = std::isnan(ms.nb);
        if (nan) ms.err = "NaN right-hand side";
        // a zero right-hand side is solved by x = 0; its launches stay off
        const bool idle = nan || !(ms.nb > 0.0) || cfg.max_iter <= 0;
        k::fcg_arm(s_, m_[q].W->scal.p, ms.thr, idle ? 1 : 0);
        if (idle) finish_member((int)q);
        else ms.active = true;
        left += ms.active;
      }
      phase_ = left ? 2 : 3;
      continue;
    }
    // phase 2: one iteration queued beyond the awaited one unless every
    // active member's residual history predicts that the awaited one converges
    bool ahead = false;
    for (const MemberState &ms : st_) ahead = ahead || (ms.active && ms.ahead);
    while (issued_ < max_iter_ && issued_ < it_ + (ahead ? 2 : 1)) issue(issued_++);
    if (!ready(ev_[it_ & 1])) return false;
    int left = 0;
    for (size_t q = 0; q < m_.size(); ++q) {
      MemberState &ms = st_[q];
      if (!ms.active) continue;
      const volatile double *slot = m_[q].W->hring + 2 * (it_ & 1);
      if (slot[1] != (double)it_) throw Error(UFVM_ERR_STATE, "FPCG residual ring out of step");
      const double rr2 = slot[0];
      ms.rr = std::sqrt(rr2);
      ms.its = it_ + 1;
      if (std::isnan(ms.rr)) {
        // the member's own solve would stop here with this error; the device
        // flag keeps its launches off from now on
        ms.err = "FPCG breakdown (NaN residual)";
        k::fcg_arm(s_, m_[q].W->scal.p, ms.thr, 1);
        finish_member((int)q);
        continue;
      }
      if (rr2 <= ms.thr || ms.its >= max_iter_) {
        finish_member((int)q);
        continue;
      }
      if (ms.prev2 > 0.0) ms.ahead = rr2 * (rr2 / ms.prev2) > 4.0 * ms.thr;
      ms.prev2 = rr2;
      ++left;
    }
    ++it_;
    if (left == 0) phase_ = 3;
  }
}

void pcg_solve_group(ufvm_ctx *ctx, const std::vector<System *> &sys, const double *const *b,
                     double *const *x, int32_t *iters, double *relres) {
  if (sys.empty()) throw Error(UFVM_ERR_ARG, "an empty group");
  if (!b || !x) throw Error(UFVM_ERR_ARG, "null right-hand sides or solutions");
  std::vector<FcgGroup::Member> m;
  for (size_t q = 0; q < sys.size(); ++q) {
    System &S = *sys[q];
    if (!S.pcg.ready || S.pcg.stale || !S.pcg.work0)
      throw Error(UFVM_ERR_STATE, "call ufvm_pcg_setup first (or refresh after new coefficients)");
    if (!b[q] || !x[q]) throw Error(UFVM_ERR_ARG, "null right-hand side or solution");
    for (size_t o = 0; o < q; ++o)
      if (sys[o] == sys[q])
        throw Error(UFVM_ERR_ARG, "a system appears twice (ufvm_pcg_solve_batch solves several "
                                  "right-hand sides of one system)");
    SolveWork &W = *S.pcg.work0;
    UFVM_HIP(hipMemcpyAsync(W.bref.p, b[q], (size_t)W.n0 * sizeof(double),
                            hipMemcpyHostToDevice, ctx->stream));
    m.push_back({&S, &W, W.bref.p, W.xref.p, x[q]});
  }
  FcgGroup grp(ctx, m, ctx->stream);
  grp.start();
  while (!grp.poll(true)) {
  }
  for (int q = 0; q < grp.size(); ++q) {
    if (!grp.member_error(q).empty()) throw Error(UFVM_ERR_NUMERIC, grp.member_error(q));
    if (iters) iters[q] = grp.iterations(q);
    if (relres) relres[q] = grp.relres(q);
  }
}

}  // namespace detail
}  // namespace ufvm
