// kernels.hip -- gfx950 kernels of the uFVM pressure-solve path.
//
// All of this work is sparse and HBM-bound: gathers over an unstructured
// cell graph, no dense contraction (MFMA is not used).  Two families:
//
//  * reference-order FP64 kernels.  They sum every row in the reference's
//    cconn order with separately rounded multiplies and adds (the library is
//    built with -ffp-contract=off), so residuals, restriction, prolongation,
//    Galerkin coarse operators and the level-scheduled lexicographic SOR /
//    DILU sweeps reproduce the reference arithmetic bit for bit.
//  * multicolour / precision-form kernels, templated on the value type, used
//    by the FP32-or-FP64 V-cycle preconditioner and the FP64 flexible CG.
//
// Reductions are deterministic: per-block partial sums in a fixed tree order,
// then one block sums the partials in a fixed order.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "reduce.h"

namespace ufvm {
namespace k {

// ------------------------------------------------------------ reference mode

// v - sum over the CSR entries [k0, k1) of anb[k] * xv(col[k]), in entry order
// (the reference rows' arithmetic).  The entries are taken in chunks whose
// index and value loads are issued together, then their neighbour loads, so a
// row waits for two round trips per chunk instead of two per entry; a chunk
// of N slots holding cnt entries repeats its last entry in the other slots
// (same address) and does not sum them.
template <int N, typename XV>
__device__ __forceinline__ double csr_chunk(const int32_t *__restrict__ col,
                                            const double *__restrict__ anb, const XV &xv,
                                            int32_t k0, int32_t cnt, double v) {
  int32_t c[N];
  double a[N], xj[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const int32_t e = k0 + (u < cnt ? u : cnt - 1);
    c[u] = col[e];
    a[u] = anb[e];
  }
#pragma unroll
  for (int u = 0; u < N; ++u) xj[u] = xv(c[u]);
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const double t = v - a[u] * xj[u];
    v = u < cnt ? t : v;
  }
  return v;
}

template <typename XV>
__device__ __forceinline__ double csr_row_sum(const int32_t *__restrict__ col,
                                              const double *__restrict__ anb, const XV &xv,
                                              int32_t k0, int32_t k1, double v) {
  if (k1 <= k0) return v;
  if (k1 - k0 <= 8) return csr_chunk<8>(col, anb, xv, k0, k1 - k0, v);
  for (; k0 + 16 < k1; k0 += 16) v = csr_chunk<16>(col, anb, xv, k0, 16, v);
  return csr_chunk<16>(col, anb, xv, k0, k1 - k0, v);
}

// The DILU rows' sums (cfdFactorizeILU.m / cfdSolveILU.m arithmetic, in list
// order) in chunks of kDiluChunk list entries: every index load of a chunk,
// then every value load, then the chunk's terms in order (slots past cnt
// repeat the last entry and are not summed).
constexpr int kDiluChunk = 8;

// v - sum over the lower-neighbour list [e0, e1) of anb[lpji[e]] * (dinv[i] *
// rv(i)), i = lnb[e] (the forward substitution)
template <typename V, typename RV>
__device__ __forceinline__ V dilu_lower_sum(const int32_t *__restrict__ lnb,
                                            const int32_t *__restrict__ lpji,
                                            const V *__restrict__ anb, const V *__restrict__ dinv,
                                            const RV &rv, int32_t e0, int32_t e1, V v) {
  constexpr int N = kDiluChunk;
  for (; e0 < e1; e0 += N) {
    const int32_t cnt = e1 - e0 < N ? e1 - e0 : N;
    int32_t ii[N], pj[N];
    V d[N], r[N], a[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const int32_t e = e0 + (u < cnt ? u : cnt - 1);
      ii[u] = lnb[e];
      pj[u] = lpji[e];
    }
#pragma unroll
    for (int u = 0; u < N; ++u) {
      d[u] = dinv[ii[u]];
      r[u] = rv(ii[u]);
      a[u] = anb[pj[u]];
    }
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const V mat1 = d[u] * r[u];
      const V t = v - a[u] * mat1;
      v = u < cnt ? t : v;
    }
  }
  return v;
}

// d - sum over the lower-neighbour list [e0, e1), skipping neighbour
// skip_row, of (anb[lpji[e]] * dv(i)) * anb[lpij[e]], i = lnb[e] (the DILU
// diagonal of cfdFactorizeILU.m)
template <typename DV>
__device__ __forceinline__ double dilu_factor_sum(const int32_t *__restrict__ lnb,
                                                  const int32_t *__restrict__ lpji,
                                                  const int32_t *__restrict__ lpij,
                                                  const double *__restrict__ anb, const DV &dv,
                                                  int32_t skip_row, int32_t e0, int32_t e1,
                                                  double d) {
  constexpr int N = kDiluChunk;
  for (; e0 < e1; e0 += N) {
    const int32_t cnt = e1 - e0 < N ? e1 - e0 : N;
    int32_t ii[N], pj[N], pi[N];
    double dj[N], aji[N], aij[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const int32_t e = e0 + (u < cnt ? u : cnt - 1);
      ii[u] = lnb[e];
      pj[u] = lpji[e];
      pi[u] = lpij[e];
    }
#pragma unroll
    for (int u = 0; u < N; ++u) {
      dj[u] = dv(ii[u]);
      aji[u] = anb[pj[u]];
      aij[u] = anb[pi[u]];
    }
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const double t = d - (aji[u] * dj[u]) * aij[u];
      d = (u < cnt && ii[u] != skip_row) ? t : d;
    }
  }
  return d;
}

// v - sum over the CSR entries [k0, k1) with column j > i of anb[k] * rv(j)
// (the backward substitution)
template <typename V, typename RV>
__device__ __forceinline__ V dilu_upper_sum(const int32_t *__restrict__ col,
                                            const V *__restrict__ anb, const RV &rv, int32_t i,
                                            int32_t k0, int32_t k1, V v) {
  constexpr int N = kDiluChunk;
  for (; k0 < k1; k0 += N) {
    const int32_t cnt = k1 - k0 < N ? k1 - k0 : N;
    int32_t c[N];
    V a[N], r[N];
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const int32_t e = k0 + (u < cnt ? u : cnt - 1);
      c[u] = col[e];
      a[u] = anb[e];
    }
#pragma unroll
    for (int u = 0; u < N; ++u) r[u] = rv(c[u]);
#pragma unroll
    for (int u = 0; u < N; ++u) {
      const V t = v - a[u] * r[u];
      v = (u < cnt && c[u] > i) ? t : v;
    }
  }
  return v;
}

__global__ void __launch_bounds__(kBlock)
k_residual_ref(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
               const double *__restrict__ ac, const double *__restrict__ anb,
               const double *__restrict__ bc, const double *__restrict__ x,
               double *__restrict__ r, double *__restrict__ partial,
               const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  __shared__ double lds[kBlock / kWave];
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double a = 0.0;
  if (i < n) {
    const double s = csr_row_sum(col, anb, [x](int32_t j) { return x[j]; }, ptr[i], ptr[i + 1],
                                 bc[i] - ac[i] * x[i]);
    r[i] = s;
    a = fabs(s);
  }
  a = block_sum(a, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
}

// ---- the reference level's rows as colour-class sliced-ELL chunks (slot
// row t of chunk t / 64 is row row[t], -1 for padding rows; entries in cconn
// order, padding entries col = -1).  Each row's arithmetic is that of
// k_residual_ref / k_gs_rows.
// v - sum over the N slots from e0 of val x[col], padding (col -1) skipped;
// the run's loads are issued together, padding reading x[0]
template <int N>
__device__ __forceinline__ double csell_run(const int32_t *__restrict__ col,
                                            const double *__restrict__ val, const double *x,
                                            int32_t e0, double v) {
  int32_t j[N];
  double a[N], xj[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = col[e0 + u * 64];
    a[u] = val[e0 + u * 64];
  }
#pragma unroll
  for (int u = 0; u < N; ++u) xj[u] = x[j[u] >= 0 ? j[u] : 0];
#pragma unroll
  for (int u = 0; u < N; ++u) asm volatile("" : "+v"(xj[u]));
#pragma unroll
  for (int u = 0; u < N; ++u) v = j[u] >= 0 ? v - a[u] * xj[u] : v;
  return v;
}

__device__ __forceinline__ double csell_row(const int32_t *__restrict__ off,
                                            const int32_t *__restrict__ col,
                                            const double *__restrict__ val, const double *x,
                                            int32_t t, double v) {
  const int32_t q = t >> 6;
  const int32_t base = off[q] + (t & 63), w = (off[q + 1] - off[q]) >> 6;
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun) v = csell_run<kSellRun>(col, val, x, base + k * 64, v);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: return csell_run<1>(col, val, x, e0, v);
    case 2: return csell_run<2>(col, val, x, e0, v);
    case 3: return csell_run<3>(col, val, x, e0, v);
    case 4: return csell_run<4>(col, val, x, e0, v);
    case 5: return csell_run<5>(col, val, x, e0, v);
    case 6: return csell_run<6>(col, val, x, e0, v);
    case 7: return csell_run<7>(col, val, x, e0, v);
    default: return v;
  }
}

__global__ void __launch_bounds__(kBlock)
k_residual_csell(int64_t nslot, const int32_t *__restrict__ row, const int32_t *__restrict__ off,
                 const int32_t *__restrict__ col, const double *__restrict__ val,
                 const double *__restrict__ ac, const double *__restrict__ bc,
                 const double *__restrict__ x, double *__restrict__ r,
                 double *__restrict__ partial, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  __shared__ double lds[kBlock / kWave];
  const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double a = 0.0;
  const int32_t i = t < nslot ? row[t] : -1;
  if (i >= 0) {
    const double s = csell_row(off, col, val, x, (int32_t)t, bc[i] - ac[i] * x[i]);
    r[i] = s;
    a = fabs(s);
  }
  a = block_sum(a, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = a;
}

void residual_csell(hipStream_t s, int64_t nslot, const int32_t *row, const int32_t *off,
                    const int32_t *col, const double *val, const double *ac, const double *bc,
                    const double *x, double *r, double *partial, const double *stop) {
  if (nslot <= 0) return;
  hipLaunchKernelGGL(k_residual_csell, dim3(nblocks(nslot)), dim3(kBlock), 0, s, nslot, row, off,
                     col, val, ac, bc, x, r, partial, stop);
}

// Gauss-Seidel rows of the chunks [q0, q1) (one colour class)
__global__ void __launch_bounds__(kBlock)
k_gs_csell(int32_t q0, int32_t q1, const int32_t *__restrict__ row,
           const int32_t *__restrict__ off, const int32_t *__restrict__ col,
           const double *__restrict__ val, const double *__restrict__ ac,
           const double *__restrict__ b, double *x, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  const int32_t t = q0 * 64 + blockIdx.x * kBlock + threadIdx.x;
  if (t >= q1 * 64) return;
  const int32_t i = row[t];
  if (i < 0) return;
  x[i] = csell_row(off, col, val, x, t, b[i]) / ac[i];
}

void gs_csell(hipStream_t s, int32_t q0, int32_t q1, const int32_t *row, const int32_t *off,
              const int32_t *col, const double *val, const double *ac, const double *b,
              double *x, const double *stop) {
  if (q1 <= q0) return;
  hipLaunchKernelGGL(k_gs_csell, dim3(nblocks((int64_t)(q1 - q0) * 64)), dim3(kBlock), 0, s, q0,
                     q1, row, off, col, val, ac, b, x, stop);
}

void residual_ref(hipStream_t s, int32_t n, const int32_t *ptr, const int32_t *col,
                  const double *ac, const double *anb, const double *bc, const double *x,
                  double *r, double *partial, const double *stop) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_residual_ref, dim3(nblocks(n)), dim3(kBlock), 0, s, n, ptr, col, ac, anb,
                     bc, x, r, partial, stop);
}

// Device-side stopping test of a smoothing call (mean |r| after each sweep,
// cfdSolveAlgebraicSystem.m:25-27): sums the residual kernel's per-block |r|
// partials as reduce_final does, then
//   phase 0 (initial): st[0] = st[1] = m, st[2] = !(m > rrf m), st[3] |= isnan(m)
//   phase 1 (after a sweep; nothing once st[2] is set): st[1] = m,
//           st[2] = !(m > rrf st[0])
// with m = sum / n.  st[2] is the stop flag the sweep kernels test.
__global__ void __launch_bounds__(kBlock)
k_alg_test(const double *__restrict__ partial, int32_t np, int32_t n, double rrf, int phase,
           double *st) {
  __shared__ double lds[kBlock / kWave];
  if (phase == 1 && st[2] != 0.0) return;
  double v = 0.0;
  for (int32_t b = threadIdx.x; b < np; b += kBlock) v += partial[b];
  v = block_sum(v, lds);
  if (threadIdx.x == 0) {
    const double m = v / n;
    st[1] = m;
    if (phase == 0) {
      st[0] = m;
      st[2] = (m > rrf * m) ? 0.0 : 1.0;
      if (isnan(m)) st[3] = 1.0;
    } else if (!(m > rrf * st[0])) {
      st[2] = 1.0;
    }
  }
}

void alg_test(hipStream_t s, const double *partial, int32_t np, int32_t n, double rrf, int phase,
              double *st) {
  hipLaunchKernelGGL(k_alg_test, dim3(1), dim3(kBlock), 0, s, partial, np, n, rrf, phase, st);
}

__global__ void __launch_bounds__(kBlock)
k_reduce_groups(const double *__restrict__ partial, int32_t nblk, int32_t width,
                double *__restrict__ out) {
  __shared__ double lds[kBlock / kWave];
  // one block per output component; each thread sums a strided subset
  const int c = blockIdx.x;
  double v = 0.0;
  for (int32_t b = threadIdx.x; b < nblk; b += kBlock) v += partial[(int64_t)b * width + c];
  v = block_sum(v, lds);
  if (threadIdx.x == 0) out[c] = v;
}

void reduce_final(hipStream_t s, const double *partial, int32_t np, double *out) {
  hipLaunchKernelGGL(k_reduce_groups, dim3(1), dim3(kBlock), 0, s, partial, np, 1, out);
}

void reduce_groups(hipStream_t s, const double *partial, int32_t nblk, int32_t width,
                   double *out) {
  hipLaunchKernelGGL(k_reduce_groups, dim3(width), dim3(kBlock), 0, s, partial, nblk, width,
                     out);
}

__device__ __forceinline__ void gs_row(int32_t i, const int32_t *__restrict__ ptr,
                                       const int32_t *__restrict__ col,
                                       const double *__restrict__ ac,
                                       const double *__restrict__ anb, const double *b,
                                       double *x) {
  const double v = csr_row_sum(col, anb, [x](int32_t j) { return x[j]; }, ptr[i], ptr[i + 1], b[i]);
  x[i] = v / ac[i];
}

__global__ void __launch_bounds__(kBlock)
k_gs_rows(const int32_t *__restrict__ rows, int32_t count, const int32_t *__restrict__ ptr,
          const int32_t *__restrict__ col, const double *__restrict__ ac,
          const double *__restrict__ anb, const double *__restrict__ b, double *x,
          const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  const int32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  gs_row(rows[t], ptr, col, ac, anb, b, x);
}

// A whole level-scheduled sweep in one workgroup: the rows of each level in
// parallel, a workgroup barrier between levels (levels lp[q] .. lp[q+1] of
// rows).  For schedules of many small levels this replaces a launch per
// level; every row is computed exactly as by the per-level kernel.
__global__ void __launch_bounds__(kSweepBlock)
k_gs_sweep_block(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
                 const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                 const double *__restrict__ ac, const double *__restrict__ anb, const double *b,
                 double *x, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  for (int32_t q = 0; q < nlev; ++q) {
    for (int32_t t = lp[q] + threadIdx.x; t < lp[q + 1]; t += kSweepBlock)
      gs_row(rows[t], ptr, col, ac, anb, b, x);
    __syncthreads();
  }
}

// The forward then backward sweep (SOR) of a level whose n unknowns fit in
// LDS: x is loaded once, every level reads and writes it there (the level
// chain runs at LDS latency), and it is stored back at the end.  Row
// arithmetic as gs_row.
__global__ void __launch_bounds__(kSweepBlock)
k_sor_lds(const int32_t *__restrict__ fr, const int32_t *__restrict__ flp, int32_t nf,
          const int32_t *__restrict__ br, const int32_t *__restrict__ blp, int32_t nbk, int32_t n,
          const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
          const double *__restrict__ ac, const double *__restrict__ anb,
          const double *__restrict__ b, double *__restrict__ x, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  extern __shared__ double xs[];
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) xs[i] = x[i];
  __syncthreads();
  auto sweep = [&](const int32_t *rows, const int32_t *lp, int32_t nlev) {
    for (int32_t q = 0; q < nlev; ++q) {
      for (int32_t t = lp[q] + threadIdx.x; t < lp[q + 1]; t += kSweepBlock) {
        const int32_t i = rows[t];
        const double v =
            csr_row_sum(col, anb, [](int32_t j) { return xs[j]; }, ptr[i], ptr[i + 1], b[i]);
        xs[i] = v / ac[i];
      }
      __syncthreads();
    }
  };
  sweep(fr, flp, nf);
  sweep(br, blp, nbk);
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) x[i] = xs[i];
}

void sor_lds(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf, const int32_t *br,
             const int32_t *blp, int32_t nbk, int32_t n, const int32_t *ptr, const int32_t *col,
             const double *ac, const double *anb, const double *b, double *x, const double *stop) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_sor_lds, dim3(1), dim3(kSweepBlock), (size_t)n * sizeof(double), s, fr, flp,
                     nf, br, blp, nbk, n, ptr, col, ac, anb, b, x, stop);
}

// The forward then backward sweep (SOR) of a small level held whole in LDS:
// the workgroup copies the unknowns, right-hand side, diagonal, the rows as
// ELL (W slots, slot u of row i at u * n + i, 16-bit columns, lengths) and
// both schedules into LDS, then one wave runs every level.  No level needs a
// workgroup barrier (a wave's LDS accesses complete in order) and no global
// load is left on the level chain.  Row arithmetic as gs_row.
// doubles of LDS a staged level takes (sor_staged_lds bytes, rounded up)
__host__ __device__ inline size_t sor_staged_doubles(int32_t n, int32_t w, int32_t nf,
                                                     int32_t nbk) {
  const int32_t ws = w <= 4 ? 4 : w <= 8 ? 8 : 16;
  const size_t bytes = (size_t)n * (3 * sizeof(double) + ws * (sizeof(double) + sizeof(uint16_t)) +
                                    3 * sizeof(int32_t)) +
                       (size_t)(nf + nbk + 2) * sizeof(int32_t);
  return (bytes + sizeof(double) - 1) / sizeof(double);
}

template <int W>
struct StagedLevel {
  double *xs, *bs, *as, *an;
  int32_t *ln, *frs, *brs, *flps, *blps;
  uint16_t *cs;
  int32_t n, nf, nbk;
  // carve the LDS (sor_staged_lds bytes) and copy the level in (every thread)
  __device__ void stage(double *lds, const int32_t *__restrict__ fr,
                        const int32_t *__restrict__ flp, int32_t nf_,
                        const int32_t *__restrict__ br, const int32_t *__restrict__ blp,
                        int32_t nbk_, int32_t n_, const int32_t *__restrict__ ptr,
                        const int32_t *__restrict__ col, const double *__restrict__ ac,
                        const double *__restrict__ anb, const double *__restrict__ b,
                        const double *__restrict__ x) {
    n = n_;
    nf = nf_;
    nbk = nbk_;
    xs = lds;
    bs = xs + n;
    as = bs + n;
    an = as + n;
    ln = reinterpret_cast<int32_t *>(an + (size_t)W * n);
    frs = ln + n;
    brs = frs + n;
    flps = brs + n;
    blps = flps + nf + 1;
    cs = reinterpret_cast<uint16_t *>(blps + nbk + 1);
    for (int32_t i = threadIdx.x; i < n; i += blockDim.x) {
      xs[i] = x[i];
      bs[i] = b[i];
      as[i] = ac[i];
      const int32_t k0 = ptr[i], len = ptr[i + 1] - k0;
      ln[i] = len;
#pragma unroll
      for (int u = 0; u < W; ++u) {
        const int32_t k = k0 + (u < len ? u : 0);
        cs[u * n + i] = (uint16_t)(u < len ? col[k] : i);
        an[u * n + i] = u < len ? anb[k] : 0.0;
      }
      frs[i] = fr[i];
      brs[i] = br[i];
    }
    for (int32_t q = threadIdx.x; q <= nf; q += blockDim.x) flps[q] = flp[q];
    for (int32_t q = threadIdx.x; q <= nbk; q += blockDim.x) blps[q] = blp[q];
  }
  // forward then backward sweep by the calling wave (lane 0..63)
  __device__ void sweep(int32_t lane) {
    auto one = [&](const int32_t *rows, const int32_t *lps, int32_t nlev) {
      for (int32_t q = 0; q < nlev; ++q) {
        for (int32_t t = lps[q] + lane; t < lps[q + 1]; t += 64) {
          const int32_t i = rows[t], len = ln[i];
          double v = bs[i];
#pragma unroll
          for (int u = 0; u < W; ++u) {
            const double tv = v - an[u * n + i] * xs[cs[u * n + i]];
            v = u < len ? tv : v;
          }
          xs[i] = v / as[i];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      }
    };
    one(frs, flps, nf);
    one(brs, blps, nbk);
  }
  // row i's residual, k_residual_ref's arithmetic
  __device__ double residual(int32_t i) const {
    double v = bs[i] - as[i] * xs[i];
    const int32_t len = ln[i];
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const double tv = v - an[u * n + i] * xs[cs[u * n + i]];
      v = u < len ? tv : v;
    }
    return v;
  }
};

template <int W>
__global__ void __launch_bounds__(kSweepBlock)
k_sor_staged(const int32_t *__restrict__ fr, const int32_t *__restrict__ flp, int32_t nf,
             const int32_t *__restrict__ br, const int32_t *__restrict__ blp, int32_t nbk,
             int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
             const double *__restrict__ ac, const double *__restrict__ anb,
             const double *__restrict__ b, double *__restrict__ x,
             const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  extern __shared__ double st_lds[];
  StagedLevel<W> L;
  L.stage(st_lds, fr, flp, nf, br, blp, nbk, n, ptr, col, ac, anb, b, x);
  __syncthreads();
  if (threadIdx.x >= 64) return;
  L.sweep(threadIdx.x);
  for (int32_t i = threadIdx.x; i < n; i += 64) x[i] = L.xs[i];
}

// A whole smoothing call with the device-side stopping test (the initial
// residual and test, then up to n_iter sweeps, each followed by its residual
// and test) on a staged level, in one launch.  The mean |r| is summed as
// k_residual_ref's blocks and k_alg_test sum it (per kBlock rows: wave sums
// added in wave order; then the block sums, strided over kBlock threads, as
// wave sums added in order), so st[] gets the values those kernels give.
// r holds the last residual computed; st as k_alg_test leaves it.
template <int W>
__global__ void __launch_bounds__(kSweepBlock)
k_sor_smooth_staged(const int32_t *__restrict__ fr, const int32_t *__restrict__ flp, int32_t nf,
                    const int32_t *__restrict__ br, const int32_t *__restrict__ blp, int32_t nbk,
                    int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                    const double *__restrict__ ac, const double *__restrict__ anb,
                    const double *__restrict__ b, double *__restrict__ x, double *__restrict__ r,
                    int32_t n_iter, double rrf, double *st) {
  extern __shared__ double st_lds[];
  StagedLevel<W> L;
  L.stage(st_lds, fr, flp, nf, br, blp, nbk, n, ptr, col, ac, anb, b, x);
  const int32_t np = (n + kBlock - 1) / kBlock;
  double *wsum = st_lds + sor_staged_doubles(n, W, nf, nbk);  // 4 per block of kBlock rows
  double *wv = wsum + 4 * np;                                  // kBlock / kWave
  __syncthreads();
  const int32_t lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  constexpr int kWpb = kBlock / 64;          // waves per residual block
  constexpr int kVb = kSweepBlock / kBlock;  // residual blocks per pass
  auto mean_abs_residual = [&]() -> double {
    for (int32_t p0 = 0; p0 < np; p0 += kVb) {
      const int32_t vb = p0 + wid / kWpb, w = wid % kWpb, i = vb * kBlock + w * 64 + lane;
      double a = 0.0;
      if (vb < np && i < n) {
        const double v = L.residual(i);
        r[i] = v;
        a = fabs(v);
      }
      a = wave_sum(a);
      if (lane == 0 && vb < np) wsum[vb * kWpb + w] = a;
    }
    __syncthreads();
    if (threadIdx.x < kBlock) {
      double v = 0.0;
      for (int32_t blk = threadIdx.x; blk < np; blk += kBlock) {
        double p = 0;
        for (int w = 0; w < kWpb; ++w) p += wsum[blk * kWpb + w];
        v += p;
      }
      v = wave_sum(v);
      if (lane == 0) wv[wid] = v;
    }
    __syncthreads();
    double t = 0;
    for (int w = 0; w < kBlock / 64; ++w) t += wv[w];
    __syncthreads();
    return t / n;
  };
  double m = mean_abs_residual();
  const double m0 = m;
  bool stopped = !(m > rrf * m);
  const bool nan = isnan(m);
  for (int32_t it = 1; it <= n_iter && !stopped; ++it) {
    if (threadIdx.x < 64) L.sweep(lane);
    __syncthreads();
    m = mean_abs_residual();
    if (!(m > rrf * m0)) stopped = true;
  }
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) x[i] = L.xs[i];
  if (threadIdx.x == 0) {
    st[0] = m0;
    st[1] = m;
    st[2] = stopped ? 1.0 : 0.0;
    if (nan) st[3] = 1.0;
  }
}

size_t sor_smooth_staged_lds(int32_t n, int32_t w, int32_t nf, int32_t nbk) {
  const int32_t np = (n + kBlock - 1) / kBlock;
  return (sor_staged_doubles(n, w, nf, nbk) + 4 * (size_t)np + kBlock / 64) * sizeof(double);
}

bool sor_smooth_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                       const int32_t *br, const int32_t *blp, int32_t nbk, int32_t n, int32_t w,
                       const int32_t *ptr, const int32_t *col, const double *ac,
                       const double *anb, const double *b, double *x, double *r, int32_t n_iter,
                       double rrf, double *st) {
  if (n <= 0) return false;
  const size_t lds = sor_smooth_staged_lds(n, w, nf, nbk);
#define UFVM_SMOOTH(W)                                                                      \
  hipLaunchKernelGGL(k_sor_smooth_staged<W>, dim3(1), dim3(kSweepBlock), lds, s, fr, flp, nf, \
                     br, blp, nbk, n, ptr, col, ac, anb, b, x, r, n_iter, rrf, st)
  if (w <= 4)
    UFVM_SMOOTH(4);
  else if (w <= 8)
    UFVM_SMOOTH(8);
  else
    UFVM_SMOOTH(16);
#undef UFVM_SMOOTH
  return hipGetLastError() == hipSuccess;
}

size_t sor_staged_lds(int32_t n, int32_t w, int32_t nf, int32_t nbk) {
  const int32_t ws = w <= 4 ? 4 : w <= 8 ? 8 : 16;
  return (size_t)n * (3 * sizeof(double) + ws * (sizeof(double) + sizeof(uint16_t)) +
                      3 * sizeof(int32_t)) +
         (size_t)(nf + nbk + 2) * sizeof(int32_t);
}

bool sor_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                const int32_t *br, const int32_t *blp, int32_t nbk, int32_t n, int32_t w,
                const int32_t *ptr, const int32_t *col, const double *ac, const double *anb,
                const double *b, double *x, const double *stop) {
  if (n <= 0) return true;
  const size_t lds = sor_staged_lds(n, w, nf, nbk);
#define UFVM_STAGED(W)                                                                          \
  hipLaunchKernelGGL(k_sor_staged<W>, dim3(1), dim3(kSweepBlock), lds, s, fr, flp, nf, br, blp, \
                     nbk, n, ptr, col, ac, anb, b, x, stop)
  if (w <= 4)
    UFVM_STAGED(4);
  else if (w <= 8)
    UFVM_STAGED(8);
  else
    UFVM_STAGED(16);
#undef UFVM_STAGED
  return hipGetLastError() == hipSuccess;
}

// The DILU forward and backward substitutions of a small level held whole in
// LDS and run by one wave, as k_sor_staged: rc, dinv, the lower lists (W
// slots: 16-bit neighbour, a_ji) and the entries above the diagonal (16-bit
// column, a_ij, cconn order) are copied in, the substitutions run with no
// workgroup barrier, and rc = z, x += z are written at the end.  Row
// arithmetic as dilu_fwd_row / dilu_bwd_row.
__host__ __device__ inline int dilu_staged_slots(int32_t w) {
  return w <= 2 ? 2 : w <= 4 ? 4 : w <= 8 ? 8 : 16;
}

size_t dilu_staged_lds(int32_t n, int32_t w, int32_t nf, int32_t nbk) {
  const int ws = dilu_staged_slots(w);
  return (size_t)n * (2 * sizeof(double) + 2 * ws * (sizeof(double) + sizeof(uint16_t)) +
                      4 * sizeof(int32_t)) +
         (size_t)(nf + nbk + 2) * sizeof(int32_t);
}

template <int W>
__global__ void __launch_bounds__(kSweepBlock)
k_dilu_staged(const int32_t *__restrict__ fr, const int32_t *__restrict__ flp, int32_t nf,
              const int32_t *__restrict__ br, const int32_t *__restrict__ blp, int32_t nbk,
              int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
              const double *__restrict__ anb, const int32_t *__restrict__ lptr,
              const int32_t *__restrict__ lnb, const int32_t *__restrict__ lpji,
              const double *__restrict__ dinv, double *__restrict__ rc, double *__restrict__ x,
              const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  extern __shared__ double dl_lds[];
  double *rs = dl_lds, *dv = rs + n, *la = dv + n, *ua = la + (size_t)W * n;
  int32_t *ll = reinterpret_cast<int32_t *>(ua + (size_t)W * n);
  int32_t *ul = ll + n, *frs = ul + n, *brs = frs + n, *flps = brs + n, *blps = flps + nf + 1;
  uint16_t *lc = reinterpret_cast<uint16_t *>(blps + nbk + 1), *uc = lc + (size_t)W * n;
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) {
    rs[i] = rc[i];
    dv[i] = dinv[i];
    const int32_t e0 = lptr[i], m = lptr[i + 1] - e0;
    ll[i] = m;
#pragma unroll
    for (int u = 0; u < W; ++u) {
      const int32_t e = e0 + (u < m ? u : 0);
      lc[u * n + i] = (uint16_t)(u < m ? lnb[e] : i);
      la[u * n + i] = u < m ? anb[lpji[e]] : 0.0;
    }
    int32_t c = 0;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int32_t j = col[k];
      if (j > i) {
        uc[c * n + i] = (uint16_t)j;
        ua[c * n + i] = anb[k];
        ++c;
      }
    }
    ul[i] = c;
    for (int32_t u = c; u < W; ++u) {
      uc[u * n + i] = (uint16_t)i;
      ua[u * n + i] = 0.0;
    }
    frs[i] = fr[i];
    brs[i] = br[i];
  }
  for (int32_t q = threadIdx.x; q <= nf; q += kSweepBlock) flps[q] = flp[q];
  for (int32_t q = threadIdx.x; q <= nbk; q += kSweepBlock) blps[q] = blp[q];
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int32_t lane = threadIdx.x;
  auto wave_step = [] {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  };
  for (int32_t q = 0; q < nf; ++q) {
    for (int32_t t = flps[q] + lane; t < flps[q + 1]; t += 64) {
      const int32_t j = frs[t], m = ll[j];
      double v = rs[j];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        const int32_t i = lc[u * n + j];
        const double mat1 = dv[i] * rs[i];
        const double tv = v - la[u * n + j] * mat1;
        v = u < m ? tv : v;
      }
      rs[j] = v;
    }
    wave_step();
  }
  for (int32_t q = 0; q < nbk; ++q) {
    for (int32_t t = blps[q] + lane; t < blps[q + 1]; t += 64) {
      const int32_t i = brs[t], m = ul[i];
      double v = rs[i];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        const double tv = v - ua[u * n + i] * rs[uc[u * n + i]];
        v = u < m ? tv : v;
      }
      rs[i] = dv[i] * v;
    }
    wave_step();
  }
  for (int32_t i = lane; i < n; i += 64) {
    const double z = rs[i];
    rc[i] = z;
    x[i] = x[i] + z;
  }
}

// The DILU factor of a small level held whole in LDS and run by one wave:
// the diagonal and each row's lower list without the skipped neighbour (W
// slots: 16-bit neighbour, a_ji, a_ij) are copied in, dinv is built level by
// level in LDS and written at the end.  Row arithmetic as dilu_factor_row.
size_t dilu_factor_staged_lds(int32_t n, int32_t w, int32_t nf) {
  const int ws = dilu_staged_slots(w);
  return (size_t)n * (2 * sizeof(double) + ws * (2 * sizeof(double) + sizeof(uint16_t)) +
                      2 * sizeof(int32_t)) +
         (size_t)(nf + 1) * sizeof(int32_t);
}

template <int W>
__global__ void __launch_bounds__(kSweepBlock)
k_dilu_factor_staged(const int32_t *__restrict__ fr, const int32_t *__restrict__ flp, int32_t nf,
                     int32_t n, const double *__restrict__ ac, const double *__restrict__ anb,
                     const int32_t *__restrict__ lptr, const int32_t *__restrict__ lnb,
                     const int32_t *__restrict__ lpji, const int32_t *__restrict__ lpij,
                     int32_t skip_row, double *__restrict__ dinv) {
  extern __shared__ double df_lds[];
  double *dv = df_lds, *d0 = dv + n, *aji = d0 + n, *aij = aji + (size_t)W * n;
  int32_t *ll = reinterpret_cast<int32_t *>(aij + (size_t)W * n);
  int32_t *frs = ll + n, *flps = frs + n;
  uint16_t *lc = reinterpret_cast<uint16_t *>(flps + nf + 1);
  for (int32_t j = threadIdx.x; j < n; j += kSweepBlock) {
    d0[j] = ac[j];
    int32_t c = 0;
    for (int32_t e = lptr[j]; e < lptr[j + 1]; ++e) {
      const int32_t i = lnb[e];
      if (i != skip_row) {
        lc[c * n + j] = (uint16_t)i;
        aji[c * n + j] = anb[lpji[e]];
        aij[c * n + j] = anb[lpij[e]];
        ++c;
      }
    }
    ll[j] = c;
    for (int32_t u = c; u < W; ++u) {
      lc[u * n + j] = (uint16_t)j;
      aji[u * n + j] = 0.0;
      aij[u * n + j] = 0.0;
    }
    frs[j] = fr[j];
  }
  for (int32_t q = threadIdx.x; q <= nf; q += kSweepBlock) flps[q] = flp[q];
  __syncthreads();
  if (threadIdx.x >= 64) return;
  const int32_t lane = threadIdx.x;
  for (int32_t q = 0; q < nf; ++q) {
    for (int32_t t = flps[q] + lane; t < flps[q + 1]; t += 64) {
      const int32_t j = frs[t], m = ll[j];
      double d = d0[j];
#pragma unroll
      for (int u = 0; u < W; ++u) {
        const double tv = d - (aji[u * n + j] * dv[lc[u * n + j]]) * aij[u * n + j];
        d = u < m ? tv : d;
      }
      dv[j] = 1.0 / d;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  for (int32_t j = lane; j < n; j += 64) dinv[j] = dv[j];
}

bool dilu_factor_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                        int32_t n, int32_t w, const double *ac, const double *anb,
                        const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
                        const int32_t *lpij, int32_t skip_row, double *dinv) {
  if (n <= 0) return true;
  const size_t lds = dilu_factor_staged_lds(n, w, nf);
#define UFVM_FSTAGED(W)                                                                           \
  hipLaunchKernelGGL(k_dilu_factor_staged<W>, dim3(1), dim3(kSweepBlock), lds, s, fr, flp, nf, n, \
                     ac, anb, lptr, lnb, lpji, lpij, skip_row, dinv)
  switch (dilu_staged_slots(w)) {
    case 2: UFVM_FSTAGED(2); break;
    case 4: UFVM_FSTAGED(4); break;
    case 8: UFVM_FSTAGED(8); break;
    default: UFVM_FSTAGED(16); break;
  }
#undef UFVM_FSTAGED
  return hipGetLastError() == hipSuccess;
}

bool dilu_staged(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf,
                 const int32_t *br, const int32_t *blp, int32_t nbk, int32_t n, int32_t w,
                 const int32_t *ptr, const int32_t *col, const double *anb, const int32_t *lptr,
                 const int32_t *lnb, const int32_t *lpji, const double *dinv, double *rc,
                 double *x, const double *stop) {
  if (n <= 0) return true;
  const size_t lds = dilu_staged_lds(n, w, nf, nbk);
#define UFVM_DSTAGED(W)                                                                          \
  hipLaunchKernelGGL(k_dilu_staged<W>, dim3(1), dim3(kSweepBlock), lds, s, fr, flp, nf, br, blp, \
                     nbk, n, ptr, col, anb, lptr, lnb, lpji, dinv, rc, x, stop)
  switch (dilu_staged_slots(w)) {
    case 2: UFVM_DSTAGED(2); break;
    case 4: UFVM_DSTAGED(4); break;
    case 8: UFVM_DSTAGED(8); break;
    default: UFVM_DSTAGED(16); break;
  }
#undef UFVM_DSTAGED
  return hipGetLastError() == hipSuccess;
}

// DILU forward and backward substitution (cfdSolveILU's, as dilu_fwd_row /
// dilu_bwd_row) of a level whose rc fits in LDS; x += z in memory
__global__ void __launch_bounds__(kSweepBlock)
k_dilu_lds(const int32_t *__restrict__ fr, const int32_t *__restrict__ flp, int32_t nf,
           const int32_t *__restrict__ br, const int32_t *__restrict__ blp, int32_t nbk, int32_t n,
           const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
           const double *__restrict__ anb, const int32_t *__restrict__ lptr,
           const int32_t *__restrict__ lnb, const int32_t *__restrict__ lpji,
           const double *__restrict__ dinv, double *__restrict__ rc, double *__restrict__ x,
           const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  extern __shared__ double rs[];
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) rs[i] = rc[i];
  __syncthreads();
  for (int32_t q = 0; q < nf; ++q) {
    for (int32_t t = flp[q] + threadIdx.x; t < flp[q + 1]; t += kSweepBlock) {
      const int32_t j = fr[t];
      rs[j] = dilu_lower_sum(lnb, lpji, anb, dinv, [](int32_t i) { return rs[i]; }, lptr[j],
                             lptr[j + 1], rs[j]);
    }
    __syncthreads();
  }
  for (int32_t q = 0; q < nbk; ++q) {
    for (int32_t t = blp[q] + threadIdx.x; t < blp[q + 1]; t += kSweepBlock) {
      const int32_t i = br[t];
      double v = rs[i];
      if (i < n - 1)
        v = dilu_upper_sum(col, anb, [](int32_t j) { return rs[j]; }, i, ptr[i], ptr[i + 1], v);
      const double z = dinv[i] * v;
      rs[i] = z;
      x[i] = x[i] + z;
    }
    __syncthreads();
  }
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) rc[i] = rs[i];
}

void dilu_lds(hipStream_t s, const int32_t *fr, const int32_t *flp, int32_t nf, const int32_t *br,
              const int32_t *blp, int32_t nbk, int32_t n, const int32_t *ptr, const int32_t *col,
              const double *anb, const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
              const double *dinv, double *rc, double *x, const double *stop) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dilu_lds, dim3(1), dim3(kSweepBlock), (size_t)n * sizeof(double), s, fr,
                     flp, nf, br, blp, nbk, n, ptr, col, anb, lptr, lnb, lpji, dinv, rc, x, stop);
}

// DILU factor (dilu_factor_row) of a level whose dinv fits in LDS
__global__ void __launch_bounds__(kSweepBlock)
k_dilu_factor_lds(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
                  int32_t n, const double *__restrict__ ac, const double *__restrict__ anb,
                  const int32_t *__restrict__ lptr, const int32_t *__restrict__ lnb,
                  const int32_t *__restrict__ lpji, const int32_t *__restrict__ lpij,
                  int32_t skip_row, double *__restrict__ dinv) {
  extern __shared__ double ds[];
  for (int32_t q = 0; q < nlev; ++q) {
    for (int32_t t = lp[q] + threadIdx.x; t < lp[q + 1]; t += kSweepBlock) {
      const int32_t j = rows[t];
      const double d = dilu_factor_sum(lnb, lpji, lpij, anb, [](int32_t i) { return ds[i]; },
                                       skip_row, lptr[j], lptr[j + 1], ac[j]);
      ds[j] = 1.0 / d;
    }
    __syncthreads();
  }
  for (int32_t i = threadIdx.x; i < n; i += kSweepBlock) dinv[i] = ds[i];
}

void dilu_factor_lds(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                     int32_t n, const double *ac, const double *anb, const int32_t *lptr,
                     const int32_t *lnb, const int32_t *lpji, const int32_t *lpij,
                     int32_t skip_row, double *dinv) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dilu_factor_lds, dim3(1), dim3(kSweepBlock), (size_t)n * sizeof(double), s,
                     rows, lp, nlev, n, ac, anb, lptr, lnb, lpji, lpij, skip_row, dinv);
}

// ---- pipelined one-workgroup sweeps over schedule-ordered ELL copies.  The
// sweep runs in one workgroup (one CU), so its cost per level is set by how
// many cache lines that CU's loads touch and by one round trip: a CSR row's
// entries sit at the row's own offsets, so the lanes of a level (rows of
// arbitrary index) each touch their own lines for every entry's column and
// value.  Here slot u of schedule position s sits at u * n + s: the lanes of
// a level (consecutive positions) read consecutive words, and only the
// neighbour values x[c] and the right-hand side b[i] are gathered.  Positions
// past a row's length hold column 0 and weight 0 (never summed).
// Thread t owns position lps[q] + t of level q; level q+1's slots, length,
// diagonal and right-hand side and level q+2's row are in flight while level
// q computes.  Row arithmetic as gs_row (the same terms in the same order).

template <int NE>
struct PipeRow {
  int32_t i, len;
  double b, d;
  int32_t c[NE];
  double a[NE];
  // position s (< 0: none) holding row i; slots past w read slot w - 1
  __device__ __forceinline__ void load(int32_t s, int32_t row, int32_t n, int32_t w,
                                       const int32_t *__restrict__ ecol,
                                       const int32_t *__restrict__ elen,
                                       const double *__restrict__ eanb,
                                       const double *__restrict__ eac,
                                       const double *__restrict__ bv) {
    i = row;
    if (s >= 0) {
      len = elen[s];
      d = eac[s];
      b = bv[row];
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const size_t e = (size_t)(u < w ? u : w - 1) * n + s;
        c[u] = ecol[e];
        a[u] = eanb[e];
      }
    } else {
      len = 0;
      d = 1.0;
      b = 0.0;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        c[u] = 0;
        a[u] = 0.0;
      }
    }
  }
};

template <int NE>
__global__ void __launch_bounds__(kPipeBlock)
k_gs_sweep_pipe(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
                int32_t n, int32_t w, const int32_t *__restrict__ ecol,
                const int32_t *__restrict__ elen, const double *__restrict__ eanb,
                const double *__restrict__ eac, const double *__restrict__ b, double *x,
                const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  extern __shared__ int32_t lps[];
  for (int32_t q = threadIdx.x; q <= nlev; q += kPipeBlock) lps[q] = lp[q];
  __syncthreads();
  const int32_t tid = threadIdx.x;
  auto pos_at = [&](int32_t q) -> int32_t {
    if (q >= nlev) return -1;
    const int32_t t = lps[q] + tid;
    return t < lps[q + 1] ? t : -1;
  };
  int32_t s0 = pos_at(0), s1 = pos_at(1);
  PipeRow<NE> cur;
  cur.load(s0, s0 >= 0 ? rows[s0] : -1, n, w, ecol, elen, eanb, eac, b);
  int32_t i1 = s1 >= 0 ? rows[s1] : -1;
  for (int32_t q = 0; q < nlev; ++q) {
    // in flight: level q+1's slots, level q+2's row
    const int32_t s2 = pos_at(q + 2);
    const int32_t i2 = s2 >= 0 ? rows[s2] : -1;
    PipeRow<NE> nxt;
    nxt.load(s1, i1, n, w, ecol, elen, eanb, eac, b);
    // level q (threads without a row compute on column 0 and store nothing)
    {
      double xj[NE];
#pragma unroll
      for (int u = 0; u < NE; ++u) xj[u] = x[cur.c[u]];
      double v = cur.b;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const double t = v - cur.a[u] * xj[u];
        v = u < cur.len ? t : v;
      }
      for (int32_t u = NE; u < cur.len; ++u) {
        const size_t e = (size_t)u * n + s0;
        v = v - eanb[e] * x[ecol[e]];
      }
      if (s0 >= 0) x[cur.i] = v / cur.d;
    }
    __syncthreads();
    cur = nxt;
    s0 = s1;
    s1 = s2;
    i1 = i2;
  }
}

// eanb / eac for one schedule: slot u of position s from row rows[s]'s
// entry u (0 past the row's length), the diagonal of that row
__global__ void k_pipe_values(const int32_t *__restrict__ rows, int32_t n, int32_t w,
                              const int32_t *__restrict__ ptr, const double *__restrict__ anb,
                              const double *__restrict__ ac, double *__restrict__ eanb,
                              double *__restrict__ eac) {
  const int32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int32_t i = rows[s], k0 = ptr[i], len = ptr[i + 1] - k0;
  eac[s] = ac[i];
  for (int32_t u = 0; u < w; ++u) eanb[(size_t)u * n + s] = u < len ? anb[k0 + u] : 0.0;
}

void pipe_values(hipStream_t s, const int32_t *rows, int32_t n, int32_t w, const int32_t *ptr,
                 const double *anb, const double *ac, double *eanb, double *eac) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_pipe_values, dim3((n + 255) / 256), dim3(256), 0, s, rows, n, w, ptr, anb,
                     ac, eanb, eac);
}

size_t gs_sweep_pipe_lds(int32_t nlev) { return (size_t)(nlev + 1) * sizeof(int32_t); }

void gs_sweep_pipe(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev, int32_t n,
                   int32_t w, const int32_t *ecol, const int32_t *elen, const double *eanb,
                   const double *eac, const double *b, double *x, const double *stop) {
  if (nlev <= 0) return;
  const size_t lds = gs_sweep_pipe_lds(nlev);
#define UFVM_PIPE(NE)                                                                           \
  hipLaunchKernelGGL(k_gs_sweep_pipe<NE>, dim3(1), dim3(kPipeBlock), lds, s, rows, lp, nlev, n, \
                     w, ecol, elen, eanb, eac, b, x, stop)
  if (w <= 8)
    UFVM_PIPE(8);
  else if (w <= 16)
    UFVM_PIPE(16);
  else if (w <= 24)
    UFVM_PIPE(24);
  else
    UFVM_PIPE(32);
#undef UFVM_PIPE
}

// The DILU substitutions (dilu_fwd_row / dilu_bwd_row arithmetic) in the same
// pipelined form, over ELL copies in each direction's schedule order: the
// forward copy holds a row's lower-neighbour list (columns lnb, values
// anb[lpji]), the backward copy its CSR entries with column above the row
// (cconn order).  esrc names each slot's anb entry (-1 past the row's
// length: column 0, value 0).  The forward sweep stores y_i = dinv_i * rc_i
// beside rc_i, so a lower neighbour's term takes one gather; the product is
// the one the row function forms.
template <int NE, bool BWD>
struct DiluPipeRow {
  int32_t i, len;
  double v0, d, x0;
  int32_t c[NE];
  double a[NE];
  __device__ __forceinline__ void load(int32_t s, int32_t row, int32_t n, int32_t w,
                                       const int32_t *__restrict__ ecol,
                                       const int32_t *__restrict__ elen,
                                       const double *__restrict__ eval,
                                       const double *__restrict__ dinv, const double *rc,
                                       const double *x) {
    i = row;
    if (s >= 0) {
      len = elen[s];
      v0 = rc[row];
      d = dinv[row];
      x0 = BWD ? x[row] : 0.0;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const size_t e = (size_t)(u < w ? u : w - 1) * n + s;
        c[u] = ecol[e];
        a[u] = eval[e];
      }
    } else {
      len = 0;
      v0 = 0.0;
      d = 0.0;
      x0 = 0.0;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        c[u] = 0;
        a[u] = 0.0;
      }
    }
  }
};

// forward (BWD false): rc_j -= a_ji y_i over the lower list, y_j = dinv_j rc_j;
// backward: rc_i -= a_ij rc_j over the upper entries, z = dinv_i rc_i,
// rc_i = z, x_i += z
template <int NE, bool BWD>
__global__ void __launch_bounds__(kPipeBlock)
k_dilu_pipe(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
            int32_t n, int32_t w, const int32_t *__restrict__ ecol,
            const int32_t *__restrict__ elen, const double *__restrict__ eval,
            const double *__restrict__ dinv, double *rc, double *y, double *x,
            const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  extern __shared__ int32_t lps[];
  for (int32_t q = threadIdx.x; q <= nlev; q += kPipeBlock) lps[q] = lp[q];
  __syncthreads();
  const int32_t tid = threadIdx.x;
  auto pos_at = [&](int32_t q) -> int32_t {
    if (q >= nlev) return -1;
    const int32_t t = lps[q] + tid;
    return t < lps[q + 1] ? t : -1;
  };
  double *g = BWD ? rc : y;  // the gathered neighbour values
  int32_t s0 = pos_at(0), s1 = pos_at(1);
  DiluPipeRow<NE, BWD> cur;
  cur.load(s0, s0 >= 0 ? rows[s0] : -1, n, w, ecol, elen, eval, dinv, rc, x);
  int32_t i1 = s1 >= 0 ? rows[s1] : -1;
  for (int32_t q = 0; q < nlev; ++q) {
    const int32_t s2 = pos_at(q + 2);
    const int32_t i2 = s2 >= 0 ? rows[s2] : -1;
    DiluPipeRow<NE, BWD> nxt;
    nxt.load(s1, i1, n, w, ecol, elen, eval, dinv, rc, x);
    {
      double gj[NE];
#pragma unroll
      for (int u = 0; u < NE; ++u) gj[u] = g[cur.c[u]];
      double v = cur.v0;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const double t = v - cur.a[u] * gj[u];
        v = u < cur.len ? t : v;
      }
      for (int32_t u = NE; u < cur.len; ++u) {
        const size_t e = (size_t)u * n + s0;
        v = v - eval[e] * g[ecol[e]];
      }
      if (s0 >= 0) {
        if (BWD) {
          const double z = cur.d * v;
          rc[cur.i] = z;
          x[cur.i] = cur.x0 + z;
        } else {
          rc[cur.i] = v;
          y[cur.i] = cur.d * v;
        }
      }
    }
    __syncthreads();
    cur = nxt;
    s0 = s1;
    s1 = s2;
    i1 = i2;
  }
}

// The DILU factor (dilu_factor_row arithmetic) pipelined the same way over an
// ELL copy of the lower-neighbour lists in forward-schedule order, without
// the skipped neighbour: slot values a_ji (eji) and a_ij (eij), the
// neighbour's dinv gathered.
template <int NE>
struct FactorPipeRow {
  int32_t i, len;
  double d0;
  int32_t c[NE];
  double aji[NE], aij[NE];
  __device__ __forceinline__ void load(int32_t s, int32_t row, int32_t n, int32_t w,
                                       const int32_t *__restrict__ ecol,
                                       const int32_t *__restrict__ elen,
                                       const double *__restrict__ eji,
                                       const double *__restrict__ eij,
                                       const double *__restrict__ ac) {
    i = row;
    if (s >= 0) {
      len = elen[s];
      d0 = ac[row];
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const size_t e = (size_t)(u < w ? u : w - 1) * n + s;
        c[u] = ecol[e];
        aji[u] = eji[e];
        aij[u] = eij[e];
      }
    } else {
      len = 0;
      d0 = 1.0;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        c[u] = 0;
        aji[u] = 0.0;
        aij[u] = 0.0;
      }
    }
  }
};

template <int NE>
__global__ void __launch_bounds__(kPipeBlock)
k_dilu_factor_pipe(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
                   int32_t n, int32_t w, const int32_t *__restrict__ ecol,
                   const int32_t *__restrict__ elen, const double *__restrict__ eji,
                   const double *__restrict__ eij, const double *__restrict__ ac, double *dinv) {
  extern __shared__ int32_t lps[];
  for (int32_t q = threadIdx.x; q <= nlev; q += kPipeBlock) lps[q] = lp[q];
  __syncthreads();
  const int32_t tid = threadIdx.x;
  auto pos_at = [&](int32_t q) -> int32_t {
    if (q >= nlev) return -1;
    const int32_t t = lps[q] + tid;
    return t < lps[q + 1] ? t : -1;
  };
  int32_t s0 = pos_at(0), s1 = pos_at(1);
  FactorPipeRow<NE> cur;
  cur.load(s0, s0 >= 0 ? rows[s0] : -1, n, w, ecol, elen, eji, eij, ac);
  int32_t i1 = s1 >= 0 ? rows[s1] : -1;
  for (int32_t q = 0; q < nlev; ++q) {
    const int32_t s2 = pos_at(q + 2);
    const int32_t i2 = s2 >= 0 ? rows[s2] : -1;
    FactorPipeRow<NE> nxt;
    nxt.load(s1, i1, n, w, ecol, elen, eji, eij, ac);
    {
      double dj[NE];
#pragma unroll
      for (int u = 0; u < NE; ++u) dj[u] = dinv[cur.c[u]];
      double d = cur.d0;
#pragma unroll
      for (int u = 0; u < NE; ++u) {
        const double t = d - (cur.aji[u] * dj[u]) * cur.aij[u];
        d = u < cur.len ? t : d;
      }
      for (int32_t u = NE; u < cur.len; ++u) {
        const size_t e = (size_t)u * n + s0;
        d = d - (eji[e] * dinv[ecol[e]]) * eij[e];
      }
      if (s0 >= 0) dinv[cur.i] = 1.0 / d;
    }
    __syncthreads();
    cur = nxt;
    s0 = s1;
    s1 = s2;
    i1 = i2;
  }
}

void dilu_factor_pipe(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                      int32_t n, int32_t w, const int32_t *ecol, const int32_t *elen,
                      const double *eji, const double *eij, const double *ac, double *dinv) {
  if (nlev <= 0) return;
  const size_t lds = gs_sweep_pipe_lds(nlev);
  if (w <= 8)
    hipLaunchKernelGGL(k_dilu_factor_pipe<8>, dim3(1), dim3(kPipeBlock), lds, s, rows, lp, nlev,
                       n, w, ecol, elen, eji, eij, ac, dinv);
  else if (w <= 16)
    hipLaunchKernelGGL(k_dilu_factor_pipe<16>, dim3(1), dim3(kPipeBlock), lds, s, rows, lp, nlev,
                       n, w, ecol, elen, eji, eij, ac, dinv);
  else
    hipLaunchKernelGGL(k_dilu_factor_pipe<24>, dim3(1), dim3(kPipeBlock), lds, s, rows, lp, nlev,
                       n, w, ecol, elen, eji, eij, ac, dinv);
}

// eval[e] = anb[esrc[e]] (0 where esrc is -1)
__global__ void k_pipe_gather(int64_t m, const int32_t *__restrict__ esrc,
                              const double *__restrict__ anb, double *__restrict__ eval) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const int32_t k = esrc[e];
  eval[e] = k >= 0 ? anb[k] : 0.0;
}

void pipe_gather(hipStream_t s, int64_t m, const int32_t *esrc, const double *anb, double *eval) {
  if (m <= 0) return;
  hipLaunchKernelGGL(k_pipe_gather, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, m, esrc,
                     anb, eval);
}

void dilu_pipe(hipStream_t s, bool backward, const int32_t *rows, const int32_t *lp, int32_t nlev,
               int32_t n, int32_t w, const int32_t *ecol, const int32_t *elen, const double *eval,
               const double *dinv, double *rc, double *y, double *x, const double *stop) {
  if (nlev <= 0) return;
  const size_t lds = gs_sweep_pipe_lds(nlev);
#define UFVM_DPIPE(NE)                                                                            \
  do {                                                                                            \
    if (backward)                                                                                 \
      hipLaunchKernelGGL((k_dilu_pipe<NE, true>), dim3(1), dim3(kPipeBlock), lds, s, rows, lp,   \
                         nlev, n, w, ecol, elen, eval, dinv, rc, y, x, stop);                     \
    else                                                                                          \
      hipLaunchKernelGGL((k_dilu_pipe<NE, false>), dim3(1), dim3(kPipeBlock), lds, s, rows, lp,  \
                         nlev, n, w, ecol, elen, eval, dinv, rc, y, x, stop);                     \
  } while (0)
  if (w <= 8)
    UFVM_DPIPE(8);
  else if (w <= 16)
    UFVM_DPIPE(16);
  else
    UFVM_DPIPE(24);
#undef UFVM_DPIPE
}


void gs_sweep_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                    const int32_t *ptr, const int32_t *col, const double *ac, const double *anb,
                    const double *b, double *x, const double *stop) {
  if (nlev <= 0) return;
  hipLaunchKernelGGL(k_gs_sweep_block, dim3(1), dim3(kSweepBlock), 0, s, rows, lp, nlev, ptr,
                     col, ac, anb, b, x, stop);
}

void gs_rows(hipStream_t s, const int32_t *rows, int32_t count, const int32_t *ptr,
             const int32_t *col, const double *ac, const double *anb, const double *b,
             double *x, const double *stop) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_gs_rows, dim3(nblocks(count)), dim3(kBlock), 0, s, rows, count, ptr, col,
                     ac, anb, b, x, stop);
}

// d_j = ac_j - sum_{i<j, increasing i, i != skip} (a_ji * dinv_i) * a_ij ; dinv_j = 1/d_j
__device__ __forceinline__ void dilu_factor_row(int32_t j, const double *__restrict__ ac,
                                                const double *__restrict__ anb,
                                                const int32_t *__restrict__ lptr,
                                                const int32_t *__restrict__ lnb,
                                                const int32_t *__restrict__ lpji,
                                                const int32_t *__restrict__ lpij,
                                                int32_t skip_row, double *dinv) {
  const double d = dilu_factor_sum(lnb, lpji, lpij, anb, [dinv](int32_t i) { return dinv[i]; },
                                   skip_row, lptr[j], lptr[j + 1], ac[j]);
  dinv[j] = 1.0 / d;
}

__global__ void __launch_bounds__(kBlock)
k_dilu_factor(const int32_t *__restrict__ rows, int32_t count, const double *__restrict__ ac,
              const double *__restrict__ anb, const int32_t *__restrict__ lptr,
              const int32_t *__restrict__ lnb, const int32_t *__restrict__ lpji,
              const int32_t *__restrict__ lpij, int32_t skip_row, double *dinv) {
  const int32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  dilu_factor_row(rows[t], ac, anb, lptr, lnb, lpji, lpij, skip_row, dinv);
}

__global__ void __launch_bounds__(kSweepBlock)
k_dilu_factor_block(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp,
                    int32_t nlev, const double *__restrict__ ac, const double *__restrict__ anb,
                    const int32_t *__restrict__ lptr, const int32_t *__restrict__ lnb,
                    const int32_t *__restrict__ lpji, const int32_t *__restrict__ lpij,
                    int32_t skip_row, double *dinv) {
  for (int32_t q = 0; q < nlev; ++q) {
    for (int32_t t = lp[q] + threadIdx.x; t < lp[q + 1]; t += kSweepBlock)
      dilu_factor_row(rows[t], ac, anb, lptr, lnb, lpji, lpij, skip_row, dinv);
    __syncthreads();
  }
}

void dilu_factor_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                       const double *ac, const double *anb, const int32_t *lptr,
                       const int32_t *lnb, const int32_t *lpji, const int32_t *lpij,
                       int32_t skip_row, double *dinv) {
  if (nlev <= 0) return;
  hipLaunchKernelGGL(k_dilu_factor_block, dim3(1), dim3(kSweepBlock), 0, s, rows, lp, nlev, ac,
                     anb, lptr, lnb, lpji, lpij, skip_row, dinv);
}

void dilu_factor_rows(hipStream_t s, const int32_t *rows, int32_t count, const double *ac,
                      const double *anb, const int32_t *lptr, const int32_t *lnb,
                      const int32_t *lpji, const int32_t *lpij, int32_t skip_row, double *dinv) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_dilu_factor, dim3(nblocks(count)), dim3(kBlock), 0, s, rows, count, ac, anb,
                     lptr, lnb, lpji, lpij, skip_row, dinv);
}

// ILU(0) (IKJ form) of the rows [r0, r1) of one colour class, in a numbering
// where earlier colours are smaller indices: F holds A's off-diagonals on
// entry and the factor's on return (L~ = strictly lower part, U~ = strictly
// upper part, M = (D + L~) D^-1 (D + U~)).  For each lower neighbour k of j
// in increasing order, l = F_jk dinv_k, then every listed entry (k, c), c > k,
// updates the diagonal (c == j) or F_jc.  Rows of one colour never read each
// other's values, and rows of earlier colours are final.
__global__ void __launch_bounds__(kBlock)
k_ilu0_factor(int32_t r0, int32_t r1, const double *__restrict__ ac,
              const int32_t *__restrict__ lptr, const int32_t *__restrict__ lnb,
              const int32_t *__restrict__ lpji, const int32_t *__restrict__ hptr,
              const int32_t *__restrict__ hsrc, const int32_t *__restrict__ hdst, double *F,
              double *dinv) {
  const int32_t j = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (j >= r1) return;
  double d = ac[j];
  for (int32_t o = lptr[j]; o < lptr[j + 1]; ++o) {
    const double l = F[lpji[o]] * dinv[lnb[o]];
    for (int32_t h = hptr[o]; h < hptr[o + 1]; ++h) {
      const int32_t t = hdst[h];
      const double u = F[hsrc[h]];
      if (t < 0) d = d - l * u;
      else F[t] = F[t] - l * u;
    }
  }
  dinv[j] = 1.0 / d;
}

void ilu0_factor_rows(hipStream_t s, int32_t r0, int32_t r1, const double *ac,
                      const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
                      const int32_t *hptr, const int32_t *hsrc, const int32_t *hdst, double *F,
                      double *dinv) {
  if (r1 <= r0) return;
  hipLaunchKernelGGL(k_ilu0_factor, dim3(nblocks(r1 - r0)), dim3(kBlock), 0, s, r0, r1, ac, lptr,
                     lnb, lpji, hptr, hsrc, hdst, F, dinv);
}

// rc_j -= a_ji * (dinv_i * rc_i) for lower i in increasing order
template <typename V>
__device__ __forceinline__ void dilu_fwd_row(int32_t j, const V *__restrict__ anb,
                                             const int32_t *__restrict__ lptr,
                                             const int32_t *__restrict__ lnb,
                                             const int32_t *__restrict__ lpji,
                                             const V *__restrict__ dinv, V *rc) {
  rc[j] = dilu_lower_sum(lnb, lpji, anb, dinv, [rc](int32_t i) { return rc[i]; }, lptr[j],
                         lptr[j + 1], rc[j]);
}

template <typename V>
__global__ void __launch_bounds__(kBlock)
k_dilu_fwd(const int32_t *__restrict__ rows, int32_t count, const V *__restrict__ anb,
           const int32_t *__restrict__ lptr, const int32_t *__restrict__ lnb,
           const int32_t *__restrict__ lpji, const V *__restrict__ dinv, V *rc,
           const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  const int32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  dilu_fwd_row<V>(rows[t], anb, lptr, lnb, lpji, dinv, rc);
}

template <typename V>
__global__ void __launch_bounds__(kSweepBlock)
k_dilu_fwd_block(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
                 const V *__restrict__ anb, const int32_t *__restrict__ lptr,
                 const int32_t *__restrict__ lnb, const int32_t *__restrict__ lpji,
                 const V *__restrict__ dinv, V *rc, const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  for (int32_t q = 0; q < nlev; ++q) {
    for (int32_t t = lp[q] + threadIdx.x; t < lp[q + 1]; t += kSweepBlock)
      dilu_fwd_row<V>(rows[t], anb, lptr, lnb, lpji, dinv, rc);
    __syncthreads();
  }
}

template <typename V>
void dilu_fwd_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                    const V *anb, const int32_t *lptr, const int32_t *lnb, const int32_t *lpji,
                    const V *dinv, V *rc, const double *stop) {
  if (nlev <= 0) return;
  hipLaunchKernelGGL(k_dilu_fwd_block<V>, dim3(1), dim3(kSweepBlock), 0, s, rows, lp, nlev, anb,
                     lptr, lnb, lpji, dinv, rc, stop);
}

template <typename V>
void dilu_fwd_rows(hipStream_t s, const int32_t *rows, int32_t count, const V *anb,
                   const int32_t *lptr, const int32_t *lnb, const int32_t *lpji, const V *dinv,
                   V *rc, const double *stop) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_dilu_fwd<V>, dim3(nblocks(count)), dim3(kBlock), 0, s, rows, count, anb,
                     lptr, lnb, lpji, dinv, rc, stop);
}

// rc_i -= a_ij rc_j (j > i, cconn order) ; z = dinv_i rc_i ; rc_i = z ; x_i += z
template <typename V>
__device__ __forceinline__ void dilu_bwd_row(int32_t i, int32_t n,
                                             const int32_t *__restrict__ ptr,
                                             const int32_t *__restrict__ col,
                                             const V *__restrict__ anb,
                                             const V *__restrict__ dinv, V *rc, V *x) {
  V v = rc[i];
  if (i < n - 1) v = dilu_upper_sum(col, anb, [rc](int32_t j) { return rc[j]; }, i, ptr[i], ptr[i + 1], v);
  const V z = dinv[i] * v;
  rc[i] = z;
  x[i] = x[i] + z;
}

template <typename V>
__global__ void __launch_bounds__(kBlock)
k_dilu_bwd(const int32_t *__restrict__ rows, int32_t count, int32_t n,
           const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
           const V *__restrict__ anb, const V *__restrict__ dinv, V *rc, V *x,
           const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  const int32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= count) return;
  dilu_bwd_row<V>(rows[t], n, ptr, col, anb, dinv, rc, x);
}

template <typename V>
__global__ void __launch_bounds__(kSweepBlock)
k_dilu_bwd_block(const int32_t *__restrict__ rows, const int32_t *__restrict__ lp, int32_t nlev,
                 int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                 const V *__restrict__ anb, const V *__restrict__ dinv, V *rc, V *x,
                 const double *__restrict__ stop) {
  if (stop && *stop != 0.0) return;  // stopped by the device-side test
  for (int32_t q = 0; q < nlev; ++q) {
    for (int32_t t = lp[q] + threadIdx.x; t < lp[q + 1]; t += kSweepBlock)
      dilu_bwd_row<V>(rows[t], n, ptr, col, anb, dinv, rc, x);
    __syncthreads();
  }
}

template <typename V>
void dilu_bwd_block(hipStream_t s, const int32_t *rows, const int32_t *lp, int32_t nlev,
                    int32_t n, const int32_t *ptr, const int32_t *col, const V *anb,
                    const V *dinv, V *rc, V *x, const double *stop) {
  if (nlev <= 0) return;
  hipLaunchKernelGGL(k_dilu_bwd_block<V>, dim3(1), dim3(kSweepBlock), 0, s, rows, lp, nlev, n,
                     ptr, col, anb, dinv, rc, x, stop);
}

template <typename V>
void dilu_bwd_rows(hipStream_t s, const int32_t *rows, int32_t count, int32_t n,
                   const int32_t *ptr, const int32_t *col, const V *anb, const V *dinv, V *rc,
                   V *x, const double *stop) {
  if (count <= 0) return;
  hipLaunchKernelGGL(k_dilu_bwd<V>, dim3(nblocks(count)), dim3(kBlock), 0, s, rows, count, n,
                     ptr, col, anb, dinv, rc, x, stop);
}

template void dilu_fwd_block<double>(hipStream_t, const int32_t *, const int32_t *, int32_t,
                                     const double *, const int32_t *, const int32_t *,
                                     const int32_t *, const double *, double *, const double *);
template void dilu_fwd_block<float>(hipStream_t, const int32_t *, const int32_t *, int32_t,
                                    const float *, const int32_t *, const int32_t *,
                                    const int32_t *, const float *, float *, const double *);
template void dilu_bwd_block<double>(hipStream_t, const int32_t *, const int32_t *, int32_t,
                                     int32_t, const int32_t *, const int32_t *, const double *,
                                     const double *, double *, double *, const double *);
template void dilu_bwd_block<float>(hipStream_t, const int32_t *, const int32_t *, int32_t,
                                    int32_t, const int32_t *, const int32_t *, const float *,
                                    const float *, float *, float *, const double *);
template void dilu_fwd_rows<double>(hipStream_t, const int32_t *, int32_t, const double *,
                                    const int32_t *, const int32_t *, const int32_t *,
                                    const double *, double *, const double *);
template void dilu_fwd_rows<float>(hipStream_t, const int32_t *, int32_t, const float *,
                                   const int32_t *, const int32_t *, const int32_t *,
                                   const float *, float *, const double *);
template void dilu_bwd_rows<double>(hipStream_t, const int32_t *, int32_t, int32_t,
                                    const int32_t *, const int32_t *, const double *,
                                    const double *, double *, double *, const double *);
template void dilu_bwd_rows<float>(hipStream_t, const int32_t *, int32_t, int32_t,
                                   const int32_t *, const int32_t *, const float *,
                                   const float *, float *, float *, const double *);

template <typename V>
__global__ void __launch_bounds__(kBlock)
k_restrict(int32_t nc, const int32_t *__restrict__ cptr, const int32_t *__restrict__ child,
           const V *__restrict__ r, V *__restrict__ bc) {
  const int32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nc) return;
  V s = 0;
  for (int32_t c = cptr[p]; c < cptr[p + 1]; ++c) s = s + r[child[c]];
  bc[p] = s;
}

void restrict_ref(hipStream_t s, int32_t nc, const int32_t *cptr, const int32_t *child,
                  const double *r, double *bc_c) {
  if (nc <= 0) return;
  hipLaunchKernelGGL(k_restrict<double>, dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, cptr, child,
                     r, bc_c);
}


template <typename V>
__global__ void __launch_bounds__(kBlock)
k_prolong(int32_t n, const int32_t *__restrict__ parent, const V *__restrict__ xc,
          V *__restrict__ x) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + xc[parent[i]];
}

void prolong_ref(hipStream_t s, int32_t n, const int32_t *parent, const double *xc, double *x) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_prolong<double>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, parent, xc, x);
}

template <typename V>
__global__ void __launch_bounds__(kBlock)
k_prolong_from(int32_t r0, int32_t n, const int32_t *__restrict__ parent,
               const V *__restrict__ xc, V *__restrict__ x) {
  const int32_t i = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  x[i] = x[i] + xc[parent[i]];
}

template <typename V>
void prolong_fast(hipStream_t s, int32_t r0, int32_t n, const int32_t *parent, const V *xc,
                  V *x) {
  if (n <= r0) return;
  hipLaunchKernelGGL(k_prolong_from<V>, dim3(nblocks(n - r0)), dim3(kBlock), 0, s, r0, n, parent,
                     xc, x);
}

// Galerkin coarse operator, one thread per coarse value (the nc diagonal
// values, then the off-diagonal slots): the value is 0 plus its sources in
// list order, which is the reference's accumulation order for that value.
// Sources e >= 0 are fine anb[e], e < 0 fine ac[-1 - e].  Four sources are
// loaded ahead of their sums.
__global__ void __launch_bounds__(kBlock)
k_coarse_lhs(int32_t nout, int32_t nc, const int32_t *__restrict__ ptr,
             const int32_t *__restrict__ src, const double *__restrict__ fac,
             const double *__restrict__ fanb, double *__restrict__ cac,
             double *__restrict__ canb) {
  const int32_t o = blockIdx.x * kBlock + threadIdx.x;
  if (o >= nout) return;
  const int32_t m0 = ptr[o], m1 = ptr[o + 1];
  double v = 0.0;
  int32_t m = m0;
  for (; m + 4 <= m1; m += 4) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int32_t e = src[m + u];
      x[u] = e >= 0 ? fanb[e] : fac[-1 - e];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) v = v + x[u];
  }
  for (; m < m1; ++m) {
    const int32_t e = src[m];
    v = v + (e >= 0 ? fanb[e] : fac[-1 - e]);
  }
  if (o < nc) cac[o] = v;
  else canb[o - nc] = v;
}

void coarse_lhs(hipStream_t s, int32_t nout, int32_t nc, const int32_t *ptr, const int32_t *src,
                const double *fac, const double *fanb, double *cac, double *canb) {
  if (nout <= 0) return;
  hipLaunchKernelGGL(k_coarse_lhs, dim3(nblocks(nout)), dim3(kBlock), 0, s, nout, nc, ptr, src,
                     fac, fanb, cac, canb);
}

// --------------------------------------------------------- multicolour mode

// ---- kernels on the sliced-ELL layout: rows in chunks of 64, chunk q holds
// its entries column-major at off[q] + k*64 + lane for k < (off[q+1]-off[q])/64
// (a wavefront reads 64 consecutive words per k).  Each row's entries keep its
// cconn order and are followed by padding entries (value 0; their column is
// the row itself, or in the 16-bit form the row's last real column), so the
// loops run a uniform trip count with no exit test and the loads of four
// entries are issued before their products are summed.  Column indices come
// through ColSrc: int32 per entry, or a chunk base plus a uint16 per entry.

// A row's entries are taken in runs: the run's index and value loads are
// issued together, then its neighbour gathers, then its products in entry
// order, so each run waits for one round trip to its entries and one to its
// neighbours.  Runs have kSellRun entries, the last one the remainder (w is
// the same for the 64 rows of a chunk, so a wavefront takes one branch).

// v -/+ sum over the N entries from slot e0 of val x[col] (NEG: subtract);
// ZA: neighbours with index >= zero_above read as 0
template <int N, typename V, bool NEG, bool ZA, bool NT, typename CS>
__device__ __forceinline__ V sell_run(const CS &col, const V *__restrict__ val, const V *x,
                                      int32_t e0, V v, int32_t zero_above) {
  int32_t j[N];
  V a[N], xj[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = NT ? col.nt(e0 + u * 64) : col[e0 + u * 64];
    a[u] = NT ? __builtin_nontemporal_load(val + e0 + u * 64) : val[e0 + u * 64];
  }
#pragma unroll
  for (int u = 0; u < N; ++u) xj[u] = x[j[u]];
  if (ZA) {
    // the gathers stay issued together here, not moved under the guards
#pragma unroll
    for (int u = 0; u < N; ++u) asm volatile("" : "+v"(xj[u]));
#pragma unroll
    for (int u = 0; u < N; ++u) xj[u] = j[u] < zero_above ? xj[u] : V(0);
  }
#pragma unroll
  for (int u = 0; u < N; ++u) v = NEG ? v - a[u] * xj[u] : v + a[u] * xj[u];
  return v;
}

// v -/+ sum_k val_k x[col_k] over the row's w entries from slot base, in
// entry order
template <typename V, bool NEG, bool ZA, typename CS, bool NT = false>
__device__ __forceinline__ V sell_row(const CS &col, const V *__restrict__ val,
                                      const V *x, int32_t base, int32_t w, V v,
                                      int32_t zero_above) {
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun)
    v = sell_run<kSellRun, V, NEG, ZA, NT>(col, val, x, base + k * 64, v, zero_above);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: return sell_run<1, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    case 2: return sell_run<2, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    case 3: return sell_run<3, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    case 4: return sell_run<4, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    case 5: return sell_run<5, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    case 6: return sell_run<6, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    case 7: return sell_run<7, V, NEG, ZA, NT>(col, val, x, e0, v, zero_above);
    default: return v;
  }
}

// Wide rows.  The coarse levels' rows hold 13-16 entries on average (up to
// ~35), so runs of kSellRun wait for two or three round trips per row.  The
// wide form issues the index and value loads of every entry of the row at
// once, then every neighbour gather (xv(j): the neighbour's value), then the
// products in entry order: the arithmetic and its order are those of
// sell_row.  The row's w entries are loaded as N (w rounded up to a multiple
// of kSellRun); slots past w repeat entry w-1 (same address, in bounds) and
// are not summed, so the loads form one block with no branches between them.
// w is the same for the 64 rows of a chunk, so a wavefront takes one branch.
constexpr int kWideMax = 32;

template <int N, typename V, bool NEG, typename CS, typename XV>
__device__ __forceinline__ V sell_row_n(const CS &col, const V *__restrict__ val, const XV &xv,
                                        int32_t base, int32_t w, V v) {
  int32_t j[N];
  V a[N], xj[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const int32_t e = base + (u < w ? u : w - 1) * 64;
    j[u] = col[e];
    a[u] = val[e];
  }
#pragma unroll
  for (int u = 0; u < N; ++u) xj[u] = xv(j[u]);
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const V t = NEG ? v - a[u] * xj[u] : v + a[u] * xj[u];
    v = u < w ? t : v;
  }
  return v;
}

// v -/+ sum over the row's w entries (kSellRun < w <= kWideMax) in one round
template <typename V, bool NEG, typename CS, typename XV>
__device__ __forceinline__ V sell_row_wide(const CS &col, const V *__restrict__ val, const XV &xv,
                                           int32_t base, int32_t w, V v) {
  static_assert(kWideMax == 4 * kSellRun, "wide buckets");
  if (w <= 2 * kSellRun) return sell_row_n<2 * kSellRun, V, NEG>(col, val, xv, base, w, v);
  if (w <= 3 * kSellRun) return sell_row_n<3 * kSellRun, V, NEG>(col, val, xv, base, w, v);
  return sell_row_n<4 * kSellRun, V, NEG>(col, val, xv, base, w, v);
}

// Gauss-Seidel update of the contiguous rows [r0, r1) of one colour class.
// Rows of one colour are mutually independent, so every read of x[col] sees
// the value from an earlier colour pass.  Neighbours with index >= zero_above
// belong to colours not yet visited in a sweep that started from x = 0 and
// are read as 0 (so x needs no clearing first).
// row i of the colour class ending at r1
template <typename V, bool PK>
__device__ __forceinline__ void mcgs_row(int32_t i, int32_t r1, const int32_t *__restrict__ off,
                                         const SellCols &cols, const V *__restrict__ val,
                                         const V *__restrict__ ac, const V *__restrict__ b, V *x,
                                         int32_t zero_above) {
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  const V d = ac[i];
  // zero_above beyond this colour's rows: every neighbour holds a value
  const V v = (zero_above <= r1)
                  ? sell_row<V, true, true>(col, val, x, base, w, b[i], zero_above)
                  : sell_row<V, true, false>(col, val, x, base, w, b[i], zero_above);
  x[i] = v / d;
}

template <typename V, bool PK>
__global__ void __launch_bounds__(kBlock)
k_mcgs_sell(int32_t r0, int32_t r1, const int32_t *__restrict__ off, SellCols cols,
            const V *__restrict__ val, const V *__restrict__ ac, const V *__restrict__ b, V *x,
            int32_t zero_above) {
  const int32_t i = (r0 & ~63) + xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i < r0 || i >= r1) return;
  mcgs_row<V, PK>(i, r1, off, cols, val, ac, b, x, zero_above);
}

template <typename V>
void mcgs_sell(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off, const SellCols &cols,
               const V *val, const V *ac, const V *b, V *x, int32_t zero_above) {
  if (r1 <= r0) return;
  const int32_t a = r0 & ~63;
  with_cols(cols, [&](auto pk) {
    hipLaunchKernelGGL((k_mcgs_sell<V, decltype(pk)::value>), dim3(nblocks(r1 - a)),
                       dim3(kBlock), 0, s, r0, r1, off, cols, val, ac, b, x, zero_above);
  });
}

// ---- finest level with FP16 off-diagonals scaled by their row's diagonal
// (s_ij = a_ij / a_ii, |s_ij| <= 1 on diagonally dominant rows), arithmetic
// in FP32: a Gauss-Seidel row is x_i = b_i / a_ii - sum_j s_ij x_j, a
// residual r_i = b_i + a_ii (-x_i - sum_j s_ij x_j).  Rows in runs of
// kSellRun as sell_row (same padding: value 0).

template <int N, bool ZA, bool NT, typename CS>
__device__ __forceinline__ float hsell_run(const CS &col, const _Float16 *__restrict__ val,
                                           const float *x, int32_t e0, float v,
                                           int32_t zero_above) {
  int32_t j[N];
  float a[N], xj[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = NT ? col.nt(e0 + u * 64) : col[e0 + u * 64];
    a[u] = (float)(NT ? __builtin_nontemporal_load(val + e0 + u * 64) : val[e0 + u * 64]);
  }
#pragma unroll
  for (int u = 0; u < N; ++u) xj[u] = x[j[u]];
  if (ZA) {
#pragma unroll
    for (int u = 0; u < N; ++u) asm volatile("" : "+v"(xj[u]));
#pragma unroll
    for (int u = 0; u < N; ++u) xj[u] = j[u] < zero_above ? xj[u] : 0.0f;
  }
#pragma unroll
  for (int u = 0; u < N; ++u) v = v - a[u] * xj[u];
  return v;
}

// v - sum over the row's w entries from slot base of s x
template <bool ZA, bool NT, typename CS>
__device__ __forceinline__ float hsell_row(const CS &col, const _Float16 *__restrict__ val,
                                           const float *x, int32_t base, int32_t w, float v,
                                           int32_t zero_above) {
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun)
    v = hsell_run<kSellRun, ZA, NT>(col, val, x, base + k * 64, v, zero_above);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: return hsell_run<1, ZA, NT>(col, val, x, e0, v, zero_above);
    case 2: return hsell_run<2, ZA, NT>(col, val, x, e0, v, zero_above);
    case 3: return hsell_run<3, ZA, NT>(col, val, x, e0, v, zero_above);
    case 4: return hsell_run<4, ZA, NT>(col, val, x, e0, v, zero_above);
    case 5: return hsell_run<5, ZA, NT>(col, val, x, e0, v, zero_above);
    case 6: return hsell_run<6, ZA, NT>(col, val, x, e0, v, zero_above);
    case 7: return hsell_run<7, ZA, NT>(col, val, x, e0, v, zero_above);
    default: return v;
  }
}

template <bool PK>
__global__ void __launch_bounds__(kBlock)
k_mcgs_sell_h(int32_t r0, int32_t r1, const int32_t *__restrict__ off, SellCols cols,
              const _Float16 *__restrict__ val, const float *__restrict__ ac,
              const float *__restrict__ b, float *x, int32_t zero_above) {
  const int32_t i = (r0 & ~63) + xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i < r0 || i >= r1) return;
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  const float v0 = b[i] / ac[i];
  x[i] = (zero_above <= r1) ? hsell_row<true, false>(col, val, x, base, w, v0, zero_above)
                            : hsell_row<false, false>(col, val, x, base, w, v0, zero_above);
}

void mcgs_sell_h(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off, const SellCols &cols,
                 const _Float16 *val, const float *ac, const float *b, float *x,
                 int32_t zero_above) {
  if (r1 <= r0) return;
  const int32_t a = r0 & ~63;
  with_cols(cols, [&](auto pk) {
    hipLaunchKernelGGL((k_mcgs_sell_h<decltype(pk)::value>), dim3(nblocks(r1 - a)), dim3(kBlock),
                       0, s, r0, r1, off, cols, val, ac, b, x, zero_above);
  });
}

template <bool PK>
__global__ void __launch_bounds__(kBlock)
k_residual_sell_h(int32_t n, int32_t zero_rows, const int32_t *__restrict__ off, SellCols cols,
                  const _Float16 *__restrict__ val, const float *__restrict__ ac,
                  const float *__restrict__ b, const float *__restrict__ x,
                  float *__restrict__ r) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (i < zero_rows) {
    r[i] = 0.0f;
    return;
  }
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  r[i] = b[i] + ac[i] * hsell_row<false, true>(col, val, x, base, w, -x[i], 0);
}

void residual_sell_h(hipStream_t s, int32_t n, int32_t zero_rows, const int32_t *off,
                     const SellCols &cols, const _Float16 *val, const float *ac, const float *b,
                     const float *x, float *r) {
  if (n <= 0) return;
  with_cols(cols, [&](auto pk) {
    hipLaunchKernelGGL((k_residual_sell_h<decltype(pk)::value>), dim3(nblocks(n)), dim3(kBlock),
                       0, s, n, zero_rows, off, cols, val, ac, b, x, r);
  });
}

// out16[slot] = val[slot] / ac[row] rounded to FP16, over every slot of the
// n rows' sliced-ELL chunks (padding slots hold 0)
__global__ void __launch_bounds__(kBlock)
k_scale_half(int32_t n, const int32_t *__restrict__ off, const double *__restrict__ val,
             const double *__restrict__ ac, _Float16 *__restrict__ out16) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const double d = ac[i];
  for (int32_t k = 0; k < w; ++k) out16[base + k * 64] = (_Float16)(val[base + k * 64] / d);
}

void scale_half(hipStream_t s, int32_t n, const int32_t *off, const double *val, const double *ac,
                _Float16 *out16) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_half, dim3(nblocks(n)), dim3(kBlock), 0, s, n, off, val, ac, out16);
}

// ---- DILU in colour order on the sliced-ELL copy of a level whose colour
// classes are contiguous row ranges (earlier colours = lower rows).  Forward
// pass over the colour rows [r0, r1):
//   rc_i = (b_i - ac_i x_i - sum_k a_ik x_k) - sum_{k < r0} a_ik (dinv_k rc_k)
// both sums in entry order (ZERO: x = 0, the residual part is b_i).
// Backward pass over [r0, r1), colours downward:
//   z_i = dinv_i (rc_i - sum_{k >= r1} a_ik rc_k),  rc_i = z_i,  x_i += z_i
// (ZERO: x_i = z_i).  Padding entries (col = the row) fall in neither sum.
// For ILU(0) the sums over lower / upper neighbours use the factor's values
// instead of A's: SEP takes the lower ones from lval while the residual keeps
// val (A); with ZERO, or for the backward pass, the caller passes the factor
// as val.
template <typename V, bool ZERO, bool SEP, bool PK>
__global__ void __launch_bounds__(kBlock)
k_mcilu_fwd_sell(int32_t r0, int32_t r1, const int32_t *__restrict__ off, SellCols cols,
                 const V *__restrict__ val, const V *__restrict__ lval, const V *__restrict__ ac,
                 const V *__restrict__ b, const V *__restrict__ x, const V *__restrict__ dinv,
                 V *rc) {
  const int32_t i = (r0 & ~63) + xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i < r0 || i >= r1) return;
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  V res = ZERO ? b[i] : b[i] - ac[i] * x[i];
  V low = V(0);
  int32_t k = 0;
  for (; k + 4 <= w; k += 4) {
    int32_t j[4];
    V a[4], f[4], xj[4], lj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      j[u] = col[base + (k + u) * 64];
      a[u] = val[base + (k + u) * 64];
      f[u] = SEP ? lval[base + (k + u) * 64] : a[u];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!ZERO) xj[u] = x[j[u]];
      lj[u] = j[u] < r0 ? dinv[j[u]] * rc[j[u]] : V(0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (!ZERO) res = res - a[u] * xj[u];
      if (j[u] < r0) low = low + f[u] * lj[u];
    }
  }
  for (; k < w; ++k) {
    const int32_t j = col[base + k * 64];
    const V a = val[base + k * 64];
    const V f = SEP ? lval[base + k * 64] : a;
    if (!ZERO) res = res - a * x[j];
    if (j < r0) low = low + f * (dinv[j] * rc[j]);
  }
  rc[i] = res - low;
}

template <typename V, bool ZERO, bool PK>
__global__ void __launch_bounds__(kBlock)
k_mcilu_bwd_sell(int32_t r0, int32_t r1, const int32_t *__restrict__ off, SellCols cols,
                 const V *__restrict__ val, const V *__restrict__ dinv, V *rc, V *x) {
  const int32_t i = (r0 & ~63) + xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i < r0 || i >= r1) return;
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  V v = rc[i];
  int32_t k = 0;
  for (; k + 4 <= w; k += 4) {
    int32_t j[4];
    V a[4], rj[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      j[u] = col[base + (k + u) * 64];
      a[u] = val[base + (k + u) * 64];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) rj[u] = j[u] >= r1 ? rc[j[u]] : V(0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (j[u] >= r1) v = v - a[u] * rj[u];
  }
  for (; k < w; ++k) {
    const int32_t j = col[base + k * 64];
    if (j >= r1) v = v - val[base + k * 64] * rc[j];
  }
  const V z = dinv[i] * v;
  rc[i] = z;
  x[i] = ZERO ? z : x[i] + z;
}

template <typename V>
void mcilu_fwd_sell(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off,
                    const SellCols &cols, const V *val, const V *lval, const V *ac, const V *b,
                    const V *x, const V *dinv, V *rc, bool zero) {
  if (r1 <= r0) return;
  const dim3 g(nblocks(r1 - (r0 & ~63))), t(kBlock);
  // from zero only the lower sum reads values
  const V *v0 = (zero && lval) ? lval : val;
  with_cols(cols, [&](auto pk) {
    constexpr bool PK = decltype(pk)::value;
    if (zero)
      hipLaunchKernelGGL((k_mcilu_fwd_sell<V, true, false, PK>), g, t, 0, s, r0, r1, off, cols,
                         v0, v0, ac, b, x, dinv, rc);
    else if (lval && lval != val)
      hipLaunchKernelGGL((k_mcilu_fwd_sell<V, false, true, PK>), g, t, 0, s, r0, r1, off, cols,
                         val, lval, ac, b, x, dinv, rc);
    else
      hipLaunchKernelGGL((k_mcilu_fwd_sell<V, false, false, PK>), g, t, 0, s, r0, r1, off, cols,
                         val, val, ac, b, x, dinv, rc);
  });
}

template <typename V>
void mcilu_bwd_sell(hipStream_t s, int32_t r0, int32_t r1, const int32_t *off,
                    const SellCols &cols, const V *val, const V *dinv, V *rc, V *x, bool zero) {
  if (r1 <= r0) return;
  const dim3 g(nblocks(r1 - (r0 & ~63))), t(kBlock);
  with_cols(cols, [&](auto pk) {
    constexpr bool PK = decltype(pk)::value;
    if (zero)
      hipLaunchKernelGGL((k_mcilu_bwd_sell<V, true, PK>), g, t, 0, s, r0, r1, off, cols, val,
                         dinv, rc, x);
    else
      hipLaunchKernelGGL((k_mcilu_bwd_sell<V, false, PK>), g, t, 0, s, r0, r1, off, cols, val,
                         dinv, rc, x);
  });
}

template void mcilu_fwd_sell<float>(hipStream_t, int32_t, int32_t, const int32_t *,
                                    const SellCols &, const float *, const float *,
                                    const float *, const float *, const float *, const float *,
                                    float *, bool);
template void mcilu_fwd_sell<double>(hipStream_t, int32_t, int32_t, const int32_t *,
                                     const SellCols &, const double *, const double *,
                                     const double *, const double *, const double *,
                                     const double *, double *, bool);
template void mcilu_bwd_sell<float>(hipStream_t, int32_t, int32_t, const int32_t *,
                                    const SellCols &, const float *, const float *, float *,
                                    float *, bool);
template void mcilu_bwd_sell<double>(hipStream_t, int32_t, int32_t, const int32_t *,
                                     const SellCols &, const double *, const double *, double *,
                                     double *, bool);

// r = b - A x; rows below zero_rows get r = 0 without computing it (the
// colour class a Gauss-Seidel sweep solved last, whose residual is rounding
// noise)
template <typename V, bool PK, bool WIDE = false>
__device__ __forceinline__ V residual_val(int32_t i, const int32_t *__restrict__ off,
                                          const SellCols &cols, const V *__restrict__ val,
                                          const V *__restrict__ ac, const V *b, const V *x) {
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  const V v0 = b[i] - ac[i] * x[i];
  if (WIDE && w > kSellRun && w <= kWideMax)
    return sell_row_wide<V, true>(col, val, [x](int32_t j) { return x[j]; }, base, w, v0);
  return sell_row<V, true, false, ColSrc<PK>, true>(col, val, x, base, w, v0, 0);
}

template <typename V, bool PK, bool WIDE = false>
__device__ __forceinline__ void residual_row(int32_t i, const int32_t *__restrict__ off,
                                             const SellCols &cols, const V *__restrict__ val,
                                             const V *__restrict__ ac, const V *b, const V *x,
                                             V *r) {
  r[i] = residual_val<V, PK, WIDE>(i, off, cols, val, ac, b, x);
}

template <typename V, bool PK, bool WIDE>
__global__ void __launch_bounds__(kBlock)
k_residual_sell(int32_t n, int32_t zero_rows, const int32_t *__restrict__ off, SellCols cols,
                const V *__restrict__ val, const V *__restrict__ ac, const V *__restrict__ b,
                const V *__restrict__ x, V *__restrict__ r) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (i < zero_rows) {
    r[i] = V(0);
    return;
  }
  residual_row<V, PK, WIDE>(i, off, cols, val, ac, b, x, r);
}

template <typename V>
void residual_sell(hipStream_t s, int32_t n, int32_t zero_rows, const int32_t *off,
                   const SellCols &cols, const V *val, const V *ac, const V *b, const V *x,
                   V *r, bool wide) {
  if (n <= 0) return;
  with_cols(cols, [&](auto pk) {
    constexpr bool PK = decltype(pk)::value;
    if (wide)
      hipLaunchKernelGGL((k_residual_sell<V, PK, true>), dim3(nblocks(n)), dim3(kBlock), 0, s, n,
                         zero_rows, off, cols, val, ac, b, x, r);
    else
      hipLaunchKernelGGL((k_residual_sell<V, PK, false>), dim3(nblocks(n)), dim3(kBlock), 0, s,
                         n, zero_rows, off, cols, val, ac, b, x, r);
  });
}

// v_j = x_j (+ xc[parent_j] when PRO: prolongation fused into the sweep)
template <typename V, bool PRO>
__device__ __forceinline__ V jac_val(const V *x, const int32_t *__restrict__ parent, const V *xc,
                                     int32_t j) {
  return PRO ? x[j] + xc[parent[j]] : x[j];
}

// sell_run for the Jacobi sweep: v - sum over the N entries from slot e0 of
// val * (x + xc[parent] when PRO)
template <int N, typename V, bool PRO, typename CS>
__device__ __forceinline__ V jacobi_run(const CS &col, const V *__restrict__ val, const V *x,
                                        const int32_t *__restrict__ parent, const V *xc,
                                        int32_t e0, V v) {
  int32_t j[N];
  V a[N], xj[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = col[e0 + u * 64];
    a[u] = val[e0 + u * 64];
  }
#pragma unroll
  for (int u = 0; u < N; ++u) xj[u] = jac_val<V, PRO>(x, parent, xc, j[u]);
#pragma unroll
  for (int u = 0; u < N; ++u) v = v - a[u] * xj[u];
  return v;
}

// one damped Jacobi sweep: out = v + omega (b - A v) / ac, v = x or, with
// PRO, x + xc[parent] (the coarse correction added on the fly)
template <typename V, bool PRO, bool PK, bool WIDE = false>
__device__ __forceinline__ void jacobi_row(int32_t i, const int32_t *__restrict__ off,
                                           const SellCols &cols, const V *__restrict__ val,
                                           const V *__restrict__ ac, const V *b, const V *x,
                                           const int32_t *__restrict__ parent, const V *xc,
                                           V *out, V omega) {
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  const V vi = jac_val<V, PRO>(x, parent, xc, i), d = ac[i];
  V v = b[i] - d * vi;
  if (WIDE && w > kSellRun && w <= kWideMax) {
    v = sell_row_wide<V, true>(
        col, val, [x, parent, xc](int32_t j) { return jac_val<V, PRO>(x, parent, xc, j); }, base,
        w, v);
    out[i] = vi + omega * v / d;
    return;
  }
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun)
    v = jacobi_run<kSellRun, V, PRO>(col, val, x, parent, xc, base + k * 64, v);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: v = jacobi_run<1, V, PRO>(col, val, x, parent, xc, e0, v); break;
    case 2: v = jacobi_run<2, V, PRO>(col, val, x, parent, xc, e0, v); break;
    case 3: v = jacobi_run<3, V, PRO>(col, val, x, parent, xc, e0, v); break;
    case 4: v = jacobi_run<4, V, PRO>(col, val, x, parent, xc, e0, v); break;
    case 5: v = jacobi_run<5, V, PRO>(col, val, x, parent, xc, e0, v); break;
    case 6: v = jacobi_run<6, V, PRO>(col, val, x, parent, xc, e0, v); break;
    case 7: v = jacobi_run<7, V, PRO>(col, val, x, parent, xc, e0, v); break;
    default: break;
  }
  out[i] = vi + omega * v / d;
}

template <typename V, bool PRO, bool PK, bool WIDE>
__global__ void __launch_bounds__(kBlock)
k_jacobi_sell(int32_t n, const int32_t *__restrict__ off, SellCols cols,
              const V *__restrict__ val, const V *__restrict__ ac, const V *__restrict__ b,
              const V *__restrict__ x, const int32_t *__restrict__ parent,
              const V *__restrict__ xc, V *__restrict__ out, V omega) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  jacobi_row<V, PRO, PK, WIDE>(i, off, cols, val, ac, b, x, parent, xc, out, omega);
}

template <typename V>
void jacobi_sell(hipStream_t s, int32_t n, const int32_t *off, const SellCols &cols,
                 const V *val, const V *ac, const V *b, const V *x, const int32_t *parent,
                 const V *xc, V *out, V omega, bool wide) {
  if (n <= 0) return;
  const dim3 g(nblocks(n)), t(kBlock);
  with_cols(cols, [&](auto pk) {
    constexpr bool PK = decltype(pk)::value;
    if (parent && wide)
      hipLaunchKernelGGL((k_jacobi_sell<V, true, PK, true>), g, t, 0, s, n, off, cols, val, ac, b,
                         x, parent, xc, out, omega);
    else if (parent)
      hipLaunchKernelGGL((k_jacobi_sell<V, true, PK, false>), g, t, 0, s, n, off, cols, val, ac,
                         b, x, parent, xc, out, omega);
    else if (wide)
      hipLaunchKernelGGL((k_jacobi_sell<V, false, PK, true>), g, t, 0, s, n, off, cols, val, ac,
                         b, x, parent, xc, out, omega);
    else
      hipLaunchKernelGGL((k_jacobi_sell<V, false, PK, false>), g, t, 0, s, n, off, cols, val, ac,
                         b, x, parent, xc, out, omega);
  });
}

// Block-local multicolour Gauss-Seidel (coarse levels).  Rows are numbered
// block-major: block q owns rows [bptr[q], bptr[q+1]) (at most kBRows), sorted
// by a colouring of the block's own subgraph; boff[q*(ncol+1)+c] is the first
// row of colour c in block q.  The block's unknowns live in LDS during the
// sweeps: neighbours inside the block are read there (Gauss-Seidel), those
// outside from x_in (Jacobi coupling: their values from before this launch).
// Each sweep visits the colours upward, then downward skipping the last one
// (re-solving it would repeat the same arithmetic).  Initial in-block values:
//   mode 0: zero (pre-smoothing from x = 0; out-of-block reads are 0 too)
//   mode 1: x_in
//   mode 2: x_in + xc[parent] (prolongation fused into post-smoothing)
// Result: x_out[bptr[q] .. bptr[q+1]).  x_out must not alias x_in.
template <typename V>
__global__ void __launch_bounds__(kBlock)
k_bmcgs(const int32_t *__restrict__ bptr, const int32_t *__restrict__ boff, int32_t ncol,
        const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
        const V *__restrict__ ac, const V *__restrict__ anb, const V *__restrict__ b,
        const V *__restrict__ x_in, const int32_t *__restrict__ parent,
        const V *__restrict__ xc, V *__restrict__ x_out, int mode, int nsweep) {
  __shared__ V xs[kBRows];
  const int32_t q = blockIdx.x;
  const int32_t r0 = bptr[q], r1 = bptr[q + 1];
  for (int32_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
    V v = V(0);
    if (mode >= 1) v = x_in[i];
    if (mode == 2) v = v + xc[parent[i]];
    xs[i - r0] = v;
  }
  __syncthreads();
  const int32_t *off = boff + (int64_t)q * (ncol + 1);
  auto pass = [&](int c) {
    for (int32_t i = off[c] + threadIdx.x; i < off[c + 1]; i += kBlock) {
      V v = b[i];
      for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        const int32_t j = col[k];
        V xj;
        if (j >= r0 && j < r1) xj = xs[j - r0];
        else if (mode == 0) xj = V(0);
        else if (mode == 1) xj = x_in[j];
        else xj = x_in[j] + xc[parent[j]];
        v = v - anb[k] * xj;
      }
      xs[i - r0] = v / ac[i];
    }
    __syncthreads();
  };
  for (int s = 0; s < nsweep; ++s) {
    for (int c = (s == 0 ? 0 : 1); c < ncol; ++c) pass(c);
    for (int c = ncol - 2; c >= 0; --c) pass(c);
  }
  for (int32_t i = r0 + threadIdx.x; i < r1; i += kBlock) x_out[i] = xs[i - r0];
}

template <typename V>
void bmcgs(hipStream_t s, int32_t nblk, const int32_t *bptr, const int32_t *boff, int32_t ncol,
           const int32_t *ptr, const int32_t *col, const V *ac, const V *anb, const V *b,
           const V *x_in, const int32_t *parent, const V *xc, V *x_out, int mode, int nsweep) {
  if (nblk <= 0 || nsweep <= 0) return;
  hipLaunchKernelGGL(k_bmcgs<V>, dim3(nblk), dim3(kBlock), 0, s, bptr, boff, ncol, ptr, col, ac,
                     anb, b, x_in, parent, xc, x_out, mode, nsweep);
}

// restriction fused with the first Jacobi sweep from zero on the coarse level:
// bc[p] = sum of children's residuals; xc[p] = w bc[p] / ac[p]
// Restriction over children stored like the sliced-ELL matrix: coarse rows in
// chunks of 64, chunk q's child lists column-major at off[q] + k*64 + lane,
// padded with -1.  bc[p] = scale * (sum of the children's residuals in list
// order, increasing reference fine index); JAC: also the first Jacobi sweep
// from zero on the coarse level, xc[p] = w bc[p] / ac[p].  The coarse cycle is
// linear in bc, so scale multiplies the whole coarse correction.
// s + the residuals of the N child slots from e0 in list order (padding -1
// adds 0); the run's loads are issued together, padding reading child 0
template <int N, typename V>
__device__ __forceinline__ V restrict_run(const int32_t *__restrict__ child, const V *r, int32_t e0,
                                          V s) {
  int32_t c[N];
  V v[N];
#pragma unroll
  for (int u = 0; u < N; ++u) c[u] = child[e0 + u * 64];
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] = r[c[u] >= 0 ? c[u] : 0];
#pragma unroll
  for (int u = 0; u < N; ++u) asm volatile("" : "+v"(v[u]));
#pragma unroll
  for (int u = 0; u < N; ++u) s = s + (c[u] >= 0 ? v[u] : V(0));
  return s;
}

// the sum of restrict_run over a child list of w slots (kSellRun < w <= N)
// in one round: every slot's index load, then every residual load, then the
// sum in list order; slots past w repeat slot w-1 and are not summed
template <int N, typename V>
__device__ __forceinline__ V restrict_wide(const int32_t *__restrict__ child, const V *r,
                                           int32_t base, int32_t w) {
  int32_t c[N];
  V v[N];
#pragma unroll
  for (int u = 0; u < N; ++u) c[u] = child[base + (u < w ? u : w - 1) * 64];
#pragma unroll
  for (int u = 0; u < N; ++u) v[u] = r[c[u] >= 0 ? c[u] : 0];
  V s = 0;
#pragma unroll
  for (int u = 0; u < N; ++u) {
    const V t = s + (c[u] >= 0 ? v[u] : V(0));
    s = u < w ? t : s;
  }
  return s;
}

template <typename V, bool JAC>
__device__ __forceinline__ void restrict_row(int32_t p, const int32_t *__restrict__ off,
                                             const int32_t *__restrict__ child, const V *r,
                                             const V *__restrict__ acc, V *bc, V *xc, V omega,
                                             V scale) {
  const int32_t q = p >> 6;
  const int32_t base = off[q] + (p & 63), w = (off[q + 1] - off[q]) >> 6;
  V s = 0;
  if (w > kSellRun && w <= kWideMax) {
    if (w <= 2 * kSellRun) s = restrict_wide<2 * kSellRun>(child, r, base, w);
    else if (w <= 3 * kSellRun) s = restrict_wide<3 * kSellRun>(child, r, base, w);
    else s = restrict_wide<4 * kSellRun>(child, r, base, w);
    const V t = scale * s;
    bc[p] = t;
    if (JAC) xc[p] = omega * t / acc[p];
    return;
  }
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun) s = restrict_run<kSellRun>(child, r, base + k * 64, s);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: s = restrict_run<1>(child, r, e0, s); break;
    case 2: s = restrict_run<2>(child, r, e0, s); break;
    case 3: s = restrict_run<3>(child, r, e0, s); break;
    case 4: s = restrict_run<4>(child, r, e0, s); break;
    case 5: s = restrict_run<5>(child, r, e0, s); break;
    case 6: s = restrict_run<6>(child, r, e0, s); break;
    case 7: s = restrict_run<7>(child, r, e0, s); break;
    default: break;
  }
  const V t = scale * s;
  bc[p] = t;
  if (JAC) xc[p] = omega * t / acc[p];
}

template <typename V, bool JAC>
__global__ void __launch_bounds__(kBlock)
k_restrict_sell(int32_t nc, const int32_t *__restrict__ off, const int32_t *__restrict__ child,
                const V *__restrict__ r, const V *__restrict__ acc, V *__restrict__ bc,
                V *__restrict__ xc, V omega, V scale) {
  const int32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nc) return;
  restrict_row<V, JAC>(p, off, child, r, acc, bc, xc, omega, scale);
}

template <typename V>
void restrict_jacobi(hipStream_t s, int32_t nc, const int32_t *off, const int32_t *child,
                     const V *r, const V *acc, V *bc, V *xc, V omega, V scale) {
  if (nc <= 0) return;
  hipLaunchKernelGGL((k_restrict_sell<V, true>), dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, off,
                     child, r, acc, bc, xc, omega, scale);
}

template <typename V>
void restrict_fast(hipStream_t s, int32_t nc, const int32_t *off, const int32_t *child,
                   const V *r, V *bc, V scale) {
  if (nc <= 0) return;
  hipLaunchKernelGGL((k_restrict_sell<V, false>), dim3(nblocks(nc)), dim3(kBlock), 0, s, nc, off,
                     child, r, nullptr, bc, nullptr, V(0), scale);
}


template <typename V>
__global__ void __launch_bounds__(kBlock)
k_residual_fast(int32_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ col,
                const V *__restrict__ ac, const V *__restrict__ anb, const V *__restrict__ b,
                const V *__restrict__ x, V *__restrict__ r) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  V v = b[i] - ac[i] * x[i];
  for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) v = v - anb[k] * x[col[k]];
  r[i] = v;
}

template <typename V>
void residual_fast(hipStream_t s, int32_t n, const int32_t *ptr, const int32_t *col, const V *ac,
                   const V *anb, const V *b, const V *x, V *r) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_residual_fast<V>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, ptr, col, ac,
                     anb, b, x, r);
}

// x = Ainv * b, one wavefront per row, FP64 accumulation
template <typename V>
__device__ __forceinline__ void dense_row(int32_t row, int lane, int32_t n,
                                          const double *__restrict__ ainv, const V *b, V *x) {
  const double *a = ainv + (int64_t)row * n;
  // four independent partial sums keep four loads in flight per lane
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int32_t j = lane;
  for (; j + 3 * kWave < n; j += 4 * kWave) {
    s0 += a[j] * (double)b[j];
    s1 += a[j + kWave] * (double)b[j + kWave];
    s2 += a[j + 2 * kWave] * (double)b[j + 2 * kWave];
    s3 += a[j + 3 * kWave] * (double)b[j + 3 * kWave];
  }
  for (; j < n; j += kWave) s0 += a[j] * (double)b[j];
  double s = (s0 + s1) + (s2 + s3);
  s = wave_sum(s);
  if (lane == 0) x[row] = (V)s;
}

template <typename V>
__global__ void __launch_bounds__(kBlock)
k_dense_gemv(int32_t n, const double *__restrict__ ainv, const V *__restrict__ b,
             V *__restrict__ x) {
  const int32_t row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= n) return;
  dense_row<V>(row, lane, n, ainv, b, x);
}

template <typename V>
void dense_solve(hipStream_t s, int32_t n, const double *ainv, const V *b, V *x) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_dense_gemv<V>, dim3((n + kBlock / kWave - 1) / (kBlock / kWave)),
                     dim3(kBlock), 0, s, n, ainv, b, x);
}

template <typename V>
__global__ void k_fill(int64_t n, V *x, V v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = v;
}

template <typename V>
void fill(hipStream_t s, int64_t n, V *x, V v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill<V>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, x, v);
}

template <typename A, typename B>
__global__ void k_convert(int64_t n, const A *__restrict__ a, B *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}

template <typename A, typename B>
void convert(hipStream_t s, int64_t n, const A *a, B *b) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_convert<A, B>), dim3(nblocks(n)), dim3(kBlock), 0, s, n, a, b);
}

__global__ void k_gather(int32_t n, const int32_t *__restrict__ idx, const double *__restrict__ src,
                         double *__restrict__ dst) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_scatter(int32_t n, const int32_t *__restrict__ idx,
                          const double *__restrict__ src, double *__restrict__ dst) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}
void permute_gather(hipStream_t s, int32_t n, const int32_t *idx, const double *src, double *dst) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather, dim3(nblocks(n)), dim3(kBlock), 0, s, n, idx, src, dst);
}
void permute_scatter(hipStream_t s, int32_t n, const int32_t *idx, const double *src,
                     double *dst) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scatter, dim3(nblocks(n)), dim3(kBlock), 0, s, n, idx, src, dst);
}

// ---------------------------------------------------------- dense inverse

__global__ void k_csr_to_dense(int32_t n, const int32_t *__restrict__ ptr,
                               const int32_t *__restrict__ col, const double *__restrict__ ac,
                               const double *__restrict__ anb, double shift, double *a) {
  const int32_t i = blockIdx.x;
  double *row = a + (int64_t)i * n;
  for (int32_t j = threadIdx.x; j < n; j += blockDim.x) row[j] = shift;
  __syncthreads();
  if (threadIdx.x == 0) {
    row[i] += ac[i];
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) row[col[k]] += anb[k];
  }
}

void csr_to_dense(hipStream_t s, int32_t n, const int32_t *ptr, const int32_t *col,
                  const double *ac, const double *anb, double shift, double *a) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_csr_to_dense, dim3(n), dim3(kBlock), 0, s, n, ptr, col, ac, anb, shift, a);
}

// ---- blocked Gauss-Jordan inversion (no pivoting: the coarsest operators are
// symmetric positive definite, shifted when singular):
// per diagonal block K of at most kGjB rows, P = A_KK^-1, then
//   A_KJ <- P A_KJ,  A_IJ <- A_IJ - A_IK (P A_KJ),  A_IK <- -A_IK P,  A_KK <- P
// for I, J outside K: one block step is kGjB one-column steps at once, and
// the rank-kGjB update replaces kGjB launches over the whole matrix.

constexpr int kGjB = 64;
constexpr int kGjT = 256;  // threads of the block-step kernels: 16 x 16, a 4 x 4 tile each

// P = inverse of the b x b diagonal block at (k0, k0) by b one-column
// Gauss-Jordan steps (scale row k so the pivot becomes 1/p, subtract column
// k's multiple of row k from every other row).  The block, padded to 64 x 64
// with the identity (which the steps leave unchanged), lives in registers as
// 4 x 4 tiles; step k's raw row and column go through double-buffered LDS, so
// each step needs one barrier.
__global__ void __launch_bounds__(kGjT)
k_gjb_diag(int32_t n, int32_t k0, int32_t b, const double *__restrict__ a,
           double *__restrict__ pbuf) {
  __shared__ double rowb[2][kGjB], colb[2][kGjB];
  const int32_t ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  double t[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int32_t i = ty * 4 + r, j = tx * 4 + c;
      t[r][c] = (i < b && j < b) ? a[(int64_t)(k0 + i) * n + k0 + j] : (i == j ? 1.0 : 0.0);
    }
  for (int32_t k = 0; k < b; ++k) {
    const int q = k & 1;
    if (ty == (k >> 2)) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
        if (r == (k & 3))
#pragma unroll
          for (int c = 0; c < 4; ++c) rowb[q][tx * 4 + c] = t[r][c];
    }
    if (tx == (k >> 2)) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c == (k & 3))
#pragma unroll
          for (int r = 0; r < 4; ++r) colb[q][ty * 4 + r] = t[r][c];
    }
    __syncthreads();
    const double inv = 1.0 / rowb[q][k];
    double rk[4], ck[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int32_t j = tx * 4 + c;
      rk[c] = ((j == k) ? 1.0 : rowb[q][j]) * inv;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) ck[r] = colb[q][ty * 4 + r];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int32_t i = ty * 4 + r;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int32_t j = tx * 4 + c;
        t[r][c] = (i == k) ? rk[c] : ((j == k) ? 0.0 : t[r][c]) - ck[r] * rk[c];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int32_t i = ty * 4 + r, j = tx * 4 + c;
      if (i < b && j < b) pbuf[i * b + j] = t[r][c];
    }
}

// rbuf[s][j] = sum_r P[s][r] A[k0+r][j] (the row panel times P, r in
// order), and cbuf[j][s] = A[j][k0+s] (the old column panel): thread j of
// grid row g handles s = 16 g .. 16 g + 15; P is staged in LDS
__global__ void __launch_bounds__(kGjT)
k_gjb_panels(int32_t n, int32_t k0, int32_t b, const double *__restrict__ a,
             const double *__restrict__ pbuf, double *__restrict__ rbuf,
             double *__restrict__ cbuf) {
  __shared__ double ps[16][kGjB];
  const int32_t s0 = blockIdx.y * 16;
  for (int32_t e = threadIdx.x; e < 16 * kGjB; e += kGjT) {
    const int32_t s = s0 + e / kGjB, r = e % kGjB;
    ps[e / kGjB][r] = (s < b && r < b) ? pbuf[s * b + r] : 0.0;
  }
  __syncthreads();
  const int32_t j = blockIdx.x * kGjT + threadIdx.x;
  if (j >= n) return;
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = 0.0;
  int32_t r = 0;
  for (; r + 8 <= b; r += 8) {  // eight loads in flight before their products
    double x[8];
#pragma unroll
    for (int w = 0; w < 8; ++w) x[w] = a[(int64_t)(k0 + r + w) * n + j];
#pragma unroll
    for (int w = 0; w < 8; ++w)
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] += ps[u][r + w] * x[w];
  }
  for (; r < b; ++r) {
    const double x = a[(int64_t)(k0 + r) * n + j];
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] += ps[u][r] * x;
  }
#pragma unroll
  for (int u = 0; u < 16; ++u)
    if (s0 + u < b) {
      rbuf[(int64_t)(s0 + u) * n + j] = v[u];
      cbuf[(int64_t)j * b + s0 + u] = a[(int64_t)j * n + k0 + s0 + u];
    }
}

// every element from the panels, 64 x 64 output tiles (a 4 x 4 tile per
// thread), both panels staged in LDS; each element sums its b terms in
// order s = 0 .. b-1
__global__ void __launch_bounds__(kGjT)
k_gjb_update(int32_t n, int32_t k0, int32_t b, double *__restrict__ a,
             const double *__restrict__ pbuf, const double *__restrict__ rbuf,
             const double *__restrict__ cbuf) {
  __shared__ double cs[kGjB][kGjB + 1];  // [row of tile][s]
  __shared__ double rs[kGjB][kGjB + 1];  // [s][column of tile]
  const int32_t i0 = blockIdx.y * kGjB, j0 = blockIdx.x * kGjB;
  for (int32_t e = threadIdx.x; e < kGjB * kGjB; e += kGjT) {
    const int32_t u = e / kGjB, w = e % kGjB;
    const int32_t i = i0 + u, s = w;
    cs[u][w] = (i < n && s < b) ? cbuf[(int64_t)i * b + s] : 0.0;
    const int32_t j = j0 + w, s2 = u;
    double r = 0.0;
    if (j < n && s2 < b)
      r = (j >= k0 && j < k0 + b) ? pbuf[s2 * b + (j - k0)] : rbuf[(int64_t)s2 * n + j];
    rs[u][w] = r;
  }
  __syncthreads();
  const int32_t ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
  double v[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int32_t i = i0 + ty * 4 + r, j = j0 + tx * 4 + c;
      const bool jk = j >= k0 && j < k0 + b;
      v[r][c] = (i < n && j < n && !jk) ? a[(int64_t)i * n + j] : 0.0;
    }
  for (int32_t s = 0; s < b; ++s) {
    double ci[4], rj[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) ci[r] = cs[ty * 4 + r][s];
#pragma unroll
    for (int c = 0; c < 4; ++c) rj[c] = rs[s][tx * 4 + c];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int c = 0; c < 4; ++c) v[r][c] -= ci[r] * rj[c];
  }
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int32_t i = i0 + ty * 4 + r, j = j0 + tx * 4 + c;
      if (i >= n || j >= n) continue;
      const bool ik = i >= k0 && i < k0 + b, jk = j >= k0 && j < k0 + b;
      double x = v[r][c];
      if (ik && jk) x = pbuf[(i - k0) * b + (j - k0)];
      else if (ik) x = rbuf[(int64_t)(i - k0) * n + j];
      a[(int64_t)i * n + j] = x;
    }
}

size_t gj_scratch(int32_t n) { return (size_t)kGjB * kGjB + 2 * (size_t)kGjB * n + 8; }

void gj_inverse(hipStream_t s, int32_t n, double *a, double *scratch) {
  double *pbuf = scratch, *rbuf = pbuf + kGjB * kGjB, *cbuf = rbuf + (size_t)kGjB * n;
  for (int32_t k0 = 0; k0 < n; k0 += kGjB) {
    const int32_t b = n - k0 < kGjB ? n - k0 : kGjB;
    const int32_t tiles = (n + kGjB - 1) / kGjB;
    hipLaunchKernelGGL(k_gjb_diag, dim3(1), dim3(kGjT), 0, s, n, k0, b, a, pbuf);
    hipLaunchKernelGGL(k_gjb_panels, dim3(nblocks(n, kGjT), (b + 15) / 16), dim3(kGjT), 0, s, n,
                       k0, b, a, pbuf, rbuf, cbuf);
    hipLaunchKernelGGL(k_gjb_update, dim3(tiles, tiles), dim3(kGjT), 0, s, n, k0, b, a, pbuf,
                       rbuf, cbuf);
  }
}

// ------------------------------------------------------------------- FPCG

// scal layout: 0 zsum 1 zw 2 zr 3 wsum 4 rsum 5 delta 6 beta 7 alpha 8 rho 9 m 10 rr
// 12 scratch (mean of b)
//
// Each FPCG kernel writes per-block partial sums; a one-block kernel then sums
// them in block order and does the scalar arithmetic (no grid-wide atomics or
// device-scope fences in the streaming kernels).

// ---- flexible CG (Notay FCG(1)) vector kernels.  Scalars live in scal[]:
//   0..4  sums of z, z.w, z.r, w, r        5  delta = p.Ap     6  beta
//   7     alpha = rho / delta              8  rho = (z - m).r  9  m (mean of z, singular)
//   10    rr = r.r                         11 convergence flag
// The streaming kernels run kFcgBlock-thread blocks, at most kFoldBlocks of
// them (grid-stride; two blocks fill a CU), and write one partial per block;
// the next kernel of the iteration re-reduces those few partials in every
// block (same order in each block, so every block derives identical scalars)
// instead of a separate reduction launch.

int fold_blocks(int32_t n) {
  const int b = nblocks(n, kFcgBlock);
  return b < kFoldBlocks ? b : kFoldBlocks;
}

// W sums of g partials each (group c at part + c * g), returned to every
// thread of every block with the same values
template <int W>
__device__ __forceinline__ void fold_sums(const double *__restrict__ part, int32_t g,
                                          double (&t)[W], double *lds) {
#pragma unroll
  for (int c = 0; c < W; ++c) t[c] = 0.0;
  for (int32_t q = threadIdx.x; q < g; q += kFcgBlock) {
#pragma unroll
    for (int c = 0; c < W; ++c) t[c] += part[(int64_t)c * g + q];
  }
  block_sums_all<kFcgBlock, W>(t, lds);
}

// The FCG kernels' bodies take the block index blk and the block count nb of
// the launch (blockIdx.x / gridDim.x for a single solve; a grouped launch
// passes each member's own count), which fix the grid-stride row assignment
// and the reduction tree.

// w = A p (sliced-ELL layout) ; partial[b] = p.w over block b's rows
template <bool PK>
__device__ __forceinline__ void spmv_dot_body(int32_t blk, int32_t nb, int32_t n,
                                              const int32_t *__restrict__ off,
                                              const SellCols &cols,
                                              const double *__restrict__ ac,
                                              const double *__restrict__ val,
                                              const double *__restrict__ p,
                                              double *__restrict__ w,
                                              double *__restrict__ partial, double *lds) {
  double d = 0.0;
  const int32_t T = spmv_trips(n, nb), c0 = xcd_block(blk, nb) * T;
  for (int32_t t = 0; t < T; ++t) {
    const int32_t i = (c0 + t) * kFcgBlock + threadIdx.x;
    if (i >= n) break;
    const int32_t q = i >> 6;
    const int32_t base = off[q] + (i & 63), wq = (off[q + 1] - off[q]) >> 6;
    const ColSrc<PK> col(cols, q);
    const double pi = p[i];
    const double v =
        sell_row<double, false, false, ColSrc<PK>, true>(col, val, p, base, wq, ac[i] * pi, 0);
    w[i] = v;
    d += pi * v;
  }
  d = block_sum_n<kFcgBlock>(d, lds);
  if (threadIdx.x == 0) partial[blk] = d;
}

template <bool PK>
__global__ void __launch_bounds__(kFcgBlock)
k_spmv_dot(int32_t n, const int32_t *__restrict__ off, SellCols cols,
           const double *__restrict__ ac,
           const double *__restrict__ val, const double *__restrict__ p, double *__restrict__ w,
           double *__restrict__ partial) {
  __shared__ double lds[kFcgBlock / kWave];
  spmv_dot_body<PK>(blockIdx.x, gridDim.x, n, off, cols, ac, val, p, w, partial, lds);
}

void spmv_dot(hipStream_t s, int32_t n, const int32_t *off, const SellCols &cols,
              const double *ac, const double *val, const double *p, double *w, double *partial) {
  with_cols(cols, [&](auto pk) {
    hipLaunchKernelGGL((k_spmv_dot<decltype(pk)::value>), dim3(fold_blocks(n)), dim3(kFcgBlock),
                       0, s, n, off, cols, ac, val, p, w, partial);
  });
}

// partials of z, z.w, z.r, w, r: partial[c * g + b] for block b of g
template <typename V>
__device__ __forceinline__ void fcg_dots_body(int32_t blk, int32_t nb, int32_t n,
                                              const V *__restrict__ z,
                                              const double *__restrict__ w,
                                              const double *__restrict__ r,
                                              double *__restrict__ partial, double *lds) {
  double v[5] = {0, 0, 0, 0, 0};
  auto add = [&](double zi, double wi, double ri) {
    v[0] += zi;
    v[1] += zi * wi;
    v[2] += zi * ri;
    v[3] += wi;
    v[4] += ri;
  };
  const int32_t G = nb * kFcgBlock;
  int32_t i = blk * kFcgBlock + threadIdx.x;
  // four grid-stride rows per trip: their loads are in flight together, and
  // they are summed in the same order as one row per trip
  for (; i + 3 * G < n; i += 4 * G) {
    double zi[4], wi[4], ri[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      zi[u] = (double)z[i + u * G];
      wi[u] = __builtin_nontemporal_load(w + i + u * G);
      ri[u] = __builtin_nontemporal_load(r + i + u * G);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) add(zi[u], wi[u], ri[u]);
  }
  for (; i < n; i += G) add((double)z[i], w[i], r[i]);
  block_sums_all<kFcgBlock, 5>(v, lds);
  if (threadIdx.x == 0)
    for (int c = 0; c < 5; ++c) partial[(int64_t)c * nb + blk] = v[c];
}

template <typename V>
__global__ void __launch_bounds__(kFcgBlock)
k_fcg_dots(int32_t n, const V *__restrict__ z, const double *__restrict__ w,
           const double *__restrict__ r, double *__restrict__ partial, BatchStrides bs) {
  __shared__ double lds[5 * kFcgBlock / kWave];
  // right-hand side blockIdx.y of a batch
  z += blockIdx.y * bs.ldv;
  w += blockIdx.y * bs.ld;
  r += blockIdx.y * bs.ld;
  partial += blockIdx.y * bs.ps;
  fcg_dots_body<V>(blockIdx.x, gridDim.x, n, z, w, r, partial, lds);
}

template <typename V>
void fcg_dots(hipStream_t s, int32_t n, const V *z, const double *w, const double *r,
              double *partial, int k, BatchStrides bs) {
  hipLaunchKernelGGL(k_fcg_dots<V>, dim3(fold_blocks(n), k), dim3(kFcgBlock), 0, s, n, z, w, r,
                     partial, bs);
}

// Folds the z-dots (m = mean of z when singular, rho = (z - m).r,
// beta = (z - m).w / delta_prev), then p = (z - m) - beta p (p = z - m on the
// first iteration).  Block 0 stores the scalars.
template <typename V>
__device__ __forceinline__ void fcg_update_p_body(int32_t blk, int32_t nb, int32_t n,
                                                  const V *__restrict__ z,
                                                  const double *__restrict__ partial, int32_t g,
                                                  int singular, int first,
                                                  double *__restrict__ scal,
                                                  double *__restrict__ p, double *lds) {
  const int32_t G = nb * kFcgBlock;
  int32_t i = blk * kFcgBlock + threadIdx.x;
  // the first four grid-stride rows are loaded before the fold, so that their
  // latency overlaps the partials' (the scalars are not needed to load them)
  double zi[4], pi[4];
  const bool ahead = i + 3 * G < n;
  if (ahead) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      zi[u] = (double)z[i + u * G];
      pi[u] = first ? 0.0 : p[i + u * G];
    }
  }
  double t[5];
  fold_sums<5>(partial, g, t, lds);
  const double m = singular ? t[0] / (double)n : 0.0;
  const double rho = t[2] - m * t[4];
  const double beta = first ? 0.0 : (t[1] - m * t[3]) / scal[5];
  if (blk == 0 && threadIdx.x == 0) {
    for (int c = 0; c < 5; ++c) scal[c] = t[c];
    scal[9] = m;
    scal[8] = rho;
    scal[6] = beta;
  }
  auto put = [&](int32_t j, double zj, double pj) {
    const double d = zj - m;
    p[j] = first ? d : d - beta * pj;
  };
  if (ahead) {
#pragma unroll
    for (int u = 0; u < 4; ++u) put(i + u * G, zi[u], pi[u]);
    i += 4 * G;
  }
  for (; i + 3 * G < n; i += 4 * G) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      zi[u] = (double)z[i + u * G];
      pi[u] = first ? 0.0 : p[i + u * G];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) put(i + u * G, zi[u], pi[u]);
  }
  for (; i < n; i += G) {
    const double zi = (double)z[i] - m;
    p[i] = first ? zi : zi - beta * p[i];
  }
}

template <typename V>
__global__ void __launch_bounds__(kFcgBlock)
k_fcg_update_p(int32_t n, const V *__restrict__ z, const double *__restrict__ partial,
               int32_t g, int singular, int first, double *__restrict__ scal,
               double *__restrict__ p, BatchStrides bs) {
  __shared__ double lds[5 * kFcgBlock / kWave];
  z += blockIdx.y * bs.ldv;
  p += blockIdx.y * bs.ld;
  partial += blockIdx.y * bs.ps;
  scal += blockIdx.y * bs.ss;
  fcg_update_p_body<V>(blockIdx.x, gridDim.x, n, z, partial, g, singular, first, scal, p, lds);
}

template <typename V>
void fcg_update_p(hipStream_t s, int32_t n, const V *z, const double *partial, int singular,
                  int first, double *scal, double *p, int k, BatchStrides bs) {
  const int g = fold_blocks(n);
  hipLaunchKernelGGL(k_fcg_update_p<V>, dim3(g, k), dim3(kFcgBlock), 0, s, n, z, partial, g,
                     singular, first, scal, p, bs);
}

// Folds delta = p.Ap (partials of k_spmv_dot) and alpha = rho / delta, then
// x += alpha p ; r -= alpha w ; rv = (V) r ; partial_out[b] = r.r.  Rows below
// n0 (colour class 0 of the finest level) also get the next V-cycle's first
// Gauss-Seidel pass from zero, z0 = rv / acv: their neighbours all lie in
// later colours, which that pass reads as 0.  Nothing changes once the
// convergence flag scal[11] is set (an iteration issued after the converged
// one leaves x and r as they were).  partial_out must not overlap partial_in.
template <typename V>
__device__ __forceinline__ void fcg_update_xr_body(
    int32_t blk, int32_t nb, int32_t n, const double *__restrict__ partial_in,
    double *__restrict__ scal, const double *__restrict__ p, const double *__restrict__ w,
    double *__restrict__ x, double *__restrict__ r, V *__restrict__ rv,
    double *__restrict__ partial_out, int32_t n0, const V *__restrict__ acv, V *__restrict__ z0,
    double *lds) {
  const int32_t G = nb * kFcgBlock;
  int32_t i = blk * kFcgBlock + threadIdx.x;
  // the first four grid-stride rows are loaded before the fold (their latency
  // overlaps the partials')
  double xi[4], pi[4], ri[4], wi[4];
  const bool ahead = i + 3 * G < n;
  if (ahead) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xi[u] = __builtin_nontemporal_load(x + i + u * G);
      pi[u] = __builtin_nontemporal_load(p + i + u * G);
      ri[u] = __builtin_nontemporal_load(r + i + u * G);
      wi[u] = __builtin_nontemporal_load(w + i + u * G);
    }
  }
  double t[1];
  fold_sums<1>(partial_in, nb, t, lds);
  const double delta = t[0];
  const double a = scal[8] / delta;
  if (blk == 0 && threadIdx.x == 0) {
    scal[5] = delta;
    scal[7] = a;
  }
  double rr = 0.0;
  auto row = [&](int32_t i, double xi, double pi, double ri0, double wi) {
    __builtin_nontemporal_store(xi + a * pi, x + i);
    const double ri = ri0 - a * wi;
    __builtin_nontemporal_store(ri, r + i);
    const V rvi = (V)ri;
    rv[i] = rvi;
    if (i < n0) z0[i] = rvi / acv[i];
    rr += ri * ri;
  };
  if (ahead) {
#pragma unroll
    for (int u = 0; u < 4; ++u) row(i + u * G, xi[u], pi[u], ri[u], wi[u]);
    i += 4 * G;
  }
  for (; i + 3 * G < n; i += 4 * G) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      xi[u] = __builtin_nontemporal_load(x + i + u * G);
      pi[u] = __builtin_nontemporal_load(p + i + u * G);
      ri[u] = __builtin_nontemporal_load(r + i + u * G);
      wi[u] = __builtin_nontemporal_load(w + i + u * G);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) row(i + u * G, xi[u], pi[u], ri[u], wi[u]);
  }
  for (; i < n; i += G) row(i, x[i], p[i], r[i], w[i]);
  rr = block_sum_n<kFcgBlock>(rr, lds);
  if (threadIdx.x == 0) partial_out[blk] = rr;
}

template <typename V>
__global__ void __launch_bounds__(kFcgBlock)
k_fcg_update_xr(int32_t n, const double *__restrict__ partial_in, double *__restrict__ scal,
                const double *__restrict__ p, const double *__restrict__ w,
                double *__restrict__ x, double *__restrict__ r, V *__restrict__ rv,
                double *__restrict__ partial_out, int32_t n0, const V *__restrict__ acv,
                V *__restrict__ z0, BatchStrides bs) {
  __shared__ double lds[kFcgBlock / kWave];
  scal += blockIdx.y * bs.ss;
  if (scal[11] != 0.0) return;
  partial_in += blockIdx.y * bs.ps;
  partial_out += blockIdx.y * bs.ps;
  p += blockIdx.y * bs.ld;
  w += blockIdx.y * bs.ld;
  x += blockIdx.y * bs.ld;
  r += blockIdx.y * bs.ld;
  rv += blockIdx.y * bs.ldv;
  z0 += blockIdx.y * bs.ldv;
  fcg_update_xr_body<V>(blockIdx.x, gridDim.x, n, partial_in, scal, p, w, x, r, rv, partial_out,
                        n0, acv, z0, lds);
}

template <typename V>
__global__ void k_gather_vals(int64_t n, const int32_t *__restrict__ idx,
                              const double *__restrict__ src, V *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t j = idx[i];
  out[i] = j >= 0 ? (V)src[j] : V(0);
}

template <typename V>
void gather_vals(hipStream_t s, int64_t n, const int32_t *idx, const double *src, V *out) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_vals<V>, dim3(nblocks(n)), dim3(kBlock), 0, s, n, idx, src, out);
}
template void gather_vals<float>(hipStream_t, int64_t, const int32_t *, const double *, float *);
template void gather_vals<double>(hipStream_t, int64_t, const int32_t *, const double *,
                                  double *);

__global__ void k_ring_put(const double *__restrict__ src, double *ring) {
  if (threadIdx.x == 0) {
    ring[0] = src[0];
    ring[1] = -1.0;
  }
}

void ring_put(hipStream_t s, const double *src, double *ring) {
  hipLaunchKernelGGL(k_ring_put, dim3(1), dim3(64), 0, s, src, ring);
}

// rr = r.r from the update's partials; sets the convergence flag scal[11]
// once rr <= scal[13], after which scal[10] is left alone; writes {rr, it} to
// ring slot it & 1 (host-visible, may be null), it = scal[14] the iteration
// counter, which it then advances.  Everything per solve lives in scal, so a
// captured iteration replays unchanged.
__device__ __forceinline__ void fcg_rr_body(const double *__restrict__ partial, int32_t g,
                                            double *scal, double *ring, double *lds) {
  double t[1];
  fold_sums<1>(partial, g, t, lds);
  if (threadIdx.x != 0) return;
  if (scal[11] == 0.0) {
    scal[10] = t[0];
    if (t[0] <= scal[13]) scal[11] = 1.0;
  }
  const double it = scal[14];
  if (ring) {
    double *slot = ring + 2 * ((int64_t)it & 1);
    slot[0] = scal[10];
    slot[1] = it;
  }
  scal[14] = it + 1.0;
}

__global__ void __launch_bounds__(kFcgBlock)
k_fcg_rr(const double *__restrict__ partial, int32_t g, double *scal, double *ring,
         BatchStrides bs) {
  __shared__ double lds[kFcgBlock / kWave];
  partial += blockIdx.y * bs.ps;
  scal += blockIdx.y * bs.ss;
  if (ring) ring += blockIdx.y * bs.rs;
  fcg_rr_body(partial, g, scal, ring, lds);
}

// arm a solve: convergence flag off (on: done, for a zero right-hand side in
// a batch), threshold thr, iteration counter 0
__global__ void k_fcg_arm(double *scal, double thr, int done) {
  if (threadIdx.x == 0) {
    scal[11] = done ? 1.0 : 0.0;
    scal[13] = thr;
    scal[14] = 0.0;
  }
}

void fcg_arm(hipStream_t s, double *scal, double thr, int done) {
  hipLaunchKernelGGL(k_fcg_arm, dim3(1), dim3(64), 0, s, scal, thr, done);
}

template <typename V>
void fcg_update_xr(hipStream_t s, int32_t n, double *scal, const double *p, const double *w,
                   double *x, double *r, V *rv, double *partial, double *ring, int32_t n0,
                   const V *acv, V *z0, int k, BatchStrides bs) {
  // the operator-product partials occupy partial[0, g); the r.r partials go
  // after them
  const int g = fold_blocks(n);
  hipLaunchKernelGGL(k_fcg_update_xr<V>, dim3(g, k), dim3(kFcgBlock), 0, s, n, partial, scal, p,
                     w, x, r, rv, partial + g, n0, acv, z0, bs);
  hipLaunchKernelGGL(k_fcg_rr, dim3(1, k), dim3(kFcgBlock), 0, s, partial + g, g, scal, ring, bs);
}

__global__ void __launch_bounds__(kBlock)
k_dot_partial(int32_t n, const double *__restrict__ a, const double *__restrict__ b,
              double *__restrict__ partial) {
  __shared__ double lds[kBlock / kWave];
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double v = (i < n) ? a[i] * b[i] : 0.0;
  v = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

void dot_partial(hipStream_t s, int32_t n, const double *a, const double *b, double *partial) {
  hipLaunchKernelGGL(k_dot_partial, dim3(nblocks(n)), dim3(kBlock), 0, s, n, a, b, partial);
}

__global__ void __launch_bounds__(kBlock)
k_sum_partial(int32_t n, const double *__restrict__ a, double *__restrict__ partial) {
  __shared__ double lds[kBlock / kWave];
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double v = (i < n) ? a[i] : 0.0;
  v = block_sum(v, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = v;
}

void sum_partial(hipStream_t s, int32_t n, const double *a, double *partial) {
  hipLaunchKernelGGL(k_sum_partial, dim3(nblocks(n)), dim3(kBlock), 0, s, n, a, partial);
}

__global__ void k_axpy_const(int32_t n, const double *__restrict__ slot, double mul,
                             double *__restrict__ x) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) x[i] = x[i] + mul * slot[0];
}

void axpy_const(hipStream_t s, int32_t n, const double *scal_slot, double mul, double *x) {
  hipLaunchKernelGGL(k_axpy_const, dim3(nblocks(n)), dim3(kBlock), 0, s, n, scal_slot, mul, x);
}

// ---------------------------------------------------------- grouped launches

template <typename V, GOp OP>
__global__ void __launch_bounds__(kBlock) k_group_row(const GArgs<V> *__restrict__ tab) {
  const GArgs<V> a = tab[blockIdx.y];
  if ((int32_t)blockIdx.x >= a.nb || *a.done != 0.0) return;
  if (OP == G_MCGS) {
    // the member's rows in k_mcgs_sell's XCD-aware order over its own blocks
    // (each member's column form is uniform over its blocks)
    const int32_t i = (a.r0 & ~63) + xcd_block(blockIdx.x, a.nb) * kBlock + threadIdx.x;
    if (i < a.r0 || i >= a.n) return;
    if (a.cols.d16) mcgs_row<V, true>(i, a.n, a.off, a.cols, a.val, a.ac, a.b, a.out, a.aux);
    else mcgs_row<V, false>(i, a.n, a.off, a.cols, a.val, a.ac, a.b, a.out, a.aux);
  } else if (OP == G_RESID) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    if (i < a.aux) a.out[i] = V(0);
    else if (a.cols.d16 && a.wide)
      residual_row<V, true, true>(i, a.off, a.cols, a.val, a.ac, a.b, a.x, a.out);
    else if (a.cols.d16) residual_row<V, true>(i, a.off, a.cols, a.val, a.ac, a.b, a.x, a.out);
    else if (a.wide) residual_row<V, false, true>(i, a.off, a.cols, a.val, a.ac, a.b, a.x, a.out);
    else residual_row<V, false>(i, a.off, a.cols, a.val, a.ac, a.b, a.x, a.out);
  } else if (OP == G_JACOBI) {
    const int32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    // the member's rows, PRO (prolongation fused) and column form, wide or in runs
    auto row = [&](auto pro, auto pk, auto wide) {
      constexpr bool PRO = decltype(pro)::value, PK = decltype(pk)::value;
      jacobi_row<V, PRO, PK, decltype(wide)::value>(i, a.off, a.cols, a.val, a.ac, a.b, a.x,
                                                    PRO ? a.parent : nullptr,
                                                    PRO ? a.xc : nullptr, a.out, a.omega);
    };
    using T = std::true_type;
    using F = std::false_type;
    if (a.parent) {
      if (a.cols.d16) a.wide ? row(T{}, T{}, T{}) : row(T{}, T{}, F{});
      else a.wide ? row(T{}, F{}, T{}) : row(T{}, F{}, F{});
    } else {
      if (a.cols.d16) a.wide ? row(F{}, T{}, T{}) : row(F{}, T{}, F{});
      else a.wide ? row(F{}, F{}, T{}) : row(F{}, F{}, F{});
    }
  } else if (OP == G_RESTRICT) {
    const int32_t p = blockIdx.x * kBlock + threadIdx.x;
    if (p >= a.n) return;
    if (a.ac)
      restrict_row<V, true>(p, a.off, a.child, a.x, a.ac, a.out, a.out2, a.omega, a.scale);
    else
      restrict_row<V, false>(p, a.off, a.child, a.x, nullptr, a.out, nullptr, V(0), a.scale);
  } else if (OP == G_PROLONG) {
    const int32_t i = a.r0 + blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    a.out[i] = a.out[i] + a.xc[a.parent[i]];
  } else {
    const int32_t row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
    if (row >= a.n) return;
    dense_row<V>(row, threadIdx.x & (kWave - 1), a.n, a.ainv, a.b, a.out);
  }
}

template <typename V, GFcgOp OP>
__global__ void __launch_bounds__(kFcgBlock) k_group_fcg(const GFcgArgs<V> *__restrict__ tab) {
  __shared__ double lds[(OP == G_DOTS || OP == G_UPDATE_P || OP == G_UPDATE_P_FIRST ? 5 : 1) *
                        kFcgBlock / kWave];
  const GFcgArgs<V> a = tab[blockIdx.y];
  if ((int32_t)blockIdx.x >= a.nb || a.scal[11] != 0.0) return;
  if (OP == G_DOTS) {
    fcg_dots_body<V>(blockIdx.x, a.nb, a.n, a.z, a.w, a.r, a.partial, lds);
  } else if (OP == G_UPDATE_P || OP == G_UPDATE_P_FIRST) {
    fcg_update_p_body<V>(blockIdx.x, a.nb, a.n, a.z, a.partial, a.nb, a.singular,
                         OP == G_UPDATE_P_FIRST, a.scal, a.p, lds);
  } else if (OP == G_SPMV) {
    if (a.cols.d16)
      spmv_dot_body<true>(blockIdx.x, a.nb, a.n, a.off, a.cols, a.ac, a.val, a.p, a.w, a.partial,
                          lds);
    else
      spmv_dot_body<false>(blockIdx.x, a.nb, a.n, a.off, a.cols, a.ac, a.val, a.p, a.w,
                           a.partial, lds);
  } else if (OP == G_UPDATE_XR) {
    // the operator-product partials occupy partial[0, nb); r.r's follow
    fcg_update_xr_body<V>(blockIdx.x, a.nb, a.n, a.partial, a.scal, a.p, a.w, a.x, a.r, a.rv,
                          a.partial + a.nb, a.n0, a.acv, a.z0, lds);
  } else {
    fcg_rr_body(a.partial + a.nb, a.nb, a.scal, a.ring, lds);
  }
}

template <typename V>
void group_row(hipStream_t s, GOp op, int32_t gx, int32_t g, const GArgs<V> *tab) {
  if (gx <= 0 || g <= 0) return;
  const dim3 grid(gx, g), blk(kBlock);
  switch (op) {
    case G_MCGS: hipLaunchKernelGGL((k_group_row<V, G_MCGS>), grid, blk, 0, s, tab); break;
    case G_RESID: hipLaunchKernelGGL((k_group_row<V, G_RESID>), grid, blk, 0, s, tab); break;
    case G_JACOBI: hipLaunchKernelGGL((k_group_row<V, G_JACOBI>), grid, blk, 0, s, tab); break;
    case G_RESTRICT: hipLaunchKernelGGL((k_group_row<V, G_RESTRICT>), grid, blk, 0, s, tab); break;
    case G_PROLONG: hipLaunchKernelGGL((k_group_row<V, G_PROLONG>), grid, blk, 0, s, tab); break;
    case G_DENSE: hipLaunchKernelGGL((k_group_row<V, G_DENSE>), grid, blk, 0, s, tab); break;
  }
}

template <typename V>
void group_fcg(hipStream_t s, GFcgOp op, int32_t gx, int32_t g, const GFcgArgs<V> *tab) {
  if (gx <= 0 || g <= 0) return;
  const dim3 grid(op == G_RR ? 1 : gx, g), blk(kFcgBlock);
  switch (op) {
    case G_DOTS: hipLaunchKernelGGL((k_group_fcg<V, G_DOTS>), grid, blk, 0, s, tab); break;
    case G_UPDATE_P: hipLaunchKernelGGL((k_group_fcg<V, G_UPDATE_P>), grid, blk, 0, s, tab); break;
    case G_UPDATE_P_FIRST:
      hipLaunchKernelGGL((k_group_fcg<V, G_UPDATE_P_FIRST>), grid, blk, 0, s, tab);
      break;
    case G_SPMV: hipLaunchKernelGGL((k_group_fcg<V, G_SPMV>), grid, blk, 0, s, tab); break;
    case G_UPDATE_XR: hipLaunchKernelGGL((k_group_fcg<V, G_UPDATE_XR>), grid, blk, 0, s, tab); break;
    case G_RR: hipLaunchKernelGGL((k_group_fcg<V, G_RR>), grid, blk, 0, s, tab); break;
  }
}

template void group_row<float>(hipStream_t, GOp, int32_t, int32_t, const GArgs<float> *);
template void group_row<double>(hipStream_t, GOp, int32_t, int32_t, const GArgs<double> *);
template void group_fcg<float>(hipStream_t, GFcgOp, int32_t, int32_t, const GFcgArgs<float> *);
template void group_fcg<double>(hipStream_t, GFcgOp, int32_t, int32_t, const GFcgArgs<double> *);

// explicit instantiations
#define UFVM_INST(V)                                                                            \
  template void bmcgs<V>(hipStream_t, int32_t, const int32_t *, const int32_t *, int32_t,       \
                         const int32_t *, const int32_t *, const V *, const V *, const V *,     \
                         const V *, const int32_t *, const V *, V *, int, int);                 \
  template void mcgs_sell<V>(hipStream_t, int32_t, int32_t, const int32_t *, const SellCols &,   \
                             const V *, const V *, const V *, V *, int32_t);                    \
  template void residual_sell<V>(hipStream_t, int32_t, int32_t, const int32_t *,                \
                                 const SellCols &, const V *, const V *, const V *, const V *,  \
                                 V *, bool);                                                    \
  template void jacobi_sell<V>(hipStream_t, int32_t, const int32_t *, const SellCols &,         \
                               const V *, const V *, const V *, const V *, const int32_t *,     \
                               const V *, V *, V, bool);                                        \
  template void restrict_jacobi<V>(hipStream_t, int32_t, const int32_t *, const int32_t *,      \
                                   const V *, const V *, V *, V *, V, V);                      \
  template void residual_fast<V>(hipStream_t, int32_t, const int32_t *, const int32_t *,        \
                                 const V *, const V *, const V *, const V *, V *);              \
  template void restrict_fast<V>(hipStream_t, int32_t, const int32_t *, const int32_t *,        \
                                 const V *, V *, V);                                            \
  template void prolong_fast<V>(hipStream_t, int32_t, int32_t, const int32_t *, const V *, V *); \
  template void dense_solve<V>(hipStream_t, int32_t, const double *, const V *, V *);           \
  template void fill<V>(hipStream_t, int64_t, V *, V);                                          \
  template void fcg_dots<V>(hipStream_t, int32_t, const V *, const double *, const double *,    \
                            double *, int, BatchStrides);                                       \
  template void fcg_update_p<V>(hipStream_t, int32_t, const V *, const double *, int, int,       \
                                double *, double *, int, BatchStrides);                         \
  template void fcg_update_xr<V>(hipStream_t, int32_t, double *, const double *, const double *, \
                                 double *, double *, V *, double *, double *, int32_t,          \
                                 const V *, V *, int, BatchStrides);
UFVM_INST(float)
UFVM_INST(double)
template void convert<double, float>(hipStream_t, int64_t, const double *, float *);
template void convert<float, double>(hipStream_t, int64_t, const float *, double *);
template void convert<double, double>(hipStream_t, int64_t, const double *, double *);
template void convert<float, float>(hipStream_t, int64_t, const float *, float *);

}  // namespace k
}  // namespace ufvm
