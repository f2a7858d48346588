// kernels_batch.hip -- multi-right-hand-side forms of the precision-form
// V-cycle kernels and of the FP64 operator product.
//
// K right-hand sides of one system are solved together: their vectors are
// stored one after another (RHS-major, vector r at offset r * ld, ld a
// multiple of 64), and every kernel runs one thread per row that loads the
// row's sliced-ELL entries once and applies them to all K vectors.  The
// matrix, the dominant HBM stream of every pass, is read once per K solves.
// Each row of each vector is computed with exactly the arithmetic, in exactly
// the order, of the single-vector kernel, so a batched solve gives each right-
// hand side the bits of its own solve.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "reduce.h"

namespace ufvm {
namespace k {

// v[r] -= sum over the N entries from slot e0 of val x_r[col] in entry order,
// the run's loads issued together (the runs of the single-vector sell_row);
// ZA: neighbours with index >= zero_above read as 0
template <int N, typename V, int K, bool ZA, typename CS>
__device__ __forceinline__ void sell_run_k(const CS &col, const V *__restrict__ val, const V *x,
                                           int64_t ld, int32_t e0, V (&v)[K],
                                           int32_t zero_above) {
  int32_t j[N];
  V a[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = col[e0 + u * 64];
    a[u] = val[e0 + u * 64];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    V xj[N];
#pragma unroll
    for (int u = 0; u < N; ++u) xj[u] = x[r * ld + j[u]];
    if (ZA) {
#pragma unroll
      for (int u = 0; u < N; ++u) asm volatile("" : "+v"(xj[u]));
#pragma unroll
      for (int u = 0; u < N; ++u) xj[u] = j[u] < zero_above ? xj[u] : V(0);
    }
#pragma unroll
    for (int u = 0; u < N; ++u) v[r] = v[r] - a[u] * xj[u];
  }
}

template <typename V, int K, bool ZA, typename CS>
__device__ __forceinline__ void sell_rows_k(const CS &col, const V *__restrict__ val,
                                            const V *x, int64_t ld, int32_t base, int32_t w,
                                            V (&v)[K], int32_t zero_above) {
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun)
    sell_run_k<kSellRun, V, K, ZA>(col, val, x, ld, base + k * 64, v, zero_above);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: sell_run_k<1, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    case 2: sell_run_k<2, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    case 3: sell_run_k<3, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    case 4: sell_run_k<4, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    case 5: sell_run_k<5, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    case 6: sell_run_k<6, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    case 7: sell_run_k<7, V, K, ZA>(col, val, x, ld, e0, v, zero_above); break;
    default: break;
  }
}

// Gauss-Seidel pass over the colour class [r0, r1) (k_mcgs_sell for K vectors)
template <typename V, int K, bool PK>
__global__ void __launch_bounds__(kBlock)
k_mcgs_sell_k(int32_t r0, int32_t r1, const int32_t *__restrict__ off, SellCols cols,
              const V *__restrict__ val, const V *__restrict__ ac, const V *__restrict__ b, V *x,
              int64_t ld, int32_t zero_above) {
  const int32_t i = (r0 & ~63) + xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
  if (i < r0 || i >= r1) return;
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  V v[K];
#pragma unroll
  for (int r = 0; r < K; ++r) v[r] = b[r * ld + i];
  if (zero_above <= r1) sell_rows_k<V, K, true>(col, val, x, ld, base, w, v, zero_above);
  else sell_rows_k<V, K, false>(col, val, x, ld, base, w, v, zero_above);
  const V d = ac[i];
#pragma unroll
  for (int r = 0; r < K; ++r) x[r * ld + i] = v[r] / d;
}

template <typename V, int K, bool PK>
__global__ void __launch_bounds__(kBlock)
k_residual_sell_k(int32_t n, int32_t zero_rows, const int32_t *__restrict__ off, SellCols cols,
                  const V *__restrict__ val, const V *__restrict__ ac, const V *__restrict__ b, const V *__restrict__ x,
                  V *__restrict__ rr, int64_t ld) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  if (i < zero_rows) {
#pragma unroll
    for (int r = 0; r < K; ++r) rr[r * ld + i] = V(0);
    return;
  }
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  const V d = ac[i];
  V v[K];
#pragma unroll
  for (int r = 0; r < K; ++r) v[r] = b[r * ld + i] - d * x[r * ld + i];
  sell_rows_k<V, K, false>(col, val, x, ld, base, w, v, 0);
#pragma unroll
  for (int r = 0; r < K; ++r) rr[r * ld + i] = v[r];
}

// v[r] -= sum over the N entries from slot e0 of val (x_r + xc_r[parent] when
// PRO), the run's loads issued together
template <int N, typename V, bool PRO, int K, typename CS>
__device__ __forceinline__ void jacobi_run_k(const CS &col, const V *__restrict__ val,
                                             const V *__restrict__ x, int64_t ld,
                                             const int32_t *__restrict__ parent,
                                             const V *__restrict__ xc, int64_t ldc, int32_t e0,
                                             V (&v)[K]) {
  int32_t j[N], pj[N];
  V a[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = col[e0 + u * 64];
    a[u] = val[e0 + u * 64];
  }
  if (PRO) {
#pragma unroll
    for (int u = 0; u < N; ++u) pj[u] = parent[j[u]];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    V xj[N];
#pragma unroll
    for (int u = 0; u < N; ++u)
      xj[u] = PRO ? x[r * ld + j[u]] + xc[r * ldc + pj[u]] : x[r * ld + j[u]];
#pragma unroll
    for (int u = 0; u < N; ++u) v[r] = v[r] - a[u] * xj[u];
  }
}

// damped Jacobi for K vectors; PRO: x + xc[parent] as the sweep's input
template <typename V, bool PRO, int K, bool PK>
__global__ void __launch_bounds__(kBlock)
k_jacobi_sell_k(int32_t n, const int32_t *__restrict__ off, SellCols cols,
                const V *__restrict__ val, const V *__restrict__ ac, const V *__restrict__ b,
                const V *__restrict__ x, int64_t ld, const int32_t *__restrict__ parent,
                const V *__restrict__ xc, int64_t ldc, V *__restrict__ out, V omega) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t q = i >> 6;
  const int32_t base = off[q] + (i & 63), w = (off[q + 1] - off[q]) >> 6;
  const ColSrc<PK> col(cols, q);
  const V d = ac[i];
  const int32_t pi = PRO ? parent[i] : 0;
  V vi[K], v[K];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    vi[r] = PRO ? x[r * ld + i] + xc[r * ldc + pi] : x[r * ld + i];
    v[r] = b[r * ld + i] - d * vi[r];
  }
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun)
    jacobi_run_k<kSellRun, V, PRO, K>(col, val, x, ld, parent, xc, ldc, base + k * 64, v);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: jacobi_run_k<1, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    case 2: jacobi_run_k<2, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    case 3: jacobi_run_k<3, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    case 4: jacobi_run_k<4, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    case 5: jacobi_run_k<5, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    case 6: jacobi_run_k<6, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    case 7: jacobi_run_k<7, V, PRO, K>(col, val, x, ld, parent, xc, ldc, e0, v); break;
    default: break;
  }
#pragma unroll
  for (int r = 0; r < K; ++r) out[r * ld + i] = vi[r] + omega * v[r] / d;
}

// s[r] += the residuals r of the N child slots from e0 in list order (padding
// -1 adds 0), the run's child loads issued together
template <int N, typename V, int K>
__device__ __forceinline__ void restrict_run_k(const int32_t *__restrict__ child,
                                               const V *__restrict__ rf, int64_t ldf, int32_t e0,
                                               V (&s)[K]) {
  int32_t c[N];
#pragma unroll
  for (int u = 0; u < N; ++u) c[u] = child[e0 + u * 64];
#pragma unroll
  for (int r = 0; r < K; ++r) {
    V v[N];
#pragma unroll
    for (int u = 0; u < N; ++u) v[u] = rf[r * ldf + (c[u] >= 0 ? c[u] : 0)];
#pragma unroll
    for (int u = 0; u < N; ++u) asm volatile("" : "+v"(v[u]));
#pragma unroll
    for (int u = 0; u < N; ++u) s[r] = s[r] + (c[u] >= 0 ? v[u] : V(0));
  }
}

// restriction over sliced-ELL child lists (k_restrict_sell for K vectors)
template <typename V, bool JAC, int K>
__global__ void __launch_bounds__(kBlock)
k_restrict_sell_k(int32_t nc, const int32_t *__restrict__ off, const int32_t *__restrict__ child,
                  const V *__restrict__ rf, int64_t ldf, const V *__restrict__ acc,
                  V *__restrict__ bc, V *__restrict__ xc, int64_t ldc, V omega, V scale) {
  const int32_t p = blockIdx.x * kBlock + threadIdx.x;
  if (p >= nc) return;
  const int32_t q = p >> 6;
  const int32_t base = off[q] + (p & 63), w = (off[q + 1] - off[q]) >> 6;
  V s[K];
#pragma unroll
  for (int r = 0; r < K; ++r) s[r] = V(0);
  int32_t k = 0;
  for (; k + kSellRun <= w; k += kSellRun)
    restrict_run_k<kSellRun, V, K>(child, rf, ldf, base + k * 64, s);
  const int32_t e0 = base + k * 64;
  switch (w - k) {
    case 1: restrict_run_k<1, V, K>(child, rf, ldf, e0, s); break;
    case 2: restrict_run_k<2, V, K>(child, rf, ldf, e0, s); break;
    case 3: restrict_run_k<3, V, K>(child, rf, ldf, e0, s); break;
    case 4: restrict_run_k<4, V, K>(child, rf, ldf, e0, s); break;
    case 5: restrict_run_k<5, V, K>(child, rf, ldf, e0, s); break;
    case 6: restrict_run_k<6, V, K>(child, rf, ldf, e0, s); break;
    case 7: restrict_run_k<7, V, K>(child, rf, ldf, e0, s); break;
    default: break;
  }
  const V a = JAC ? acc[p] : V(1);
#pragma unroll
  for (int r = 0; r < K; ++r) {
    const V t = scale * s[r];
    bc[r * ldc + p] = t;
    if (JAC) xc[r * ldc + p] = omega * t / a;
  }
}

template <typename V, int K>
__global__ void __launch_bounds__(kBlock)
k_prolong_from_k(int32_t r0, int32_t n, const int32_t *__restrict__ parent,
                 const V *__restrict__ xc, int64_t ldc, V *__restrict__ x, int64_t ld) {
  const int32_t i = r0 + blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t p = parent[i];
#pragma unroll
  for (int r = 0; r < K; ++r) x[r * ld + i] = x[r * ld + i] + xc[r * ldc + p];
}

// x_r = Ainv b_r, one wavefront per row, FP64 accumulation in the order of
// k_dense_gemv (four partial sums per lane, then the wavefront sum)
template <typename V, int K>
__global__ void __launch_bounds__(kBlock)
k_dense_gemv_k(int32_t n, const double *__restrict__ ainv, const V *__restrict__ b, int64_t ld,
               V *__restrict__ x) {
  const int32_t row = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  const int lane = threadIdx.x & (kWave - 1);
  if (row >= n) return;
  const double *a = ainv + (int64_t)row * n;
  double s0[K], s1[K], s2[K], s3[K];
#pragma unroll
  for (int r = 0; r < K; ++r) s0[r] = s1[r] = s2[r] = s3[r] = 0.0;
  int32_t j = lane;
  for (; j + 3 * kWave < n; j += 4 * kWave) {
    const double a0 = a[j], a1 = a[j + kWave], a2 = a[j + 2 * kWave], a3 = a[j + 3 * kWave];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      s0[r] += a0 * (double)b[r * ld + j];
      s1[r] += a1 * (double)b[r * ld + j + kWave];
      s2[r] += a2 * (double)b[r * ld + j + 2 * kWave];
      s3[r] += a3 * (double)b[r * ld + j + 3 * kWave];
    }
  }
  for (; j < n; j += kWave) {
    const double aj = a[j];
#pragma unroll
    for (int r = 0; r < K; ++r) s0[r] += aj * (double)b[r * ld + j];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    double s = (s0[r] + s1[r]) + (s2[r] + s3[r]);
    s = wave_sum(s);
    if (lane == 0) x[r * ld + row] = (V)s;
  }
}

// v[r] += sum over the N entries from slot e0 of val p_r[col], the run's loads
// issued together
template <int N, int K, typename CS>
__device__ __forceinline__ void spmv_run_k(const CS &col, const double *__restrict__ val,
                                           const double *__restrict__ p, int64_t ld, int32_t e0,
                                           double (&v)[K]) {
  int32_t j[N];
  double a[N];
#pragma unroll
  for (int u = 0; u < N; ++u) {
    j[u] = col[e0 + u * 64];
    a[u] = val[e0 + u * 64];
  }
#pragma unroll
  for (int r = 0; r < K; ++r) {
    double xj[N];
#pragma unroll
    for (int u = 0; u < N; ++u) xj[u] = p[r * ld + j[u]];
#pragma unroll
    for (int u = 0; u < N; ++u) v[r] = v[r] + a[u] * xj[u];
  }
}

// w_r = A p_r (FP64, sliced-ELL) and partial[r * pstride + block] = p_r.w_r
// over the block's rows: the grid-stride row assignment and the block sum of
// k_spmv_dot, per vector
template <int K, bool PK>
__global__ void __launch_bounds__(kFcgBlock)
k_spmv_dot_k(int32_t n, const int32_t *__restrict__ off, SellCols cols,
             const double *__restrict__ ac, const double *__restrict__ val,
             const double *__restrict__ p, double *__restrict__ w, int64_t ld,
             double *__restrict__ partial, int64_t pstride) {
  __shared__ double lds[K * kFcgBlock / kWave];
  double d[K];
#pragma unroll
  for (int r = 0; r < K; ++r) d[r] = 0.0;
  const int32_t T = spmv_trips(n, gridDim.x), c0 = xcd_block(blockIdx.x, gridDim.x) * T;
  for (int32_t t = 0; t < T; ++t) {
    const int32_t i = (c0 + t) * kFcgBlock + threadIdx.x;
    if (i >= n) break;
    const int32_t q = i >> 6;
    const int32_t base = off[q] + (i & 63), wq = (off[q + 1] - off[q]) >> 6;
    const ColSrc<PK> col(cols, q);
    const double a0 = ac[i];
    double pi[K], v[K];
#pragma unroll
    for (int r = 0; r < K; ++r) {
      pi[r] = p[r * ld + i];
      v[r] = a0 * pi[r];
    }
    int32_t k = 0;
    for (; k + kSellRun <= wq; k += kSellRun)
      spmv_run_k<kSellRun, K>(col, val, p, ld, base + k * 64, v);
    const int32_t e0 = base + k * 64;
    switch (wq - k) {
      case 1: spmv_run_k<1, K>(col, val, p, ld, e0, v); break;
      case 2: spmv_run_k<2, K>(col, val, p, ld, e0, v); break;
      case 3: spmv_run_k<3, K>(col, val, p, ld, e0, v); break;
      case 4: spmv_run_k<4, K>(col, val, p, ld, e0, v); break;
      case 5: spmv_run_k<5, K>(col, val, p, ld, e0, v); break;
      case 6: spmv_run_k<6, K>(col, val, p, ld, e0, v); break;
      case 7: spmv_run_k<7, K>(col, val, p, ld, e0, v); break;
      default: break;
    }
#pragma unroll
    for (int r = 0; r < K; ++r) {
      w[r * ld + i] = v[r];
      d[r] += pi[r] * v[r];
    }
  }
  block_sums_all<kFcgBlock, K>(d, lds);
  if (threadIdx.x == 0)
#pragma unroll
    for (int r = 0; r < K; ++r) partial[r * pstride + blockIdx.x] = d[r];
}

// ------------------------------------------------------------------ launchers

template <typename V>
void mcgs_sell_b(hipStream_t s, int k, int32_t r0, int32_t r1, const int32_t *off,
                 const SellCols &cols, const V *val, const V *ac, const V *b, V *x, int64_t ld,
                 int32_t zero_above) {
  if (r1 <= r0) return;
  const int32_t a = r0 & ~63;
  const dim3 g(nblocks(r1 - a)), t(kBlock);
  with_cols(cols, [&](auto pk) {
    constexpr bool P = decltype(pk)::value;
    switch (k) {
      case 2: hipLaunchKernelGGL((k_mcgs_sell_k<V, 2, P>), g, t, 0, s, r0, r1, off, cols, val, ac, b, x, ld, zero_above); break;
      case 4: hipLaunchKernelGGL((k_mcgs_sell_k<V, 4, P>), g, t, 0, s, r0, r1, off, cols, val, ac, b, x, ld, zero_above); break;
      case 8: hipLaunchKernelGGL((k_mcgs_sell_k<V, 8, P>), g, t, 0, s, r0, r1, off, cols, val, ac, b, x, ld, zero_above); break;
      default: hipLaunchKernelGGL((k_mcgs_sell_k<V, 1, P>), g, t, 0, s, r0, r1, off, cols, val, ac, b, x, ld, zero_above); break;
    }
  });
}

template <typename V>
void residual_sell_b(hipStream_t s, int k, int32_t n, int32_t zero_rows, const int32_t *off,
                     const SellCols &cols, const V *val, const V *ac, const V *b, const V *x,
                     V *r, int64_t ld) {
  if (n <= 0) return;
  const dim3 g(nblocks(n)), t(kBlock);
  with_cols(cols, [&](auto pk) {
    constexpr bool P = decltype(pk)::value;
    switch (k) {
      case 2: hipLaunchKernelGGL((k_residual_sell_k<V, 2, P>), g, t, 0, s, n, zero_rows, off, cols, val, ac, b, x, r, ld); break;
      case 4: hipLaunchKernelGGL((k_residual_sell_k<V, 4, P>), g, t, 0, s, n, zero_rows, off, cols, val, ac, b, x, r, ld); break;
      case 8: hipLaunchKernelGGL((k_residual_sell_k<V, 8, P>), g, t, 0, s, n, zero_rows, off, cols, val, ac, b, x, r, ld); break;
      default: hipLaunchKernelGGL((k_residual_sell_k<V, 1, P>), g, t, 0, s, n, zero_rows, off, cols, val, ac, b, x, r, ld); break;
    }
  });
}

template <typename V, bool PRO>
static void jacobi_b_t(hipStream_t s, int k, int32_t n, const int32_t *off, const SellCols &cols,
                       const V *val, const V *ac, const V *b, const V *x, int64_t ld,
                       const int32_t *parent, const V *xc, int64_t ldc, V *out, V omega) {
  const dim3 g(nblocks(n)), t(kBlock);
  with_cols(cols, [&](auto pk) {
    constexpr bool P = decltype(pk)::value;
    switch (k) {
      case 2: hipLaunchKernelGGL((k_jacobi_sell_k<V, PRO, 2, P>), g, t, 0, s, n, off, cols, val, ac, b, x, ld, parent, xc, ldc, out, omega); break;
      case 4: hipLaunchKernelGGL((k_jacobi_sell_k<V, PRO, 4, P>), g, t, 0, s, n, off, cols, val, ac, b, x, ld, parent, xc, ldc, out, omega); break;
      case 8: hipLaunchKernelGGL((k_jacobi_sell_k<V, PRO, 8, P>), g, t, 0, s, n, off, cols, val, ac, b, x, ld, parent, xc, ldc, out, omega); break;
      default: hipLaunchKernelGGL((k_jacobi_sell_k<V, PRO, 1, P>), g, t, 0, s, n, off, cols, val, ac, b, x, ld, parent, xc, ldc, out, omega); break;
    }
  });
}

template <typename V>
void jacobi_sell_b(hipStream_t s, int k, int32_t n, const int32_t *off, const SellCols &cols,
                   const V *val, const V *ac, const V *b, const V *x, int64_t ld,
                   const int32_t *parent, const V *xc, int64_t ldc, V *out, V omega) {
  if (n <= 0) return;
  if (parent) jacobi_b_t<V, true>(s, k, n, off, cols, val, ac, b, x, ld, parent, xc, ldc, out, omega);
  else jacobi_b_t<V, false>(s, k, n, off, cols, val, ac, b, x, ld, parent, xc, ldc, out, omega);
}

template <typename V, bool JAC>
static void restrict_b_t(hipStream_t s, int k, int32_t nc, const int32_t *off,
                         const int32_t *child, const V *r, int64_t ldf, const V *acc, V *bc,
                         V *xc, int64_t ldc, V omega, V scale) {
  const dim3 g(nblocks(nc)), t(kBlock);
  switch (k) {
    case 2: hipLaunchKernelGGL((k_restrict_sell_k<V, JAC, 2>), g, t, 0, s, nc, off, child, r, ldf, acc, bc, xc, ldc, omega, scale); break;
    case 4: hipLaunchKernelGGL((k_restrict_sell_k<V, JAC, 4>), g, t, 0, s, nc, off, child, r, ldf, acc, bc, xc, ldc, omega, scale); break;
    case 8: hipLaunchKernelGGL((k_restrict_sell_k<V, JAC, 8>), g, t, 0, s, nc, off, child, r, ldf, acc, bc, xc, ldc, omega, scale); break;
    default: hipLaunchKernelGGL((k_restrict_sell_k<V, JAC, 1>), g, t, 0, s, nc, off, child, r, ldf, acc, bc, xc, ldc, omega, scale); break;
  }
}

template <typename V>
void restrict_b(hipStream_t s, int k, int32_t nc, const int32_t *off, const int32_t *child,
                const V *r, int64_t ldf, const V *acc, V *bc, V *xc, int64_t ldc, V omega,
                V scale) {
  if (nc <= 0) return;
  if (acc) restrict_b_t<V, true>(s, k, nc, off, child, r, ldf, acc, bc, xc, ldc, omega, scale);
  else restrict_b_t<V, false>(s, k, nc, off, child, r, ldf, acc, bc, xc, ldc, omega, scale);
}

template <typename V>
void prolong_b(hipStream_t s, int k, int32_t r0, int32_t n, const int32_t *parent, const V *xc,
               int64_t ldc, V *x, int64_t ld) {
  if (n <= r0) return;
  const dim3 g(nblocks(n - r0)), t(kBlock);
  switch (k) {
    case 2: hipLaunchKernelGGL((k_prolong_from_k<V, 2>), g, t, 0, s, r0, n, parent, xc, ldc, x, ld); break;
    case 4: hipLaunchKernelGGL((k_prolong_from_k<V, 4>), g, t, 0, s, r0, n, parent, xc, ldc, x, ld); break;
    case 8: hipLaunchKernelGGL((k_prolong_from_k<V, 8>), g, t, 0, s, r0, n, parent, xc, ldc, x, ld); break;
    default: hipLaunchKernelGGL((k_prolong_from_k<V, 1>), g, t, 0, s, r0, n, parent, xc, ldc, x, ld); break;
  }
}

template <typename V>
void dense_solve_b(hipStream_t s, int k, int32_t n, const double *ainv, const V *b, int64_t ld,
                   V *x) {
  if (n <= 0) return;
  const dim3 g((n + kBlock / kWave - 1) / (kBlock / kWave)), t(kBlock);
  switch (k) {
    case 2: hipLaunchKernelGGL((k_dense_gemv_k<V, 2>), g, t, 0, s, n, ainv, b, ld, x); break;
    case 4: hipLaunchKernelGGL((k_dense_gemv_k<V, 4>), g, t, 0, s, n, ainv, b, ld, x); break;
    case 8: hipLaunchKernelGGL((k_dense_gemv_k<V, 8>), g, t, 0, s, n, ainv, b, ld, x); break;
    default: hipLaunchKernelGGL((k_dense_gemv_k<V, 1>), g, t, 0, s, n, ainv, b, ld, x); break;
  }
}

void spmv_dot_b(hipStream_t s, int k, int32_t n, const int32_t *off, const SellCols &cols,
                const double *ac, const double *val, const double *p, double *w, int64_t ld,
                double *partial, int64_t pstride) {
  const dim3 g(fold_blocks(n)), t(kFcgBlock);
  with_cols(cols, [&](auto pk) {
    constexpr bool P = decltype(pk)::value;
    switch (k) {
      case 2: hipLaunchKernelGGL((k_spmv_dot_k<2, P>), g, t, 0, s, n, off, cols, ac, val, p, w, ld, partial, pstride); break;
      case 4: hipLaunchKernelGGL((k_spmv_dot_k<4, P>), g, t, 0, s, n, off, cols, ac, val, p, w, ld, partial, pstride); break;
      case 8: hipLaunchKernelGGL((k_spmv_dot_k<8, P>), g, t, 0, s, n, off, cols, ac, val, p, w, ld, partial, pstride); break;
      default: hipLaunchKernelGGL((k_spmv_dot_k<1, P>), g, t, 0, s, n, off, cols, ac, val, p, w, ld, partial, pstride); break;
    }
  });
}

#define UFVM_INST_B(V)                                                                          \
  template void mcgs_sell_b<V>(hipStream_t, int, int32_t, int32_t, const int32_t *,             \
                               const SellCols &, const V *, const V *, const V *, V *, int64_t, \
                               int32_t);                                                        \
  template void residual_sell_b<V>(hipStream_t, int, int32_t, int32_t, const int32_t *,         \
                                   const SellCols &, const V *, const V *, const V *,           \
                                   const V *, V *, int64_t);                                    \
  template void jacobi_sell_b<V>(hipStream_t, int, int32_t, const int32_t *, const SellCols &,  \
                                 const V *, const V *, const V *, const V *, int64_t,           \
                                 const int32_t *, const V *, int64_t, V *, V);                  \
  template void restrict_b<V>(hipStream_t, int, int32_t, const int32_t *, const int32_t *,      \
                              const V *, int64_t, const V *, V *, V *, int64_t, V, V);          \
  template void prolong_b<V>(hipStream_t, int, int32_t, int32_t, const int32_t *, const V *,    \
                             int64_t, V *, int64_t);                                            \
  template void dense_solve_b<V>(hipStream_t, int, int32_t, const double *, const V *, int64_t, \
                                 V *);
UFVM_INST_B(float)
UFVM_INST_B(double)

}  // namespace k
}  // namespace ufvm
