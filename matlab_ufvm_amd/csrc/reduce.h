// reduce.h -- deterministic wavefront / block / grid reductions (device code).
#pragma once
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ufvm {
namespace k {

constexpr int kWave = 64;

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
  return v;
}

// sum over a kBlock-thread block; result valid in thread 0
template <typename T>
__device__ inline T block_sum(T v, T *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0) {
    for (int w = 0; w < kBlock / kWave; ++w) r += lds[w];
  }
  __syncthreads();
  return r;
}

template <typename T>
__device__ inline T block_max(T v, T *lds) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v = fmax(v, __shfl_down(v, off, kWave));
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T r = lds[0];
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; ++w) r = fmax(r, lds[w]);
  }
  __syncthreads();
  return r;
}

// Deterministic grid reduction of W values per block: every block writes
// its partial values, the last block to finish (counted with an atomic)
// combines all partials in block order (op[c] = 0: sum, 1: max, 2: min) and
// runs `fin` on the totals.  The counter is reset by that block, ready for
// the next launch.
template <int W, typename Fin>
__device__ inline void grid_reduce_last(double (&v)[W], const int (&op)[W], double *partial,
                                        unsigned *counter, Fin fin) {
  __shared__ double lds[kBlock / kWave];
  __shared__ bool last;
  for (int c = 0; c < W; ++c) {
    if (op[c] == 0) v[c] = block_sum(v[c], lds);
    else if (op[c] == 1) v[c] = block_max(v[c], lds);
    else v[c] = -block_max(-v[c], lds);
  }
  if (threadIdx.x == 0) {
    for (int c = 0; c < W; ++c) partial[(int64_t)blockIdx.x * W + c] = v[c];
    __threadfence();
    const unsigned t = atomicAdd(counter, 1u);
    last = (t == gridDim.x - 1);
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double tot[W];
  for (int c = 0; c < W; ++c) {
    double a = op[c] == 0 ? 0.0 : __builtin_nontemporal_load(&partial[c]);
    for (int32_t q = threadIdx.x; q < (int32_t)gridDim.x; q += kBlock) {
      const double x = __builtin_nontemporal_load(&partial[(int64_t)q * W + c]);
      if (op[c] == 0) a += x;
      else if (op[c] == 1) a = fmax(a, x);
      else a = fmin(a, x);
    }
    if (op[c] == 0) tot[c] = block_sum(a, lds);
    else if (op[c] == 1) tot[c] = block_max(a, lds);
    else tot[c] = -block_max(-a, lds);
  }
  if (threadIdx.x == 0) {
    fin(tot);
    *counter = 0u;
  }
}

template <int W, typename Fin>
__device__ inline void grid_reduce_last(double (&v)[W], double *partial, unsigned *counter,
                                        Fin fin) {
  int op[W];
  for (int c = 0; c < W; ++c) op[c] = 0;
  grid_reduce_last<W>(v, op, partial, counter, fin);
}

}  // namespace k
}  // namespace ufvm
