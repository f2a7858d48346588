// reduce.h -- deterministic wavefront / block reductions (device code).
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

// W sums over an NB-thread block, returned to every thread (each thread adds
// the wave sums in the same order, so all get the same values).  lds holds
// W * NB / kWave elements.
template <int NB, int W, typename T>
__device__ inline void block_sums_all(T (&v)[W], T *lds) {
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
  for (int c = 0; c < W; ++c) {
    v[c] = wave_sum(v[c]);
    if (lane == 0) lds[c * (NB / kWave) + wid] = v[c];
  }
  __syncthreads();
  // not unrolled across c: W x NB/kWave LDS values held at once would spill
#pragma unroll 1
  for (int c = 0; c < W; ++c) {
    T r = 0;
    for (int w = 0; w < NB / kWave; ++w) r += lds[c * (NB / kWave) + w];
    v[c] = r;
  }
  __syncthreads();
}

// sum over an NB-thread block; result valid in thread 0
template <int NB, typename T>
__device__ inline T block_sum_n(T v, T *lds) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  T r = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NB / kWave; ++w) r += lds[w];
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

}  // namespace k
}  // namespace ufvm
