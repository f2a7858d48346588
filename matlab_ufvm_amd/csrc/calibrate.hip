// calibrate.hip -- kernels of known traffic for calibrating the memory
// counters of a profiler (measurement tooling, not part of the solver): a
// stream copy with 4-byte and 8-byte lanes, and a gather through a
// permutation inside 64-element windows (every fetched line used whole).
// Each kernel moves exactly the bytes its comment states.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int kThreads = 256;

// dst[i] = src[i]: reads n * sizeof(T), writes n * sizeof(T)
template <typename T>
__global__ void __launch_bounds__(kThreads) k_cal_stream(int64_t n, const T *__restrict__ src,
                                                         T *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// dst[i] = src[idx[i]]: reads 4 n (idx) + 4 n (src), writes 4 n
__global__ void __launch_bounds__(kThreads) k_cal_gather(int64_t n, const int32_t *__restrict__ idx,
                                                         const float *__restrict__ src,
                                                         float *__restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// idx[i] = a permutation of the 64 indices of i's window (i ^ 37 inside it)
__global__ void __launch_bounds__(kThreads) k_cal_index(int64_t n, int32_t *__restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) {
    const int64_t j = (i & ~(int64_t)63) | ((i & 63) ^ 37);
    idx[i] = (int32_t)(j < n ? j : i);
  }
}

inline dim3 grid(int64_t n) { return dim3((unsigned)((n + kThreads - 1) / kThreads)); }

}  // namespace

extern "C" {

// kind 0: float stream, 1: double stream, 2: float gather; bytes is the size
// of each of the three buffers (a, b: data, c: indices for the gather);
// launches the kernel `reps` times on the null stream and waits.  Returns
// the HIP status.
int ufvm_calibration_run(int kind, void *a, void *b, void *c, int64_t bytes, int reps) {
  if (kind == 2) {
    const int64_t n = bytes / 4;
    hipLaunchKernelGGL(k_cal_index, grid(n), dim3(kThreads), 0, 0, n, (int32_t *)c);
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_cal_gather, grid(n), dim3(kThreads), 0, 0, n, (const int32_t *)c,
                         (const float *)a, (float *)b);
  } else if (kind == 1) {
    const int64_t n = bytes / 8;
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_cal_stream<double>, grid(n), dim3(kThreads), 0, 0, n,
                         (const double *)a, (double *)b);
  } else {
    const int64_t n = bytes / 4;
    for (int r = 0; r < reps; ++r)
      hipLaunchKernelGGL(k_cal_stream<float>, grid(n), dim3(kThreads), 0, 0, n,
                         (const float *)a, (float *)b);
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return (int)e;
  return (int)hipDeviceSynchronize();
}

}  // extern "C"
