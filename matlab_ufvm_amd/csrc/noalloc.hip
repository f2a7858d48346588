// noalloc.hip -- ufvm_compute_residuals_noalloc: the residual and its L1 norm
// of ufvm_compute_residuals for callers that cannot allocate device memory
// (work captured into HIP graphs, tight serving loops).  The call allocates
// nothing and does not wait for the device: r and the norm go to device
// memory the caller owns, enqueued on the context's stream, and the per-block
// partial sums of |r| live in the file-scope array `scratch` below.
#include <hip/hip_runtime.h>

#include <mutex>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "state.h"
#include "ufvm.h"

using namespace ufvm;
using namespace ufvm::detail;

namespace {

// one partial per k::kBlock rows: UFVM_NOALLOC_MAX_ROWS / 256 doubles (1 MiB),
// held on every GPU the library runs on until the process ends
constexpr int64_t kMaxRows = UFVM_NOALLOC_MAX_ROWS;
static __device__ __align__(256) double scratch[kMaxRows / k::kBlock];

std::mutex g_mutex;
std::vector<double *> g_ptr;  // per device: the address of that device's `scratch`

double *scratch_of(int device) {
  std::lock_guard<std::mutex> g(g_mutex);
  return device >= 0 && device < (int)g_ptr.size() ? g_ptr[device] : nullptr;
}

}  // namespace

namespace ufvm {
namespace detail {

// Called by ufvm_ctx_create with `device` current: looks up, once per device,
// the device address of `scratch` (host code uses only that address, never
// the array's name).
void noalloc_setup(int device) {
  std::lock_guard<std::mutex> g(g_mutex);
  if ((int)g_ptr.size() <= device) g_ptr.resize(device + 1, nullptr);
  if (g_ptr[device]) return;
  void *p = nullptr;
  UFVM_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  g_ptr[device] = static_cast<double *>(p);
}

}  // namespace detail
}  // namespace ufvm

extern "C" int ufvm_compute_residuals_noalloc(ufvm_ctx *ctx, int32_t sys, int32_t level,
                                              double *r_dev, double *l1_dev) {
  return guard(ctx, [&] {
    if (!r_dev || !l1_dev) throw Error(UFVM_ERR_ARG, "null output pointer");
    RefLevel &L = level_of(sys_of(ctx, sys), level);
    if ((int64_t)L.n > kMaxRows)
      throw Error(UFVM_ERR_ARG, "level has more rows than UFVM_NOALLOC_MAX_ROWS");
    double *part = scratch_of(ctx->device);
    if (!part) throw Error(UFVM_ERR_STATE, "no scratch address for this device");
    // the kernels and the block order of ufvm_compute_residuals: the same bits.
    // Every partial the reduction reads is written first by this call.
    k::residual_ref(ctx->stream, L.n, L.ptr.p, L.col.p, L.ac.p, L.anb.p, L.bc.p, L.dphi.p, r_dev,
                    part);
    k::reduce_final(ctx->stream, part, k::nblocks(L.n), l1_dev);
  });
}
