// ctr_residual_noalloc: the relative data residual of ctr_residual for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
// forbid hipMalloc).  Its reduction scratch is a fixed device array of this file instead of
// plan memory allocated on first use, and the fused forward projector writes no residual
// sinogram, so the call allocates nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>

#include "ctr_internal.hpp"

namespace ctr {
namespace {

// The scratch: the per-workgroup partial sums of one view chunk (2 doubles per workgroup
// of the fused forward projector), then kResidualSums doubles of per-chunk totals.  4 MiB;
// every GPU the library runs on holds its own copy until the process ends.
constexpr size_t kScratchDoubles = size_t(1) << 19;
static __device__ __align__(256) double scratch[kScratchDoubles];

// The views of one chunk, as ctr_residual (the same chunks give the same sums, bit for bit).
constexpr int kChunkViews = 256;

// Device address of each GPU's copy of scratch, looked up once per device (with that
// device current) on its first call; host code uses only these addresses, never the
// array's name.
constexpr int kMaxDevices = 64;
std::mutex g_mu;
double* g_scratch[kMaxDevices] = {};

int scratch_on(int device, double** out) {
  if (device < 0 || device >= kMaxDevices)
    return fail(CTR_ERR_ARG, "ctr_residual_noalloc: device id out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CTR_ERR_HIP, std::string("ctr_residual_noalloc: hipGetSymbolAddress: ") +
                                   hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return CTR_OK;
}

}  // namespace
}  // namespace ctr

using namespace ctr;

int ctr_residual_noalloc(ctr_plan* plan, const float* d_x, const float* d_prj, double* out) {
  Plan* P = plan ? plan->impl : nullptr;
  if (!P || !d_x || !d_prj || !out) return fail(CTR_ERR_ARG, "NULL argument");
  if (P->method != CTR_METHOD_DIRECT)
    return fail(CTR_ERR_UNSUPPORTED,
                "ctr_residual_noalloc: the direct method only (the integral-image method "
                "builds its tables in plan memory)");
  CTR_HIP(hipSetDevice(P->device));
  const int chunk = std::min(P->d.PN, kChunkViews);
  const size_t parts = 2 * (size_t)fp_partial_blocks(*P, chunk);
  if (parts + kResidualSums > kScratchDoubles)
    return fail(CTR_ERR_ARG,
                "ctr_residual_noalloc: a chunk of " + std::to_string(chunk) + " views of " +
                    std::to_string(P->d.DNU) + " x " + std::to_string(P->d.DNV) +
                    " rays needs " + std::to_string(parts + kResidualSums) +
                    " doubles of scratch; the maximum is " + std::to_string(kScratchDoubles) +
                    " (4 MiB): use ctr_residual");
  double* s = nullptr;
  if (int rc = scratch_on(P->device, &s)) return rc;
  // every partial slot and every total is written before it is read: nothing to clear
  return residual_run(*P, d_x, d_prj, nullptr, chunk, s, s + parts, P->stream, out);
}
