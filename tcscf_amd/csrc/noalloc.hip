// computeF without device allocation: tcscf_compute_F_noalloc runs the same
// contraction as tcscf_compute_F, with its workspace (the packed pair
// densities and padded coordinates) in a fixed array of this file instead of
// memory the library allocates.  For callers that cannot allocate: code under
// HIP graph capture, serving loops that must not call hipMalloc.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.hpp"

namespace {

// 128 MiB: the workspace of every precision at the 30,200-point, 30-function
// Ne grid (94 MB for BF16x5, 125 MB for FP64).  Each GPU the library runs on
// holds its own copy from module load until the process ends.
constexpr size_t kScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_scratch_ptr[kMaxDevices];  // device address of `scratch` per device, once looked up

// The device address of this device's `scratch`, looked up once per device
// with that device current.  The array's host-side name is only a stand-in
// for the symbol: host code passes this address, never the name.
int scratch_address(void** out) {
  int dev = 0;
  TC_HIP(hipGetDevice(&dev));
  TC_CHECK(dev >= 0 && dev < kMaxDevices, "tcscf_compute_F_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    TC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    TC_CHECK(p != nullptr, "tcscf_compute_F_noalloc: no device address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}

}  // namespace

extern "C" {

int tcscf_compute_F_noalloc(const double* pts1, int64_t n1, const double* pts2, const double* w2,
                            const int32_t* ord2, int64_t n2, const double* A2, int nb, int kernel,
                            const tcscf_jastrow* jastrow, int same_spin, int precision, double scale, void* F,
                            int out_dtype, void* stream) {
  TC_CHECK(n2 > 0 && nb > 0 && precision >= 0 && precision <= 4, "tcscf_compute_F_noalloc: bad arguments");
  const size_t need = tcscf_F_workspace_bytes(n2, nb, precision);
  if (need > kScratchBytes) {
    tc::set_error("tcscf_compute_F_noalloc: the workspace of this contraction (" + std::to_string(need) +
              " bytes) exceeds the fixed scratch of " + std::to_string(kScratchBytes) +
              " bytes; use tcscf_compute_F");
    return TCSCF_EINVAL;
  }
  void* ws = nullptr;
  int rc = scratch_address(&ws);
  if (rc) return rc;
  return tc::compute_F_into(pts1, n1, pts2, w2, ord2, n2, A2, nb, kernel, jastrow, same_spin, precision, scale, F,
                            out_dtype, stream, static_cast<char*>(ws));
}

}  // extern "C"
