// Shared host/device helpers for the tcscf gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <mutex>
#include <set>
#include <string>
#include <utility>

#include "../../include/tcscf_hip.h"

namespace tc {

// The BF16 split precisions (MFMA on bf16 pieces of K and of the pair densities).
inline bool is_bf16(int precision) {
  return precision == TCSCF_BF16X3 || precision == TCSCF_BF16X4 || precision == TCSCF_BF16X5;
}

void set_error(const std::string& msg);
// The calling thread's last message (what tcscf_last_error returns).
const std::string& last_error();

inline int64_t npair_of(int nb) { return (int64_t)nb * (nb + 1) / 2; }

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// Returns TCSCF_EHIP and records the message when a HIP call fails.
#define TC_HIP(call)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess) {                                                                \
      ::tc::set_error(std::string(#call) + ": " + hipGetErrorString(e_));                  \
      return TCSCF_EHIP;                                                                   \
    }                                                                                      \
  } while (0)

#define TC_CHECK(cond, msg)                                                                \
  do {                                                                                     \
    if (!(cond)) {                                                                         \
      ::tc::set_error(msg);                                                                \
      return TCSCF_EINVAL;                                                                 \
    }                                                                                      \
  } while (0)

// LDS of one gfx950 workgroup.
constexpr int kMaxLdsBytes = 160 * 1024;

// Lets kernel fn launch with up to kMaxLdsBytes of dynamic LDS on the current
// device.  hipFuncAttributeMaxDynamicSharedMemorySize is a property of the
// kernel (per device, process-wide), not of one launch, so it is set once, to
// the hardware maximum: set per plan, a later plan of the same kernel with a
// smaller footprint would lower the cap below an earlier plan's launch size.
inline hipError_t lds_opt_in(const void* fn) {
  static std::mutex mu;
  static std::set<std::pair<int, const void*>> done;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  std::lock_guard<std::mutex> lock(mu);
  if (done.count({dev, fn})) return hipSuccess;
  e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, kMaxLdsBytes);
  if (e == hipSuccess) done.insert({dev, fn});
  return e;
}

// tcscf_compute_F with the workspace given (tcscf_F_workspace_bytes of it,
// device memory) or, for ws_fixed = null, from the library's scratch slot.
int compute_F_into(const double* pts1, int64_t n1, const double* pts2, const double* w2, const int32_t* ord2,
                   int64_t n2, const double* A2, int nb, int kernel, const tcscf_jastrow* jp, int same_spin,
                   int precision, double scale, void* F, int out_dtype, void* stream, char* ws_fixed);

// Device scratch owned by the library: grown on demand, never shrunk, one per
// (device, slot).  Calls that share a slot on one device must be ordered on
// one stream (every entry point documents the slots it uses).
void* scratch(int slot, size_t bytes, int* rc);

enum ScratchSlot {
  kSlotPairs = 0,
  kSlotCoords = 1,
  kSlotSweep = 2,
  kSlotSweepOut = 3,
  kSlotSmall = 4,
  kSlotBatchMeta = 5,  // + 0 / 1: batch descriptors for the two tile heights
  kSlotQmc = 7,        // tcscf_eri_qmc6d: the per-point pair densities / r12 and the contraction
  kSlotBatchMetaV = 8, // batch descriptors of the three-component (V) tiles
  kSlotSplit = 9,      // partial sums of the pipelined BF16 kernel's split last round
  kNumSlots = 10
};

}  // namespace tc
