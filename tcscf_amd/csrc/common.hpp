// Shared host/device helpers for the tcscf gfx950 library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tcscf_hip.h"

namespace tc {

void set_error(const std::string& msg);

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
  kNumSlots = 8
};

}  // namespace tc
