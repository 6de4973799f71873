// Device helpers shared by the HIP sources (included by .hip files only).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

namespace tc {

// Raw buffer resources (gfx9 word 3: 32-bit data format, no swizzle): loads
// at byte offsets at or past the view's size return zero, so a row can be read
// with one scalar base, one lane offset and no bounds branches.
constexpr int kBufferWord3 = 0x00020000;

// A wave-uniform 64-bit integer moved to scalar registers (a loop counter the
// compiler would otherwise keep per lane, turning buffer loads of its rows
// into per-lane loops).
__device__ inline int64_t uniform_i64(int64_t v) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(uint64_t)v);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((uint64_t)v >> 32));
  return (int64_t)(((uint64_t)hi << 32) | lo);
}

template <class T>
__device__ inline __amdgpu_buffer_rsrc_t buffer_view(const T* base, int64_t elems) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<T*>(base), (short)0, (int)(elems * (int64_t)sizeof(T)),
                                           kBufferWord3);
}

// The whole byte offset goes in the per-lane (vector) offset: the range check
// covers the vector and instruction offsets, not the scalar one.
template <class T>
__device__ inline T buffer_load(__amdgpu_buffer_rsrc_t r, int byte_offset);
template <>
__device__ inline double buffer_load<double>(__amdgpu_buffer_rsrc_t r, int byte_offset) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, byte_offset, 0, 0));
}
template <>
__device__ inline float buffer_load<float>(__amdgpu_buffer_rsrc_t r, int byte_offset) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, byte_offset, 0, 0));
}

// dst[i] = src[i] for i < n: the upload of a graph-replayed SCF iteration,
// src being mapped page-locked host memory (a kernel node, not a copy node).
static __global__ __launch_bounds__(256) void copy_doubles(const double* __restrict__ src, double* __restrict__ dst,
                                                           int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) dst[i] = src[i];
}
inline dim3 copy_grid(int64_t n) { return dim3((unsigned)std::min<int64_t>(64, (n + 255) / 256)); }

}  // namespace tc
