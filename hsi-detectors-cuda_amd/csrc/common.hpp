// common.hpp - shared definitions for the gfx950 kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "hsi_mi355x.h"

namespace hsi_amd {

// Library and band limits. 32 targets covers every library the workloads use
// (1-32); 256 bands covers the 16/113/224-band sensors with room to spare and
// keeps the per-pixel band loops and statistics tiles bounded.
constexpr int kMaxTargets = 32;
constexpr int kMaxBands = 256;
// RX/ACE keep the 16x16 whitener blocks of the lower triangle in LDS as
// floats; 14 blocks (224 bands) is the most that fits beside the target tile.
constexpr int kMaxRxBands = 224;

// Wavefront width on CDNA.
constexpr int kWave = 64;

__host__ __device__ inline std::size_t div_up(std::size_t a, std::size_t b) { return (a + b - 1) / b; }
__host__ __device__ inline std::size_t round_up(std::size_t a, std::size_t b) { return div_up(a, b) * b; }

// Raw buffer view of a whole BSQ cube for buffer_load: every lane supplies a
// 32-bit byte offset and loads at or past `bytes` return 0, so kernels that
// read padded K blocks or ragged tiles need no guard around the load. The
// cube must be smaller than 4 GiB (checked by the launchers).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t cube_rsrc(const void* base, std::uint32_t bytes) {
  // 0x00020000: the gfx9 word 3 for raw (untyped) 32-bit buffer access.
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, static_cast<int>(bytes), 0x00020000);
}
__device__ __forceinline__ float buffer_elem(__amdgpu_buffer_rsrc_t r, std::uint32_t off, float) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0));
}
__device__ __forceinline__ float buffer_elem(__amdgpu_buffer_rsrc_t r, std::uint32_t off, _Float16) {
  return static_cast<float>(__builtin_bit_cast(_Float16, __builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)));
}

// Element tag of an HSI_Q16 cube: a 16-bit unsigned code q of a per-band
// affine map, decoded as fmaf(scale[b], q, offset[b]) (q16_decode). The
// buffer loads below return the code itself as a float (exact).
struct q16_t {
  std::uint16_t code;
};
__device__ __forceinline__ float buffer_elem(__amdgpu_buffer_rsrc_t r, std::uint32_t off, q16_t) {
  return static_cast<float>(static_cast<std::uint16_t>(__builtin_amdgcn_raw_buffer_load_b16(r, off, 0, 0)));
}
__device__ __forceinline__ float q16_decode(float code, float scale, float offset) {
  return __builtin_fmaf(scale, code, offset);
}
template <typename In> struct is_q16 { static constexpr bool value = false; };
template <> struct is_q16<q16_t> { static constexpr bool value = true; };

// Two adjacent elements of one band plane (the first at byte `off`, which
// must be a multiple of 2 * sizeof(element)), as floats.
__device__ __forceinline__ float2 buffer_pair(__amdgpu_buffer_rsrc_t r, std::uint32_t off, float) {
  return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}
__device__ __forceinline__ float2 buffer_pair(__amdgpu_buffer_rsrc_t r, std::uint32_t off, _Float16) {
  const unsigned w = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
  return make_float2(static_cast<float>(__builtin_bit_cast(_Float16, static_cast<std::uint16_t>(w & 0xffffu))),
                     static_cast<float>(__builtin_bit_cast(_Float16, static_cast<std::uint16_t>(w >> 16))));
}
__device__ __forceinline__ float2 buffer_pair(__amdgpu_buffer_rsrc_t r, std::uint32_t off, q16_t) {
  const unsigned w = __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
  return make_float2(static_cast<float>(w & 0xffffu), static_cast<float>(w >> 16));
}
// Whether a cube of `bands` planes of `stride` elements fits one buffer view.
__host__ __device__ inline bool fits_buffer(int bands, std::size_t stride, std::size_t elem_bytes) {
  return static_cast<std::size_t>(bands) * stride * elem_bytes < (static_cast<std::size_t>(1) << 32) - 256;
}

// float(pi/2) widened: the angle the reference reports for a pixel or target
// with no direction (include/hsi/sam.hpp:25).
constexpr double kMaxSamAngle = static_cast<double>(1.57079632679489661923f);

}  // namespace hsi_amd

// The device status word bits (HSI_STATUS_*) are public: include/hsi_mi355x.h.
