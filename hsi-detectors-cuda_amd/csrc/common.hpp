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

// float(pi/2) widened: the angle the reference reports for a pixel or target
// with no direction (include/hsi/sam.hpp:25).
constexpr double kMaxSamAngle = static_cast<double>(1.57079632679489661923f);

}  // namespace hsi_amd

// Device status word bits written by the statistics build.
#define HSI_STATUS_COV_NOT_PD 1u
#define HSI_STATUS_CORR_NOT_PD 2u
