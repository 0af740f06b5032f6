// sam_cem.hip - SAM and CEM scoring over a BSQ cube (gfx950).
//
// Both detectors are bandwidth-bound: per pixel they read `bands` elements
// once and do 2 FLOP per band per target. The kernel streams the cube with
// 16-byte loads (four FP32 or four FP16 pixels per lane per band, so one
// wave64 load instruction moves a contiguous 1 KiB / 512 B run of a band
// plane) and scores every target from that one read.
//
// Arithmetic is FP64: each float (or half) is widened exactly, every product
// x*t of two floats is exact in double, and sums run in band order - the
// same values the reference CPU scorer forms (sam_cpu.cpp:48-77,
// detectors_cpu.cpp:158-187), so scores agree to the last bit apart from the
// final libm-vs-device acos. On gfx950 an FP64 FMA issues at the same rate
// as a non-packed FP32 FMA, so exactness costs no throughput here.
//
// The target library is uniform across the wavefront: it is stored as
// doubles band-major, lib[b * NT + t], and the compiler reads each band's NT
// values with scalar loads into SGPRs that feed the FMAs directly - the MI355X
// counterpart of the reference's __constant__ broadcast (sam.cu:19-23).
#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

template <typename In, int PPT>
struct VecOf;
template <>
struct VecOf<float, 4> {
  typedef float type __attribute__((ext_vector_type(4)));
};
template <>
struct VecOf<float, 2> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct VecOf<_Float16, 4> {
  typedef _Float16 type __attribute__((ext_vector_type(4)));
};
template <>
struct VecOf<_Float16, 2> {
  typedef _Float16 type __attribute__((ext_vector_type(2)));
};

template <typename V, int PPT>
__device__ __forceinline__ void widen(const V& v, double (&x)[PPT]) {
  if constexpr (PPT == 4) {
    x[0] = static_cast<double>(static_cast<float>(v.x));
    x[1] = static_cast<double>(static_cast<float>(v.y));
    x[2] = static_cast<double>(static_cast<float>(v.z));
    x[3] = static_cast<double>(static_cast<float>(v.w));
  } else {
    x[0] = static_cast<double>(static_cast<float>(v.x));
    x[1] = static_cast<double>(static_cast<float>(v.y));
  }
}

// Stores PPT consecutive 4-byte results; the output maps are 4*PPT-byte
// aligned at every group start (checked by the launcher).
template <int PPT, typename T>
__device__ __forceinline__ void store_group(T* dst, const T (&v)[PPT]) {
  if constexpr (PPT == 4) {
    typedef T V __attribute__((ext_vector_type(4)));
    *reinterpret_cast<V*>(dst) = V{v[0], v[1], v[2], v[3]};
  } else {
    typedef T V __attribute__((ext_vector_type(2)));
    *reinterpret_cast<V*>(dst) = V{v[0], v[1]};
  }
}

// Best-match state for one pixel, in angle order without taking an acos per
// target. kind 0: nothing yet (angle +inf); 1: degenerate (pixel or target
// without direction, angle = float(pi/2) as angle_of returns); 2: a clamped
// cosine. A candidate replaces the best only when its angle is strictly
// smaller, so the first target keeps a tie, as sam_best_cpu's `a < best`.
struct SamBest {
  int kind = 0;
  double cosine = 0.0;
  int target = -1;
};

__device__ __forceinline__ void sam_offer(SamBest& best, bool degenerate, double c, int t) {
  if (degenerate) {
    if (best.kind == 0) {
      best.kind = 1;
      best.target = t;
    } else if (best.kind == 2 && kMaxSamAngle < acos(best.cosine)) {
      best.kind = 1;
      best.target = t;
    }
    return;
  }
  if (!(c == c)) return;  // NaN never wins a strict comparison
  if (best.kind == 0) {
    best.kind = 2;
    best.cosine = c;
    best.target = t;
  } else if (best.kind == 1) {
    if (acos(c) < kMaxSamAngle) {
      best.kind = 2;
      best.cosine = c;
      best.target = t;
    }
  } else if (c > best.cosine) {
    // acos is decreasing, so a larger cosine is a smaller angle - unless the
    // two cosines are close enough to round to the same angle (only within a
    // few 1e-16 of each other), where the earlier target must keep the tie.
    if (c - best.cosine > 4e-16 || acos(c) < acos(best.cosine)) {
      best.cosine = c;
      best.target = t;
    }
  }
}

// One thread's PPT consecutive pixels starting at p0 (< pixels).
template <typename In, int NT, int PPT, bool DO_SAM, bool DO_CEM>
__device__ __forceinline__ void sam_cem_group(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    const double* __restrict__ slib, const float* __restrict__ snorm,
    const double* __restrict__ srnorm, int exact_all,
    const double* __restrict__ clib, int num_targets, float* __restrict__ s_score,
    int* __restrict__ s_target, float* __restrict__ c_score, int* __restrict__ c_target, std::size_t p0) {
  using V = typename VecOf<In, PPT>::type;

  double sq[PPT];
  double sd[DO_SAM ? NT : 1][PPT];
  double cd[DO_CEM ? NT : 1][PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    sq[k] = 0.0;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if constexpr (DO_SAM) sd[t][k] = 0.0;
      if constexpr (DO_CEM) cd[t][k] = 0.0;
    }
  }

  const In* base = cube + p0;
  // Band planes in flight per lane: for SAM alone 12 at four pixels per lane
  // and 16 at two, 8 for CEM and the fused form (CEM alone 88 -> 80 us at
  // 1024x1024x113, fused 132 -> 125 us).
  constexpr int U = (DO_SAM && !DO_CEM) ? (PPT == 4 ? 12 : 16) : 8;
  int b = 0;
  for (; b + U <= bands; b += U) {
    V v[U];
    // Non-temporal: the cube is read once, so it need not displace other
    // lines in L2 and the last-level cache.
#pragma unroll
    for (int u = 0; u < U; ++u)
      v[u] = __builtin_nontemporal_load(reinterpret_cast<const V*>(base + static_cast<std::size_t>(b + u) * stride));
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double x[PPT];
      widen<V, PPT>(v[u], x);
      const double* sl = slib + static_cast<std::size_t>(b + u) * NT;
      const double* cl = clib + static_cast<std::size_t>(b + u) * NT;
#pragma unroll
      for (int k = 0; k < PPT; ++k) {
        if constexpr (DO_SAM) sq[k] = __fma_rn(x[k], x[k], sq[k]);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if constexpr (DO_SAM) sd[t][k] = __fma_rn(x[k], sl[t], sd[t][k]);
          if constexpr (DO_CEM) cd[t][k] = __fma_rn(cl[t], x[k], cd[t][k]);
        }
      }
    }
  }
  for (; b < bands; ++b) {
    const V v = *reinterpret_cast<const V*>(base + static_cast<std::size_t>(b) * stride);
    double x[PPT];
    widen<V, PPT>(v, x);
    const double* sl = slib + static_cast<std::size_t>(b) * NT;
    const double* cl = clib + static_cast<std::size_t>(b) * NT;
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      if constexpr (DO_SAM) sq[k] = __fma_rn(x[k], x[k], sq[k]);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if constexpr (DO_SAM) sd[t][k] = __fma_rn(x[k], sl[t], sd[t][k]);
        if constexpr (DO_CEM) cd[t][k] = __fma_rn(cl[t], x[k], cd[t][k]);
      }
    }
  }

  float s_out[PPT], c_out[PPT];
  int st_out[PPT], ct_out[PPT];
#pragma unroll
  for (int k = 0; k < PPT; ++k) {
    if constexpr (DO_SAM) {
      // cosine = dot / (sqrt(|x|^2) * |t|), clamped, as angle_of
      // (sam_cpu.cpp:13-18): correctly rounded sqrt, product and quotient.
      const double root = sqrt(sq[k]);
      // When every target has a direction and so does this pixel, rank on
      // c'_t = dot_t * (1/root) * (1/|t|), a few ulps from the unclamped
      // exact cosine, and take the exact quotient only for the targets that
      // could still win: those whose clamped c' is within 1e-14 of the best
      // clamped c'. That window holds every cosine whose acos can round to
      // the best angle (at most ~4.4e-16 apart). Clamping ties whole groups:
      // with a float |t| the unclamped cosine can pass 1 by ~1e-7, and every
      // target past 1 (or, if the best is at -1, every target) scores the
      // same angle, so the first must win. Otherwise every target goes
      // through the exact formula.
      const bool all_exact = exact_all || !(sq[k] > 0.0);
      const double rroot = 1.0 / root;
      double approx[NT];
      double top = -__longlong_as_double(0x7ff0000000000000ll);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        approx[t] = sd[t][k] * rroot * srnorm[t];
        if (t < num_targets && approx[t] > top) top = approx[t];
      }
      const double floor_c = (all_exact || top <= -1.0 + 1e-14)
                                 ? -__longlong_as_double(0x7ff0000000000000ll)
                                 : (top < 1.0 ? top : 1.0) - 1e-14;
      SamBest best;
      if constexpr (NT <= 16) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if (t < num_targets && (all_exact || approx[t] >= floor_c)) {
            const double norm = static_cast<double>(snorm[t]);
            const bool degenerate = sq[k] <= 0.0 || norm <= 0.0;
            double c = sd[t][k] / (root * norm);
            c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
            sam_offer(best, degenerate, c, t);
          }
        }
      } else {
        // 32 targets: the exact quotients are formed in an unrolled pass
        // (so the accumulators stay in registers) and offered in a rolled
        // one, which keeps the acos-carrying tie logic out of the unrolled
        // code; only these per-pixel candidate values may live in scratch.
        double cex[NT];
        unsigned long long take = 0;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          cex[t] = 0.0;
          if (t < num_targets && (all_exact || approx[t] >= floor_c)) {
            double c = sd[t][k] / (root * static_cast<double>(snorm[t]));
            cex[t] = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
            take |= 1ull << t;
          }
        }
#pragma unroll 1
        for (int t = 0; t < NT; ++t)
          if ((take >> t) & 1ull)
            sam_offer(best, sq[k] <= 0.0 || static_cast<double>(snorm[t]) <= 0.0, cex[t], t);
      }
      double angle = __longlong_as_double(0x7ff0000000000000ll);  // +inf
      if (best.kind == 1) angle = kMaxSamAngle;
      else if (best.kind == 2) angle = acos(best.cosine);
      s_out[k] = static_cast<float>(angle);
      st_out[k] = best.target;
    }
    if constexpr (DO_CEM) {
      double best = -__longlong_as_double(0x7ff0000000000000ll);
      int best_t = -1;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        if (t < num_targets && cd[t][k] > best) {
          best = cd[t][k];
          best_t = t;
        }
      }
      c_out[k] = static_cast<float>(best);
      ct_out[k] = best_t;
    }
  }
  // Whole groups leave as one 8- or 16-byte store per map, so a wave writes
  // full cache lines; the ragged tail of the frame stores pixel by pixel.
  if (p0 + PPT <= pixels) {
    if constexpr (DO_SAM) {
      if (s_score) store_group<PPT>(s_score + p0, s_out);
      if (s_target) store_group<PPT>(s_target + p0, st_out);
    }
    if constexpr (DO_CEM) {
      if (c_score) store_group<PPT>(c_score + p0, c_out);
      if (c_target) store_group<PPT>(c_target + p0, ct_out);
    }
  } else {
#pragma unroll
    for (int k = 0; k < PPT; ++k) {
      const std::size_t p = p0 + k;
      if (p >= pixels) break;
      if constexpr (DO_SAM) {
        if (s_score) s_score[p] = s_out[k];
        if (s_target) s_target[p] = st_out[k];
      }
      if constexpr (DO_CEM) {
        if (c_score) c_score[p] = c_out[k];
        if (c_target) c_target[p] = ct_out[k];
      }
    }
  }
}

template <typename In, int NT, int PPT, bool DO_SAM, bool DO_CEM>
__global__ __launch_bounds__(256) void sam_cem_kernel(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    const double* __restrict__ slib, const float* __restrict__ snorm,
    const double* __restrict__ srnorm, int exact_all,
    const double* __restrict__ clib, int num_targets, float* __restrict__ s_score,
    int* __restrict__ s_target, float* __restrict__ c_score, int* __restrict__ c_target) {
  // Grid-stride over pixel groups: a grid of at most four workgroups per CU
  // keeps every CU streaming to the end of the frame.
  const std::size_t step = static_cast<std::size_t>(gridDim.x) * blockDim.x * PPT;
  for (std::size_t p0 = (blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x) * PPT; p0 < pixels;
       p0 += step)
    sam_cem_group<In, NT, PPT, DO_SAM, DO_CEM>(cube, pixels, bands, stride, slib, snorm, srnorm, exact_all, clib,
                                              num_targets, s_score, s_target, c_score, c_target, p0);
}

// The fused form of up to 8 targets at two pixels per lane, held to 128
// VGPRs so its waves fit beside another frame's moments waves on a CU.
template <typename In, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void sam_cem_fused_kernel(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    const double* __restrict__ slib, const float* __restrict__ snorm,
    const double* __restrict__ srnorm, int exact_all,
    const double* __restrict__ clib, int num_targets, float* __restrict__ s_score,
    int* __restrict__ s_target, float* __restrict__ c_score, int* __restrict__ c_target) {
  const std::size_t step = static_cast<std::size_t>(gridDim.x) * blockDim.x * 2;
  for (std::size_t p0 = (blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x) * 2; p0 < pixels;
       p0 += step)
    sam_cem_group<In, NT, 2, true, true>(cube, pixels, bands, stride, slib, snorm, srnorm, exact_all, clib,
                                        num_targets, s_score, s_target, c_score, c_target, p0);
}

// Batched scenes of one target-count class (library stride NT) in one
// persistent launch: the scenes' 256*PPT-pixel blocks are numbered in scene
// order (block_begin prefix), and each workgroup takes every gridDim.x-th
// block, finding its scene by a binary search over the prefix (uniform
// across the workgroup, so it runs on the scalar unit).
// block_scene[b] names the scene of block b. PERSIST: a fixed grid walks the
// blocks with a stride; otherwise one workgroup per block, which keeps the
// single-scene kernel's register budget.
template <typename In, int NT, int PPT, bool DO_SAM, bool DO_CEM, bool PERSIST>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void sam_cem_batch_kernel(const SamSceneDev* __restrict__ scenes,
                                                            const int* __restrict__ block_scene,
                                                            unsigned total_blocks) {
  for (unsigned blk = blockIdx.x; blk < total_blocks; blk += gridDim.x) {
    const SamSceneDev& sc = scenes[block_scene[blk]];
    const std::size_t p0 = (static_cast<std::size_t>(blk - sc.block_begin) * 256 + threadIdx.x) * PPT;
    if (p0 < sc.pixels)
      // A scene without one of the maps still runs the shared loop; its
      // results for the missing detector are not stored.
      sam_cem_group<In, NT, PPT, DO_SAM, DO_CEM>(static_cast<const In*>(sc.cube), sc.pixels, sc.bands, sc.stride,
                                                sc.sam_lib, sc.sam_norms, sc.sam_rnorms, sc.exact_all, sc.cem_lib,
                                                sc.num_targets, sc.sam_score, sc.sam_target, sc.cem_score,
                                                sc.cem_target, p0);
    if constexpr (!PERSIST) break;
  }
}

// Four workgroups per CU on 256 CUs: measured on a 2048x2048x113 frame, a
// grid of this size streams at 5.6 TB/s where one workgroup per pixel group
// block (4096) reaches 5.1. SAM alone at four pixels per lane holds three
// waves per SIMD at 8 targets (155 VGPRs), so its grid there is three
// workgroups per CU.
constexpr std::size_t kSamMaxWgs = 1024;
constexpr std::size_t kSamAloneWgs = 768;

template <typename In, int NT, int PPT, bool S, bool Cm>
hipError_t launch_one(const SamCemArgs& a, hipStream_t stream, std::size_t max_wgs = kSamMaxWgs) {
  const std::size_t groups = div_up(a.pixels, PPT);
  std::size_t blocks = div_up(groups, 256);
  if (blocks > max_wgs) blocks = max_wgs;
  hipLaunchKernelGGL((sam_cem_kernel<In, NT, PPT, S, Cm>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                     static_cast<const In*>(a.cube), a.pixels, a.bands, a.stride, a.sam_lib,
                     a.sam_norms, a.sam_rnorms, a.sam_exact_all, a.cem_lib, a.num_targets,
                     a.sam_score, a.sam_target,
                     a.cem_score, a.cem_target);
  return hipGetLastError();
}

template <typename In, int NT, int PPT>
hipError_t launch_mode(const SamCemArgs& a, hipStream_t stream) {
  const bool s = a.sam_score != nullptr;
  const bool c = a.cem_score != nullptr;
  if (s && c && NT > 16) {  // fused 32-target accumulators would not fit in registers
    const hipError_t e = launch_one<In, NT, PPT, true, false>(a, stream);
    return e != hipSuccess ? e : launch_one<In, NT, PPT, false, true>(a, stream);
  }
  if (s && c) {
    if constexpr (NT <= 8) {
      std::size_t blocks = div_up(div_up(a.pixels, 2), 256);
      if (blocks > kSamMaxWgs) blocks = kSamMaxWgs;
      hipLaunchKernelGGL((sam_cem_fused_kernel<In, NT>), dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream,
                         static_cast<const In*>(a.cube), a.pixels, a.bands, a.stride, a.sam_lib, a.sam_norms,
                         a.sam_rnorms, a.sam_exact_all, a.cem_lib, a.num_targets, a.sam_score, a.sam_target,
                         a.cem_score, a.cem_target);
      return hipGetLastError();
    }
    return launch_one<In, NT, PPT, true, true>(a, stream);
  }
  // SAM alone keeps more planes in flight per wave than CEM (its epilogue
  // keeps a wave off the memory system for a while): four pixels per lane
  // and twelve planes up to 8 targets (2048x2048x113 frame: 345 -> 327 us
  // against two pixels and sixteen planes), two pixels and sixteen planes
  // for 16 and 32 targets, whose accumulators need the registers.
  if (s) {
    if constexpr (NT <= 8)
      return launch_one<In, NT, 4, true, false>(a, stream, NT == 8 ? kSamAloneWgs : kSamMaxWgs);
    return launch_one<In, NT, 2, true, false>(a, stream);
  }
  return launch_one<In, NT, PPT, false, true>(a, stream);
}

template <typename In>
hipError_t launch_typed(const SamCemArgs& a, hipStream_t stream) {
  switch (a.lib_stride) {
    case 1: return launch_mode<In, 1, 4>(a, stream);
    case 2: return launch_mode<In, 2, 4>(a, stream);
    case 4: return launch_mode<In, 4, 4>(a, stream);
    case 8: return launch_mode<In, 8, 4>(a, stream);
    case 16: return launch_mode<In, 16, 2>(a, stream);
    case 32: return launch_mode<In, 32, 2>(a, stream);
  }
  return hipErrorInvalidValue;
}

}  // namespace

template <int NT, int P, bool PERSIST>
void launch_sam_batch_passes(const SamSceneDev* scenes, const int* block_scene, unsigned total_blocks, dim3 grid,
                             hipStream_t stream) {
  const dim3 block(256);
  if constexpr (NT <= 16) {
    hipLaunchKernelGGL((sam_cem_batch_kernel<float, NT, P, true, true, PERSIST>), grid, block, 0, stream, scenes,
                       block_scene, total_blocks);
  } else {
    // 32 targets: SAM and CEM as two passes, the fused accumulators would
    // not fit in registers
    hipLaunchKernelGGL((sam_cem_batch_kernel<float, NT, P, true, false, PERSIST>), grid, block, 0, stream, scenes,
                       block_scene, total_blocks);
    hipLaunchKernelGGL((sam_cem_batch_kernel<float, NT, P, false, true, PERSIST>), grid, block, 0, stream, scenes,
                       block_scene, total_blocks);
  }
}

hipError_t launch_sam_cem_batch(int lib_stride, int ppt, const SamSceneDev* scenes, const int* block_scene,
                                unsigned total_blocks, int workgroups, hipStream_t stream) {
  if (total_blocks == 0) return hipSuccess;
  const bool persist = workgroups > 0;
  const dim3 grid(persist ? static_cast<unsigned>(workgroups) : total_blocks);
#define HSI_B(NT, P)                                                                            \
  if (lib_stride == NT && ppt == P) {                                                           \
    if (persist) launch_sam_batch_passes<NT, P, true>(scenes, block_scene, total_blocks, grid, stream); \
    else launch_sam_batch_passes<NT, P, false>(scenes, block_scene, total_blocks, grid, stream);        \
    return hipGetLastError();                                                                   \
  }
  HSI_B(1, 4) HSI_B(2, 4) HSI_B(4, 4) HSI_B(8, 4) HSI_B(16, 2) HSI_B(32, 2)
  HSI_B(1, 2) HSI_B(2, 2) HSI_B(4, 2) HSI_B(8, 2)
#undef HSI_B
  return hipErrorInvalidValue;
}

int sam_cem_lib_stride(int num_targets) {
  int nt = 1;
  while (nt < num_targets) nt *= 2;
  return nt;
}

hipError_t launch_sam_cem(const SamCemArgs& a, hipStream_t stream) {
  if (!a.sam_score && !a.cem_score) return hipSuccess;
  if (a.dtype == HSI_F16) return launch_typed<_Float16>(a, stream);
  return launch_typed<float>(a, stream);
}

}  // namespace hsi_amd
