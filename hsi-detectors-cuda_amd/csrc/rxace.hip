// rxace.hip - RX and ACE as a whitening GEMM on the gfx950 matrix cores.
//
// For a tile of 16 pixels, Z = L^-1 (X - mean) is a (B x B lower) x (B x 16)
// product. It is computed one 16-band row block at a time,
//     Z_I = sum over J <= I of W_IJ Y_J,
// with v_mfma_f64_16x16x4 (FP64 in, FP64 accumulate: the precision the CPU
// oracle uses in rx_cpu / ace_best_cpu, detectors_cpu.cpp:95-156). Each
// finished Z_I block feeds
//   RX:  sum of squares down the block's rows (||z||^2), and
//   ACE: z . u_t for up to 32 targets, as four more MFMAs per target tile
//        that consume the Z_I accumulator registers directly as the B
//        operand: v_mfma_f64_16x16x4 leaves row l/16 + 4r of Z_I in lane l,
//        register r, so in K-step r lane l holds exactly the B element
//        (k = l/16) of rows {r, 4+r, 8+r, 12+r}, and the target tile's A
//        operand is laid out with the same permutation.
// The whitener blocks and the target tile live in LDS as floats (the
// reference narrows them to float: detector.cpp:83-84) and are widened per
// MFMA; each wave keeps its pixel tile's centred spectrum in registers in the
// MFMA B-operand layout and walks tiles persistently.
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

constexpr int kRxWaves = 4;


// Combine two (score, target) candidates so the larger score wins and, on
// equal scores, the smaller target index - which is what a strict `>` scan
// in target order keeps (ace_best_cpu, detectors_cpu.cpp:147-154).
__device__ __forceinline__ void ace_merge(double& s, int& t, double s2, int t2) {
  if (t2 < 0) return;
  if (t < 0 || s2 > s || (s2 == s && t2 < t)) {
    s = s2;
    t = t2;
  }
}

// Scores tiles first, first + step, ... below the tile count of one scene
// whose whitener blocks, target tiles and mean are already in LDS.
// GATHER: score only the pixels named by a candidate list (detection records
// in raster order, pixel y * width + x), *cand_count of them (the whole frame
// if that exceeds cand_cap); each 16-record slice of the list is one tile.
// Candidate list of the GATHER form (see rxace_tiles).
struct GatherDev {
  const hsi_detection* cand = nullptr;
  const unsigned* count = nullptr;
  unsigned cap = 0;
  int width = 0;
  const float* spectra = nullptr;     // nullable: candidate spectra, row = the record's target
  std::size_t spectra_stride = 0;
};

// F32: the FP32 form - whitening and ACE's products on v_mfma_f32_16x16x4f32
// (FP32 accumulation), whose accumulator holds rows 4g + r in lane group g,
// register r (the FP64 MFMA interleaves them: rows g + 4r), so ACE's target
// operand is read permuted and its target index is 16tt + 4g + r; ||z||^2
// and the ratio stay FP64.
template <typename In, int NB, int TT, bool GATHER, bool F32 = false>
__device__ __forceinline__ void rxace_tiles(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride, int num_targets,
    const float* __restrict__ target_norm2, float* __restrict__ rx_out, float* __restrict__ ace_score,
    int* __restrict__ ace_target, const float* w_lds, const float* u_lds,
    const float* mean_lds, std::size_t first, std::size_t step, std::size_t last, const GatherDev gd) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;   // lane group: K index of the A/B operands
  const int n = lane & 15;   // pixel column within the tile
  std::size_t items = pixels;
  bool gather = GATHER;
  bool from_spectra = false;
  if constexpr (GATHER) {
    // A list that overflowed its capacity means the screen flagged more than
    // it can hold: fall back to scoring the whole frame.
    const unsigned c = *gd.count;
    if (c > gd.cap) gather = false;
    else items = c;
    from_spectra = gather && gd.spectra != nullptr;
  }
  std::size_t tiles = div_up(items, 16);
  if (last < tiles) tiles = last;
  // Pixel of tile element t*16 + n, and whether it exists.
  auto pixel_of = [&](std::size_t t, bool& ok) -> std::size_t {
    const std::size_t i = t * 16 + n;
    ok = i < items;
    if constexpr (GATHER) {
      if (!gather) return i;
      if (!ok) return 0;
      const hsi_detection d = gd.cand[i];
      // a record outside the frame names no pixel
      const std::size_t rows = pixels / static_cast<std::size_t>(gd.width);
      ok = d.x >= 0 && d.x < gd.width && d.y >= 0 && static_cast<std::size_t>(d.y) < rows;
      // with spectra, neither does one whose row lies outside [0, cap)
      if (from_spectra) ok = ok && d.target >= 0 && static_cast<unsigned>(d.target) < gd.cap;
      return ok ? static_cast<std::size_t>(d.y) * static_cast<std::size_t>(gd.width) + static_cast<std::size_t>(d.x)
                : 0;
    } else {
      return i;
    }
  };
  const std::size_t wave_id = first;
  const std::size_t waves = step;

  double uu[TT > 0 ? TT : 1][4];
  if constexpr (TT > 0) {
#pragma unroll
    for (int tt = 0; tt < TT; ++tt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int t = F32 ? 16 * tt + 4 * g + r : 16 * tt + g + 4 * r;
        uu[tt][r] = t < num_targets ? static_cast<double>(target_norm2[t]) : 0.0;
      }
  }

  // Raw inputs of a 16-pixel tile, all issued at once: band 16J + 4kk + g of
  // pixel tile*16 + n (the B-operand layout). The next tile's loads are in
  // flight while the matrix cores work on this one.
  // With candidate spectra, element i of the list reads the spectra row its
  // record names (target; bands contiguous) instead of its pixel in the cube;
  // a row outside [0, cap) names no spectrum and the element is skipped.
  // The loads are unconditional (an out-of-range element reads element 0 and
  // is zeroed at use), so the whole tile is in flight at once and the raw
  // values are converted only where they are used.
  auto load_tile = [&](std::size_t t, In (&xr)[NB][4], bool& ok) {
    bool pix_ok;
    const std::size_t pp = pixel_of(t, pix_ok);
    const In* src = cube;
    std::size_t src_stride = stride, sp = pp;
    if constexpr (GATHER && std::is_same_v<In, float>) {
      if (from_spectra) {
        const int row = pix_ok ? gd.cand[t * 16 + n].target : 0;   // checked by pixel_of
        src = gd.spectra;
        src_stride = 1;
        sp = static_cast<std::size_t>(row) * gd.spectra_stride;
      }
    }
    ok = pix_ok;
#pragma unroll
    for (int J = 0; J < NB; ++J)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int b = 16 * J + 4 * kk + g;
        const std::size_t off = (pix_ok && b < bands) ? static_cast<std::size_t>(b) * src_stride + sp : 0;
        xr[J][kk] = src[off];
      }
  };
  In cur[NB][4], nxt[NB][4];
  bool cur_ok = false, nxt_ok = false;
  if (wave_id < tiles) load_tile(wave_id, cur, cur_ok);

  for (std::size_t tile = wave_id; tile < tiles; tile += waves) {
    // Read the whitener blocks from LDS at each use rather than hoisting them
    // out of the tile loop into registers.
    __asm__ volatile("" ::: "memory");
    bool live;
    const std::size_t p = pixel_of(tile, live);
    if (tile + waves < tiles) load_tile(tile + waves, nxt, nxt_ok);
    // Z = W Y accumulated one input band block J at a time.
    typedef std::conditional_t<F32, f4, d4> Acc;
    typedef std::conditional_t<F32, float, double> Real;
    Acc z[NB];
#pragma unroll
    for (int I = 0; I < NB; ++I) z[I] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int J = 0; J < NB; ++J) {
      Real y[4];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int b = 16 * J + 4 * kk + g;
        // padded bands and pixels read 0; their whitener columns are 0 too
        const Real x = cur_ok ? static_cast<Real>(static_cast<float>(cur[J][kk])) : Real(0);
        y[kk] = (b < bands) ? x - static_cast<Real>(mean_lds[b]) : Real(0);
      }
#pragma unroll
      for (int I = J; I < NB; ++I) {
        const float* blk = w_lds + (I * (I + 1) / 2 + J) * 256;
#pragma unroll
        for (int kk = 0; kk < 4; ++kk) {
          if constexpr (F32)
            z[I] = __builtin_amdgcn_mfma_f32_16x16x4f32(blk[kk * 64 + lane], y[kk], z[I], 0, 0, 0);
          else
            z[I] = __builtin_amdgcn_mfma_f64_16x16x4f64(static_cast<double>(blk[kk * 64 + lane]), y[kk],
                                                        z[I], 0, 0, 0);
        }
      }
      // Keep each band block's whitener loads next to their MFMAs.
      __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int J = 0; J < NB; ++J)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) cur[J][kk] = nxt[J][kk];
    cur_ok = nxt_ok;

    double zz = 0.0;
    Acc zu[TT > 0 ? TT : 1];
#pragma unroll
    for (int tt = 0; tt < (TT > 0 ? TT : 1); ++tt) zu[tt] = Acc{0, 0, 0, 0};
#pragma unroll
    for (int I = 0; I < NB; ++I) {
#pragma unroll
      for (int r = 0; r < 4; ++r) zz += static_cast<double>(z[I][r]) * static_cast<double>(z[I][r]);
      if constexpr (TT > 0) {
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) {
          const float* ub = u_lds + (tt * NB + I) * 256;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            // u_lds holds U[16tt + n][16I + g + 4r] at [r][16g + n]; K-step r
            // of the FP32 form needs row 16I + 4g + r, found at [g][16r + n].
            if constexpr (F32)
              zu[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(ub[g * 64 + 16 * r + n], z[I][r], zu[tt], 0, 0, 0);
            else
              zu[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(static_cast<double>(ub[r * 64 + lane]),
                                                            z[I][r], zu[tt], 0, 0, 0);
          }
        }
      }
    }
    // ||z||^2 for pixel n: the four lane groups hold disjoint rows.
    zz += __shfl_xor(zz, 16, 64);
    zz += __shfl_xor(zz, 32, 64);

    if constexpr (TT > 0) {
      // Lane (g, n) holds z.u for targets 16*tt + g + 4*r of pixel n (FP32
      // form: 16*tt + 4*g + r).
      double best = -1.0;
      int best_t = -1;
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = F32 ? 16 * tt + 4 * g + r : 16 * tt + g + 4 * r;
          const double u2 = uu[tt][r];
          if (t < num_targets && u2 > 0.0 && zz > 0.0) {
            const double zu_t = static_cast<double>(zu[tt][r]);
            const double s = (zu_t * zu_t) / (u2 * zz);
            if (s == s) ace_merge(best, best_t, s, t);
          }
        }
      // Merge across the four lane groups of the same pixel.
#pragma unroll
      for (int off = 16; off <= 32; off <<= 1) {
        const double s2 = __shfl_xor(best, off, 64);
        const int t2 = __shfl_xor(best_t, off, 64);
        ace_merge(best, best_t, s2, t2);
      }
      if (live && g == 0) {
        const double clamped = best < 0.0 ? 0.0 : (1.0 < best ? 1.0 : best);
        if (ace_score) ace_score[p] = static_cast<float>(clamped);
        if (ace_target) ace_target[p] = best_t;
      }
    }
    if (live && g == 0 && rx_out) rx_out[p] = static_cast<float>(zz);
  }
}


template <int NB, int TT>
struct RxLds {
  static constexpr int WFLOATS = NB * (NB + 1) / 2 * 256;
  static constexpr int UFLOATS = (TT > 0 ? TT : 1) * NB * 256;
};

template <int NB, int TT>
__device__ __forceinline__ void rxace_stage(const float* __restrict__ white_blocks, const float* __restrict__ ace_u,
                                            const float* __restrict__ mean, int bands, float* w_lds, float* u_lds,
                                            float* mean_lds) {
  for (int e = threadIdx.x; e < RxLds<NB, TT>::WFLOATS; e += blockDim.x) w_lds[e] = white_blocks[e];
  if constexpr (TT > 0)
    for (int e = threadIdx.x; e < RxLds<NB, TT>::UFLOATS; e += blockDim.x) u_lds[e] = ace_u[e];
  for (int e = threadIdx.x; e < NB * 16; e += blockDim.x) mean_lds[e] = e < bands ? mean[e] : 0.f;
}

// Up to 8 row blocks the registers are held to 256 so two waves share a
// SIMD (the unconstrained allocation took 259 with accumulation registers).
__host__ __device__ constexpr int plain_min_waves(int NB) { return NB <= 8 ? 2 : 1; }

template <typename In, int NB, int TT, bool GATHER, bool F32 = false>
__global__ __launch_bounds__(kRxWaves * 64) __attribute__((amdgpu_waves_per_eu(plain_min_waves(NB)))) void rxace_plain_kernel(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    int num_targets, const float* __restrict__ mean, const float* __restrict__ white_blocks,
    const float* __restrict__ ace_u, const float* __restrict__ target_norm2,
    float* __restrict__ rx_out, float* __restrict__ ace_score, int* __restrict__ ace_target,
    const GatherDev gd) {
  __shared__ float w_lds[RxLds<NB, TT>::WFLOATS];
  __shared__ float u_lds[RxLds<NB, TT>::UFLOATS];
  __shared__ float mean_lds[NB * 16];
  rxace_stage<NB, TT>(white_blocks, ace_u, mean, bands, w_lds, u_lds, mean_lds);
  __syncthreads();
  const std::size_t wave_id = blockIdx.x * static_cast<std::size_t>(kRxWaves) + (threadIdx.x >> 6);
  rxace_tiles<In, NB, TT, GATHER, F32>(cube, pixels, bands, stride, num_targets, target_norm2, rx_out, ace_score,
                                  ace_target, w_lds, u_lds, mean_lds, wave_id,
                                  static_cast<std::size_t>(gridDim.x) * kRxWaves, ~static_cast<std::size_t>(0),
                                  gd);
}

// Batched scenes of one shape class in one launch over work items (a scene
// and a range of its tiles). PERSIST: a grid of fixed size where each
// workgroup walks a contiguous run of items, restaging the scene's whitener
// and targets in LDS when its run crosses into the next scene; otherwise one
// workgroup per item, which keeps the single-scene kernel's register budget
// (the item loop costs registers the largest whiteners do not have).
template <int NB, int TT, bool PERSIST>
__global__ __launch_bounds__(kRxWaves * 64) void rxace_batch_kernel(const RxSceneDev* __restrict__ scenes,
                                                                     const RxWorkItem* __restrict__ work,
                                                                     int num_items) {
  __shared__ float w_lds[RxLds<NB, TT>::WFLOATS];
  __shared__ float u_lds[RxLds<NB, TT>::UFLOATS];
  __shared__ float mean_lds[NB * 16];
  if constexpr (PERSIST) {
    const int per = (num_items + static_cast<int>(gridDim.x) - 1) / static_cast<int>(gridDim.x);
    const int it_end = min(num_items, (static_cast<int>(blockIdx.x) + 1) * per);
    int staged = -1;
    for (int it = static_cast<int>(blockIdx.x) * per; it < it_end; ++it) {
      const RxWorkItem w = work[it];
      const RxSceneDev& sc = scenes[w.scene];
      if (w.scene != staged) {
        __syncthreads();
        rxace_stage<NB, TT>(sc.white_blocks, sc.ace_u, sc.mean, sc.bands, w_lds, u_lds, mean_lds);
        __syncthreads();
        staged = w.scene;
      }
      rxace_tiles<float, NB, TT, false>(static_cast<const float*>(sc.cube), sc.pixels, sc.bands, sc.stride,
                                        sc.num_targets, sc.target_norm2, sc.rx, sc.ace_score, sc.ace_target, w_lds,
                                        u_lds, mean_lds, w.tile_begin + (threadIdx.x >> 6), kRxWaves, w.tile_end,
                                        GatherDev{});
    }
  } else {
    if (static_cast<int>(blockIdx.x) >= num_items) return;
    const RxWorkItem w = work[blockIdx.x];
    const RxSceneDev& sc = scenes[w.scene];
    rxace_stage<NB, TT>(sc.white_blocks, sc.ace_u, sc.mean, sc.bands, w_lds, u_lds, mean_lds);
    __syncthreads();
    rxace_tiles<float, NB, TT, false>(static_cast<const float*>(sc.cube), sc.pixels, sc.bands, sc.stride,
                                      sc.num_targets, sc.target_norm2, sc.rx, sc.ace_score, sc.ace_target, w_lds,
                                      u_lds, mean_lds, w.tile_begin + (threadIdx.x >> 6), kRxWaves, w.tile_end,
                                      GatherDev{});
  }
}

using Gather = GatherDev;

template <typename In, int NB, int TT, bool GATHER, bool F32 = false>
hipError_t launch_plain_t(const RxAceArgs& a, const Gather& gth, hipStream_t stream) {
  // A candidate list's length is known only on the device: size the grid for
  // its capacity and let the kernel stop at the count.
  const std::size_t tiles = div_up(GATHER ? static_cast<std::size_t>(gth.cap) : a.pixels, 16);
  std::size_t wgs = div_up(tiles, kRxWaves);
  // A candidate list is usually far shorter than its capacity: as many
  // workgroups as are resident at once (two per CU up to 8 row blocks, one
  // above) keeps the whitener staging per tile scored low.
  const std::size_t max_wgs = GATHER ? 256 * static_cast<std::size_t>(plain_min_waves(NB)) : 1024;
  if (wgs > max_wgs) wgs = max_wgs;
  if (wgs == 0) wgs = 1;
  hipLaunchKernelGGL((rxace_plain_kernel<In, NB, TT, GATHER, F32>), dim3(static_cast<unsigned>(wgs)),
                     dim3(kRxWaves * 64), 0, stream, static_cast<const In*>(a.cube), a.pixels,
                     a.bands, a.stride, a.num_targets, a.mean, a.white_blocks, a.ace_u,
                     a.target_norm2, a.rx, a.ace_score, a.ace_target, gth);
  return hipGetLastError();
}

template <typename In, int NB, bool GATHER, bool F32>
hipError_t launch_plain_nb(const RxAceArgs& a, const Gather& gth, hipStream_t stream) {
  if (a.num_targets == 0 || (!a.ace_score && !a.ace_target))
    return launch_plain_t<In, NB, 0, GATHER, F32>(a, gth, stream);
  if (a.num_targets <= 16) return launch_plain_t<In, NB, 1, GATHER, F32>(a, gth, stream);
  return launch_plain_t<In, NB, 2, GATHER, F32>(a, gth, stream);
}

template <typename In, bool GATHER, bool F32 = false>
hipError_t launch_plain_typed(const RxAceArgs& a, const Gather& gth, hipStream_t stream) {
  switch (rx_blocks(a.bands)) {
#define HSI_NB(N) case N: return launch_plain_nb<In, N, GATHER, F32>(a, gth, stream);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8)
    HSI_NB(9) HSI_NB(10) HSI_NB(11) HSI_NB(12) HSI_NB(13) HSI_NB(14)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Derived layouts.

__global__ void layouts_kernel(LayoutArgs a) {
  const int n = a.bands;
  const int nb = rx_blocks(n);
  const int npair = nb * (nb + 1) / 2;
  const int T = a.num_targets;
  const int tiles = T > 16 ? 2 : 1;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  const int threads = gridDim.x * blockDim.x;
  // CEM filters, band-major doubles, zero padded to lib_stride.
  if (a.cem_lib) {
    for (int e = tid; e < n * a.lib_stride; e += threads) {
      const int b = e / a.lib_stride;
      const int t = e % a.lib_stride;
      a.cem_lib[e] = t < T ? static_cast<double>(a.cem_weight[static_cast<std::size_t>(t) * n + b]) : 0.0;
    }
  }
  // Whitener blocks in A-operand order: [pair][kk][lane] =
  // W[16I + lane%16][16J + 4kk + lane/16].
  for (int e = tid; e < npair * 256; e += threads) {
    const int q = e / 256;
    const int kk = (e % 256) / 64;
    const int lane = e % 64;
    int I = 0;
    while ((I + 1) * (I + 2) / 2 <= q) ++I;
    const int J = q - I * (I + 1) / 2;
    const int row = 16 * I + (lane & 15);
    const int col = 16 * J + 4 * kk + (lane >> 4);
    a.white_blocks[e] = (row < n && col < n) ? a.whitener[static_cast<std::size_t>(row) * n + col] : 0.f;
  }
  // Whitened targets in the permuted A-operand order: [tt][I][r][lane] =
  // U[16tt + lane%16][16I + lane/16 + 4r].
  if (T > 0 && a.ace_u) {
    for (int e = tid; e < tiles * nb * 256; e += threads) {
      const int tt = e / (nb * 256);
      const int I = (e / 256) % nb;
      const int r = (e % 256) / 64;
      const int lane = e % 64;
      const int t = 16 * tt + (lane & 15);
      const int b = 16 * I + (lane >> 4) + 4 * r;
      a.ace_u[e] = (t < T && b < n) ? a.target_white[static_cast<std::size_t>(t) * n + b] : 0.f;
    }
  }
}

}  // namespace

std::size_t rxace_white_floats(int bands) {
  const int nb = rx_blocks(bands);
  return static_cast<std::size_t>(nb) * (nb + 1) / 2 * 256;
}
std::size_t rxace_target_floats(int bands, int num_targets) {
  return static_cast<std::size_t>(num_targets > 16 ? 2 : 1) * rx_blocks(bands) * 256;
}

namespace {
template <int NB, bool PERSIST>
hipError_t launch_batch_p(int tt, const RxSceneDev* scenes, const RxWorkItem* work, int n, int wgs,
                          hipStream_t stream) {
  const dim3 grid(static_cast<unsigned>(wgs)), block(kRxWaves * 64);
  if (tt == 0) hipLaunchKernelGGL((rxace_batch_kernel<NB, 0, PERSIST>), grid, block, 0, stream, scenes, work, n);
  else if (tt == 1) hipLaunchKernelGGL((rxace_batch_kernel<NB, 1, PERSIST>), grid, block, 0, stream, scenes, work, n);
  else hipLaunchKernelGGL((rxace_batch_kernel<NB, 2, PERSIST>), grid, block, 0, stream, scenes, work, n);
  return hipGetLastError();
}
template <int NB>
hipError_t launch_batch_nb(int tt, const RxSceneDev* scenes, const RxWorkItem* work, int n, int wgs,
                           hipStream_t stream) {
  // workgroups <= 0: one workgroup per item
  if (wgs <= 0) return launch_batch_p<NB, false>(tt, scenes, work, n, n, stream);
  return launch_batch_p<NB, true>(tt, scenes, work, n, wgs, stream);
}
}  // namespace

hipError_t launch_rxace_batch(int nb, int tt, const RxSceneDev* scenes, const RxWorkItem* work, int num_items,
                              int workgroups, hipStream_t stream) {
  if (num_items <= 0) return hipSuccess;
  switch (nb) {
#define HSI_NB(N) case N: return launch_batch_nb<N>(tt, scenes, work, num_items, workgroups, stream);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8)
    HSI_NB(9) HSI_NB(10) HSI_NB(11) HSI_NB(12) HSI_NB(13) HSI_NB(14)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

hipError_t launch_rxace_plain(const RxAceArgs& a, hipStream_t stream) {
  const Gather none;
  if (a.dtype == HSI_F16) return launch_plain_typed<_Float16, false>(a, none, stream);
  return launch_plain_typed<float, false>(a, none, stream);
}

hipError_t launch_rxace_fp32(const RxAceArgs& a, hipStream_t stream) {
  const Gather none;
  if (a.dtype == HSI_F16) return launch_plain_typed<_Float16, false, true>(a, none, stream);
  return launch_plain_typed<float, false, true>(a, none, stream);
}

hipError_t launch_rxace_candidates(const RxAceArgs& a, const hsi_detection* cand, const unsigned* count,
                                   unsigned capacity, int width, const float* spectra,
                                   std::size_t spectra_stride, hipStream_t stream) {
  if (a.dtype != HSI_F32) return hipErrorInvalidValue;  // candidates are rescored from the FP32 cube
  if (spectra && spectra_stride < static_cast<std::size_t>(a.bands)) return hipErrorInvalidValue;
  Gather gth;
  gth.cand = cand;
  gth.count = count;
  gth.cap = capacity;
  gth.width = width;
  gth.spectra = spectra;
  gth.spectra_stride = spectra_stride;
  return launch_plain_typed<float, true>(a, gth, stream);
}

namespace {
__global__ void fill_scores_kernel(float* __restrict__ score, int* __restrict__ target, float* __restrict__ aux,
                                   std::size_t pixels) {
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i >= pixels) return;
  score[i] = 0.f;
  if (target) target[i] = -1;
  if (aux) aux[i] = 0.f;
}

// Four pixels per thread with 16-byte stores, for maps that are all 16-byte
// aligned.
__global__ void fill_scores4_kernel(float* __restrict__ score, int* __restrict__ target, float* __restrict__ aux,
                                    std::size_t quads) {
  using i4v = int __attribute__((ext_vector_type(4)));
  const std::size_t q = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (q >= quads) return;
  reinterpret_cast<f4*>(score)[q] = f4{0.f, 0.f, 0.f, 0.f};
  if (target) reinterpret_cast<i4v*>(target)[q] = i4v{-1, -1, -1, -1};
  if (aux) reinterpret_cast<f4*>(aux)[q] = f4{0.f, 0.f, 0.f, 0.f};
}
}  // namespace

hipError_t launch_fill_scores(float* score, int* target, float* aux, std::size_t pixels, hipStream_t stream) {
  auto aligned = [](const void* p) { return reinterpret_cast<std::uintptr_t>(p) % 16 == 0; };
  std::size_t done = 0;
  if (aligned(score) && aligned(target) && aligned(aux) && pixels >= 4) {
    const std::size_t quads = pixels / 4;
    hipLaunchKernelGGL(fill_scores4_kernel, dim3(static_cast<unsigned>(div_up(quads, 256))), dim3(256), 0, stream,
                       score, target, aux, quads);
    done = 4 * quads;
  }
  if (done < pixels)   // the ragged end, or maps that are not 16-byte aligned
    hipLaunchKernelGGL(fill_scores_kernel, dim3(static_cast<unsigned>(div_up(pixels - done, 256))), dim3(256), 0,
                       stream, score + done, target ? target + done : nullptr, aux ? aux + done : nullptr,
                       pixels - done);
  return hipGetLastError();
}

hipError_t launch_layouts(const LayoutArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(layouts_kernel, dim3(64), dim3(256), 0, stream, a);
  return hipGetLastError();
}

}  // namespace hsi_amd
