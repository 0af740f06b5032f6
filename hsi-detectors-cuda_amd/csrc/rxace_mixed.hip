// rxace_mixed.hip - the precision form of RX/ACE on FP16 matrix cores.
//
// Same whitening GEMM as rxace.hip, Z = L^-1 (X - mean), but on
// v_mfma_f32_16x16x32_f16 (FP16 operands, FP32 accumulate), which issues
// sixteen times the FP64 MFMA rate on gfx950. Precision is recovered by
// splitting both operands into FP16 hi + lo parts and forming
//     W y  ~=  W_hi y_hi + W_hi y_lo + W_lo y_hi
// (22 significant bits per operand; the dropped lo*lo term is ~2^-22
// relative). To keep every part in FP16's normal range the whitener is
// rescaled by powers of two first (exact): columns by s_b so each column's
// largest entry is in [8, 16), giving y'_b = y_b * s_b of order one, then rows
// by r_i likewise; z_i = z''_i / r_i.
//
// ACE's numerator z.u_t is another MFMA per pair of 16-row blocks of Z: the
// FP32 accumulators of blocks 2P and 2P+1 are split into hi/lo FP16 and used
// directly as the B operand (lane (g, n) holds rows 4g..4g+3 of each block,
// eight K values), with the target tile's A operand permuted to match, and
// again three MFMAs (hi*hi, hi*lo, lo*hi) give the split-precision product
// that the ACE ratio needs. ||z||^2 and the ratio are formed in FP64.
//
// Operand layout of v_mfma_f32_16x16x32_f16: lane l supplies row (A) or
// column (B) l % 16 and eight K values selected by l / 16; the result lane l
// holds rows 4*(l/16)..+3 of column l % 16. Only the C layout matters for
// correctness here: A and B always use the same (lane group, element) -> band
// assignment, and a dot product does not depend on the order of its terms.
#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

// K blocks (32 bands) that meet the lower triangle of 16-row block I.
__host__ __device__ constexpr int kblocks_for(int I) { return I / 2 + 1; }
// Number of (I, J) blocks stored before row block I.
__host__ __device__ constexpr int block_prefix(int I) {
  return (I % 2 == 0) ? (I / 2) * (I / 2 + 1) : (I / 2 + 1) * (I / 2 + 1);
}

// Launch shape by row-block count: up to 8 blocks a wave keeps a whole tile
// of loads in flight (4-wave workgroups, several per CU); above that,
// 8-wave workgroups stream one K block ahead. Above 10 blocks the whitener's
// LDS copy leaves room for one workgroup per CU and MFMAs are issued for 2
// row blocks at a time; 8 waves (256 registers each) rather than 16 (128
// each) keeps the 224-band kernels free of spills (2.2x faster).
__host__ __device__ constexpr int mixed_waves(int NB) { return NB > 8 ? 8 : 4; }
__host__ __device__ constexpr int mixed_group(int NB) { return NB > 10 ? 2 : 4; }

__device__ __forceinline__ float to_f(float v) { return v; }
__device__ __forceinline__ float to_f(_Float16 v) { return static_cast<float>(v); }

__device__ __forceinline__ void split8(const float (&v)[8], h8& hi, h8& lo) {
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    const _Float16 h = static_cast<_Float16>(v[i]);
    hi[i] = h;
    lo[i] = static_cast<_Float16>(v[i] - static_cast<float>(h));
  }
}

__device__ __forceinline__ void ace_merge(double& s, int& t, double s2, int t2) {
  if (t2 < 0) return;
  if (t < 0 || s2 > s || (s2 == s && t2 < t)) {
    s = s2;
    t = t2;
  }
}

// One K block (32 bands) of the whitening for the row blocks [I0, I0 + G):
// the three split products are issued term by term across the group, so
// consecutive MFMAs write different accumulators.
template <int NB, int G>
__device__ __forceinline__ void mixed_group_mfma(const h8* __restrict__ w_lds, int lane, int J, int I0,
                                                 const h8& yh, const h8& yl, f4 (&z)[NB]) {
  h8 ah[G], al[G];
#pragma unroll
  for (int k = 0; k < G; ++k) {
    const int I = I0 + k;
    if (I < NB && J < kblocks_for(I)) {
      const h8* blk = w_lds + (block_prefix(I) + J) * 128;
      ah[k] = blk[lane];
      al[k] = blk[64 + lane];
    }
  }
#pragma unroll
  for (int k = 0; k < G; ++k)
    if (I0 + k < NB && J < kblocks_for(I0 + k))
      z[I0 + k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[k], yh, z[I0 + k], 0, 0, 0);
#pragma unroll
  for (int k = 0; k < G; ++k)
    if (I0 + k < NB && J < kblocks_for(I0 + k))
      z[I0 + k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[k], yl, z[I0 + k], 0, 0, 0);
#pragma unroll
  for (int k = 0; k < G; ++k)
    if (I0 + k < NB && J < kblocks_for(I0 + k))
      z[I0 + k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[k], yh, z[I0 + k], 0, 0, 0);
}

template <typename In, int NB, int TT>
__global__ __launch_bounds__(mixed_waves(NB) * 64) void rxace_mixed_kernel(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    int num_targets, const float* __restrict__ mean, const _Float16* __restrict__ white_hl,
    const _Float16* __restrict__ ace_hl, const float* __restrict__ col_scale,
    const float* __restrict__ row_inv, const float* __restrict__ target_inv,
    const float* __restrict__ target_norm2, float* __restrict__ rx_out,
    float* __restrict__ ace_score, int* __restrict__ ace_target) {
  constexpr int WAVES = mixed_waves(NB);
  constexpr int G = mixed_group(NB);
  constexpr int NKB = (NB + 1) / 2;           // 32-band K blocks
  constexpr int PD = NB > 8 ? 1 : NKB;        // K blocks prefetched
  constexpr int NBLK = block_prefix(NB);      // stored (I, J) blocks
  constexpr int NPAIR = (NB + 1) / 2;         // 32-row pairs for ACE
  constexpr int WH = NBLK * 1024;             // halves: hi block then lo block
  constexpr int UH = (TT > 0 ? TT : 1) * NPAIR * 1024;
  __shared__ h8 w_lds[WH / 8];
  __shared__ h8 u_lds[UH / 8];
  __shared__ float mean_lds[NKB * 32];
  __shared__ float scale_lds[NKB * 32];
  __shared__ float rinv_lds[NB * 16];
  // Per-target constants of the ACE epilogue, read per lane there: from LDS
  // they cost no global load per tile.
  __shared__ float tn2_lds[(TT > 0 ? TT : 1) * 16];
  __shared__ float tinv_lds[(TT > 0 ? TT : 1) * 16];

  {
    const h8* src = reinterpret_cast<const h8*>(white_hl);
    for (int e = threadIdx.x; e < WH / 8; e += blockDim.x) w_lds[e] = src[e];
    if constexpr (TT > 0) {
      const h8* us = reinterpret_cast<const h8*>(ace_hl);
      for (int e = threadIdx.x; e < UH / 8; e += blockDim.x) u_lds[e] = us[e];
    }
    for (int e = threadIdx.x; e < NKB * 32; e += blockDim.x) {
      mean_lds[e] = e < bands ? mean[e] : 0.f;
      scale_lds[e] = e < bands ? col_scale[e] : 0.f;
    }
    for (int e = threadIdx.x; e < NB * 16; e += blockDim.x) rinv_lds[e] = row_inv[e];
    if constexpr (TT > 0)
      for (int e = threadIdx.x; e < TT * 16; e += blockDim.x) {
        tn2_lds[e] = e < num_targets ? target_norm2[e] : 0.f;
        tinv_lds[e] = e < num_targets ? target_inv[e] : 0.f;
      }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  const std::size_t tiles = div_up(pixels, 16);
  const std::size_t wave_id = blockIdx.x * static_cast<std::size_t>(WAVES) + (threadIdx.x >> 6);
  const std::size_t waves = static_cast<std::size_t>(gridDim.x) * WAVES;

  // Raw inputs of one K block of a 16-pixel tile: band 32J + 8g + i of pixel
  // tile*16 + n, through a buffer view of the cube. Bands past the last read
  // 0 (beyond the view); pixels past the last read padding or the next
  // plane, which only ever reaches their own (unwritten) output column.
  const __amdgpu_buffer_rsrc_t rsrc = cube_rsrc(cube, static_cast<std::uint32_t>(bands * stride * sizeof(In)));
  const std::uint32_t lane_off = static_cast<std::uint32_t>((static_cast<std::size_t>(8 * g) * stride + n) * sizeof(In));
  auto load_kblock = [&](std::size_t t, int J, float (&xr)[8]) {
    const std::uint32_t base = lane_off + static_cast<std::uint32_t>((t * 16 + static_cast<std::size_t>(32 * J) * stride) * sizeof(In));
#pragma unroll
    for (int i = 0; i < 8; ++i)
      xr[i] = buffer_elem(rsrc, base + static_cast<std::uint32_t>(static_cast<std::size_t>(i) * stride * sizeof(In)), In{});
  };
  // One K block of the whitening: centre and scale, split, and the MFMAs of
  // every row block that reaches it (row blocks below 2J do not).
  auto kblock = [&](int J, const float (&raw)[8], f4 (&z)[NB]) {
    float v[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int b = 32 * J + 8 * g + i;
      v[i] = (raw[i] - mean_lds[b]) * scale_lds[b];  // padding: scale 0
    }
    h8 yh, yl;
    split8(v, yh, yl);
#pragma unroll
    for (int I0 = 2 * J; I0 < NB; I0 += G) {
      mixed_group_mfma<NB, G>(w_lds, lane, J, I0, yh, yl, z);
      // one group's whitener fragments live at a time
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // Small whiteners leave registers for a whole tile in flight (PD = NKB K
  // blocks); large ones stream one K block ahead.
  float cur[PD][8], nxt[PD][8];
  if (wave_id < tiles)
#pragma unroll
    for (int J = 0; J < PD; ++J) load_kblock(wave_id, J, cur[J]);

  for (std::size_t tile = wave_id; tile < tiles; tile += waves) {
    const std::size_t p = tile * 16 + n;
    const bool live = p < pixels;
    f4 z[NB];
#pragma unroll
    for (int I = 0; I < NB; ++I) z[I] = f4{0.f, 0.f, 0.f, 0.f};

    if constexpr (PD == NKB) {
      if (tile + waves < tiles)
#pragma unroll
        for (int J = 0; J < NKB; ++J) load_kblock(tile + waves, J, nxt[J]);
#pragma unroll
      for (int J = 0; J < NKB; ++J) {
        kblock(J, cur[J], z);
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int J = 0; J < NKB; ++J)
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[J][i] = nxt[J][i];
    } else {
#pragma unroll
      for (int J = 0; J < NKB; ++J) {
        if (J + 1 < NKB) load_kblock(tile, J + 1, nxt[0]);
        else if (tile + waves < tiles) load_kblock(tile + waves, 0, nxt[0]);
        kblock(J, cur[0], z);
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[0][i] = nxt[0][i];
        __builtin_amdgcn_sched_barrier(0);
      }
    }

    double zz = 0.0;
#pragma unroll
    for (int I = 0; I < NB; ++I)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const double zi = static_cast<double>(z[I][r]) * rinv_lds[16 * I + 4 * g + r];
        zz += zi * zi;
      }
    zz += __shfl_xor(zz, 16, 64);
    zz += __shfl_xor(zz, 32, 64);

    if constexpr (TT > 0) {
      f4 zu[TT];
#pragma unroll
      for (int tt = 0; tt < TT; ++tt) zu[tt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int P = 0; P < NPAIR; ++P) {
        float v[8];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          v[i] = z[2 * P][i];
          v[4 + i] = (2 * P + 1 < NB) ? z[2 * P + 1 < NB ? 2 * P + 1 : 0][i] : 0.f;
        }
        h8 zh, zl;
        split8(v, zh, zl);
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) {
          const h8* blk = u_lds + (tt * NPAIR + P) * 128;
          const h8 uh = blk[lane];
          const h8 ul = blk[64 + lane];
          zu[tt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(uh, zh, zu[tt], 0, 0, 0);
          zu[tt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(uh, zl, zu[tt], 0, 0, 0);
          zu[tt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ul, zh, zu[tt], 0, 0, 0);
        }
      }
      double best = -1.0;
      int best_t = -1;
#pragma unroll
      for (int tt = 0; tt < TT; ++tt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int t = 16 * tt + 4 * g + r;
          if (t < num_targets) {
            const double u2 = static_cast<double>(tn2_lds[t]);
            if (u2 > 0.0 && zz > 0.0) {
              const double zu_t = static_cast<double>(zu[tt][r]) * tinv_lds[t];
              const double s = (zu_t * zu_t) / (u2 * zz);
              if (s == s) ace_merge(best, best_t, s, t);
            }
          }
        }
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

template <typename In, int NB, int TT>
hipError_t launch_mixed_t(const RxAceArgs& a, hipStream_t stream) {
  if (!fits_buffer(a.bands, a.stride, sizeof(In))) return hipErrorInvalidValue;
  constexpr int WAVES = mixed_waves(NB);
  const std::size_t tiles = div_up(a.pixels, 16);
  std::size_t wgs = div_up(tiles, WAVES);
  const std::size_t cap = NB > 10 ? 256 : 1024;  // one resident workgroup per CU above 10 blocks
  if (wgs > cap) wgs = cap;
  hipLaunchKernelGGL((rxace_mixed_kernel<In, NB, TT>), dim3(static_cast<unsigned>(wgs)),
                     dim3(WAVES * 64), 0, stream, static_cast<const In*>(a.cube), a.pixels,
                     a.bands, a.stride, a.num_targets, a.mean, a.white_hl, a.ace_hl, a.col_scale,
                     a.row_inv, a.target_inv, a.target_norm2, a.rx, a.ace_score, a.ace_target);
  return hipGetLastError();
}

template <typename In, int NB>
hipError_t launch_mixed_nb(const RxAceArgs& a, hipStream_t stream) {
  if (a.num_targets == 0 || (!a.ace_score && !a.ace_target)) return launch_mixed_t<In, NB, 0>(a, stream);
  if (a.num_targets <= 16) return launch_mixed_t<In, NB, 1>(a, stream);
  return launch_mixed_t<In, NB, 2>(a, stream);
}

template <typename In>
hipError_t launch_mixed_typed(const RxAceArgs& a, hipStream_t stream) {
  switch (rx_blocks(a.bands)) {
#define HSI_NB(N) case N: return launch_mixed_nb<In, N>(a, stream);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8)
    HSI_NB(9) HSI_NB(10) HSI_NB(11) HSI_NB(12) HSI_NB(13) HSI_NB(14)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// Layouts. One workgroup: power-of-two scales, then the split blocks.

__device__ __forceinline__ float pow2_for(float m) {
  // 2^(4 - e) where m = f * 2^e, f in [0.5, 1): scales m into [8, 16).
  if (!(m > 0.f)) return 1.f;
  int e = 0;
  (void)frexpf(m, &e);
  return ldexpf(1.f, 4 - e);
}

__global__ __launch_bounds__(1024) void mixed_layouts_kernel(MixedLayoutArgs a) {
  const int n = a.bands;
  const int nb = rx_blocks(n);
  const int T = a.num_targets;
  const int tid = threadIdx.x;
  // Maxima of |x| kept as the bit patterns of non-negative floats, whose
  // unsigned order is the float order, so LDS atomicMax finds them in any
  // order and the result is deterministic.
  __shared__ unsigned cmax[kMaxBands], rmax[kMaxBands], tmax[kMaxTargets];
  __shared__ float cs[kMaxBands];   // column scale s_b (y' = y * s_b)
  __shared__ float rs[kMaxBands];   // row scale r_i
  __shared__ float ts[kMaxTargets]; // target scale c_t
  for (int b = tid; b < kMaxBands; b += blockDim.x) cmax[b] = rmax[b] = 0u;
  for (int t = tid; t < kMaxTargets; t += blockDim.x) tmax[t] = 0u;
  __syncthreads();
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, b = e % n;
    if (b <= i) atomicMax(&cmax[b], __float_as_uint(fabsf(a.whitener[e])));
  }
  __syncthreads();
  // W' = W / s with s = 1 / pow2_for(column max): columns land in [8, 16).
  for (int b = tid; b < kMaxBands; b += blockDim.x) cs[b] = b < n ? 1.f / pow2_for(__uint_as_float(cmax[b])) : 0.f;
  __syncthreads();
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, b = e % n;
    if (b <= i) atomicMax(&rmax[i], __float_as_uint(fabsf(a.whitener[e] / cs[b])));
  }
  __syncthreads();
  for (int i = tid; i < kMaxBands; i += blockDim.x) rs[i] = pow2_for(__uint_as_float(rmax[i]));
  __syncthreads();
  for (int e = tid; e < T * n; e += blockDim.x) {
    const int t = e / n, i = e % n;
    atomicMax(&tmax[t], __float_as_uint(fabsf(a.target_white[e] / rs[i])));
  }
  __syncthreads();
  for (int t = tid; t < kMaxTargets; t += blockDim.x) ts[t] = pow2_for(__uint_as_float(tmax[t]));
  __syncthreads();
  for (int b = tid; b < kMaxBands; b += blockDim.x) {
    a.col_scale[b] = cs[b];
    a.row_inv[b] = 1.f / rs[b];
  }
  for (int t = tid; t < kMaxTargets; t += blockDim.x) a.target_inv[t] = 1.f / ts[t];
  // Whitener blocks: [block][hi|lo][lane][8], value W''[16I + l%16][32J + 8(l/16) + i].
  const int nblk = block_prefix(nb);
  for (int e = tid; e < nblk * 512; e += blockDim.x) {
    const int blk = e / 512;
    const int lane = (e % 512) / 8;
    const int i = e % 8;
    int I = 0;
    while (block_prefix(I + 1) <= blk) ++I;
    const int J = blk - block_prefix(I);
    const int row = 16 * I + (lane & 15);
    const int col = 32 * J + 8 * (lane >> 4) + i;
    float v = 0.f;
    if (row < n && col <= row) v = a.whitener[static_cast<std::size_t>(row) * n + col] / cs[col] * rs[row];
    const _Float16 h = static_cast<_Float16>(v);
    a.white_hl[blk * 1024 + lane * 8 + i] = h;
    a.white_hl[blk * 1024 + 512 + lane * 8 + i] = static_cast<_Float16>(v - static_cast<float>(h));
  }
  // Target blocks: [tt][P][hi|lo][lane][8], value c_t * U[t][band] / r_band with
  // band = 32P + 4(l/16) + i for i < 4 and 32P + 16 + 4(l/16) + (i-4) above.
  if (T > 0 && a.ace_hl) {
    const int npair = (nb + 1) / 2;
    const int tiles = T > 16 ? 2 : 1;
    for (int e = tid; e < tiles * npair * 512; e += blockDim.x) {
      const int tt = e / (npair * 512);
      const int P = (e / 512) % npair;
      const int lane = (e % 512) / 8;
      const int i = e % 8;
      const int t = 16 * tt + (lane & 15);
      const int band = 32 * P + (i < 4 ? 4 * (lane >> 4) + i : 16 + 4 * (lane >> 4) + (i - 4));
      float v = 0.f;
      if (t < T && band < n) v = a.target_white[static_cast<std::size_t>(t) * n + band] / rs[band] * ts[t];
      const _Float16 h = static_cast<_Float16>(v);
      const int base = (tt * npair + P) * 1024;
      a.ace_hl[base + lane * 8 + i] = h;
      a.ace_hl[base + 512 + lane * 8 + i] = static_cast<_Float16>(v - static_cast<float>(h));
    }
  }
}

}  // namespace

std::size_t mixed_white_halves(int bands) {
  return static_cast<std::size_t>(block_prefix(rx_blocks(bands))) * 1024;
}
std::size_t mixed_target_halves(int bands, int num_targets) {
  return static_cast<std::size_t>(num_targets > 16 ? 2 : 1) * ((rx_blocks(bands) + 1) / 2) * 1024;
}

hipError_t launch_rxace_mixed(const RxAceArgs& a, hipStream_t stream) {
  if (a.dtype == HSI_F16) return launch_mixed_typed<_Float16>(a, stream);
  return launch_mixed_typed<float>(a, stream);
}

hipError_t launch_mixed_layouts(const MixedLayoutArgs& a, hipStream_t stream) {
  hipLaunchKernelGGL(mixed_layouts_kernel, dim3(1), dim3(1024), 0, stream, a);
  return hipGetLastError();
}

}  // namespace hsi_amd
