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
#include <cstdlib>

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

// Launch shapes. S = 1: one 16-pixel tile per wave step; up to 8 row blocks
// a wave keeps a whole tile of loads in flight (4-wave workgroups, several
// per CU); above that, 8-wave workgroups stream one K block ahead. Above 10
// blocks the whitener's LDS copy leaves room for one workgroup per CU and
// MFMAs are issued for 2 row blocks at a time; 8 waves (256 registers each)
// rather than 16 (128 each) keeps the 224-band kernels free of spills (2.2x
// faster). S = 2: two tiles per step, each whitener fragment read from LDS
// serving both, S = 1's waves, loads streamed one K block ahead, two row
// blocks per fragment group (four accumulators between dependent MFMAs:
// 224-band Q16 ACE 254 -> 249 us against one row block, the same results
// bit for bit). At 1024 x 1024 px S = 2 is the faster form
// except for ACE on FP32 cubes below 10 row blocks (224 bands, Q16: ACE 258
// against 296 us, RX 205 against 248 us; 113 bands, FP32 ACE: 180 against
// 166 us). A four-wave two-tile form with the whole next step's loads in
// flight was slower at both band counts.
__host__ __device__ constexpr int mixed_tp(int S) { return S == 1 ? 1 : 2; }
__host__ __device__ constexpr int mixed_waves(int NB, int S) { return NB > 8 ? 8 : 4; }
__host__ __device__ constexpr int mixed_group(int NB, int S) { return S == 2 ? 2 : (NB > 10 ? 2 : 4); }
__host__ __device__ constexpr int mixed_prefetch(int NB, int S) { return S == 2 || NB > 8 ? 1 : (NB + 1) / 2; }

// Prefetched raw inputs of one band, kept packed until use: one element
// (TP = 1), two FP32 pixels (float2), or two 16-bit pixels in one dword.
template <typename In, int TP> struct RawT { typedef float type; };
template <> struct RawT<float, 2> { typedef float2 type; };
template <> struct RawT<_Float16, 2> { typedef unsigned type; };
template <> struct RawT<q16_t, 2> { typedef unsigned type; };

template <typename In, int TP>
__device__ __forceinline__ typename RawT<In, TP>::type raw_load(__amdgpu_buffer_rsrc_t r, std::uint32_t off) {
  if constexpr (TP == 1) return buffer_elem(r, off, In{});
  else if constexpr (sizeof(In) == 4) return buffer_pair(r, off, In{});
  else return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}

// Tile t's value (a Q16 code stays a code here).
template <typename In, int TP>
__device__ __forceinline__ float raw_unpack(typename RawT<In, TP>::type v, int t) {
  if constexpr (TP == 1) {
    return v;
  } else if constexpr (sizeof(In) == 4) {
    return t == 0 ? v.x : v.y;
  } else {
    const std::uint16_t e = static_cast<std::uint16_t>(t == 0 ? (v & 0xffffu) : (v >> 16));
    if constexpr (is_q16<In>::value) return static_cast<float>(e);
    else return static_cast<float>(__builtin_bit_cast(_Float16, e));
  }
}

__device__ __forceinline__ void ace_merge(double& s, int& t, double s2, int t2) {
  if (t2 < 0) return;
  if (t < 0 || s2 > s || (s2 == s && t2 < t)) {
    s = s2;
    t = t2;
  }
}

// One K block (32 bands) of the whitening for the row blocks [I0, I0 + G)
// of TP pixel tiles: each whitener fragment is read from LDS once and used
// for every tile, and the three split products are issued term by term
// across the group and the tiles, so consecutive MFMAs write different
// accumulators.
template <int NB, int G, int TP>
__device__ __forceinline__ void mixed_group_mfma(const h8* __restrict__ w_lds, int lane, int J, int I0,
                                                 const h8 (&yh)[TP], const h8 (&yl)[TP], f4 (&z)[TP][NB]) {
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
#pragma unroll
    for (int t = 0; t < TP; ++t)
      if (I0 + k < NB && J < kblocks_for(I0 + k))
        z[t][I0 + k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[k], yh[t], z[t][I0 + k], 0, 0, 0);
#pragma unroll
  for (int k = 0; k < G; ++k)
#pragma unroll
    for (int t = 0; t < TP; ++t)
      if (I0 + k < NB && J < kblocks_for(I0 + k))
        z[t][I0 + k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[k], yl[t], z[t][I0 + k], 0, 0, 0);
#pragma unroll
  for (int k = 0; k < G; ++k)
#pragma unroll
    for (int t = 0; t < TP; ++t)
      if (I0 + k < NB && J < kblocks_for(I0 + k))
        z[t][I0 + k] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[k], yh[t], z[t][I0 + k], 0, 0, 0);
}

// Candidate form: the elements scored are the first *count records of a
// candidate list, each read from the spectra row its record names (target
// field; bands contiguous) and written at its pixel y * width + x. A list
// longer than its capacity overflowed: the whole frame is scored from the
// cube instead (decided here, from the count, with no host round trip).
struct MixedGather {
  const hsi_detection* cand = nullptr;
  const unsigned* count = nullptr;
  unsigned cap = 0;
  int width = 0;
  const float* spectra = nullptr;
  std::size_t spectra_stride = 0;   // floats per row
};

// A wave scores TP tiles of 16 pixels per step. TP = 1: lane (g, n) holds
// pixel n of the tile. TP = 2: a step covers 32 consecutive pixels and lane
// (g, n) loads pixels 2n and 2n + 1 with one load per band - tile 0 holds
// the even pixels, tile 1 the odd ones - so each whitener fragment read
// from LDS serves 32 pixels and each load moves twice the bytes.
// In: float, _Float16 or q16_t (decoded per band before the centring).
template <typename In, int NB, int TT, int S, bool GATHER>
__global__ __launch_bounds__(mixed_waves(NB, S) * 64) void rxace_mixed_kernel(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    int num_targets, const float* __restrict__ mean, const _Float16* __restrict__ white_hl,
    const _Float16* __restrict__ ace_hl, const float* __restrict__ col_scale,
    const float* __restrict__ row_inv, const float* __restrict__ target_inv,
    const float* __restrict__ target_norm2, const float* __restrict__ q_offset,
    const float* __restrict__ q_scale, float* __restrict__ rx_out,
    float* __restrict__ ace_score, int* __restrict__ ace_target, const MixedGather mg) {
  constexpr int TP = mixed_tp(S);
  static_assert(!GATHER || (TP == 1 && sizeof(In) == 4), "the candidate form reads FP32 rows one tile a step");
  constexpr int WAVES = mixed_waves(NB, S);
  constexpr int G = mixed_group(NB, S);
  constexpr int NKB = (NB + 1) / 2;              // 32-band K blocks
  constexpr int PD = mixed_prefetch(NB, S);      // K blocks prefetched
  constexpr int NBLK = block_prefix(NB);         // stored (I, J) blocks
  constexpr int NPAIR = (NB + 1) / 2;            // 32-row pairs for ACE
  constexpr int WH = NBLK * 1024;                // halves: hi block then lo block
  constexpr int UH = (TT > 0 ? TT : 1) * NPAIR * 1024;
  constexpr bool Q16 = is_q16<In>::value;
  __shared__ h8 w_lds[WH / 8];
  __shared__ h8 u_lds[UH / 8];
  // Centring and column scaling as one FMA per element, y' = fma(ya, raw, yc):
  // ya = s_b, yc = -mean_b s_b for F32/F16 (s_b is a power of two, so this
  // equals (x - mean_b) s_b bit for bit); for Q16, x = scale_b q + offset_b
  // folds in too: ya = scale_b s_b, yc = (offset_b - mean_b) s_b.
  __shared__ float ya_lds[NKB * 32];
  __shared__ float yc_lds[NKB * 32];
  __shared__ float rinv_lds[NB * 16];
  // Per-target constants of the ACE epilogue, read per lane there: from LDS
  // they cost no global load per tile.
  __shared__ double iu2_lds[(TT > 0 ? TT : 1) * 16];   // 1 / ||u_t||^2, 0 when not positive
  __shared__ float tinv_lds[(TT > 0 ? TT : 1) * 16];

  {
    const h8* src = reinterpret_cast<const h8*>(white_hl);
    for (int e = threadIdx.x; e < WH / 8; e += blockDim.x) w_lds[e] = src[e];
    if constexpr (TT > 0) {
      const h8* us = reinterpret_cast<const h8*>(ace_hl);
      for (int e = threadIdx.x; e < UH / 8; e += blockDim.x) u_lds[e] = us[e];
    }
    for (int e = threadIdx.x; e < NKB * 32; e += blockDim.x) {
      float a = 0.f, c = 0.f;  // padding bands: y' = 0
      if (e < bands) {
        const float sc = col_scale[e];
        if constexpr (Q16) {
          a = q_scale[e] * sc;
          c = (q_offset[e] - mean[e]) * sc;
        } else {
          a = sc;
          c = -mean[e] * sc;
        }
      }
      ya_lds[e] = a;
      yc_lds[e] = c;
    }
    for (int e = threadIdx.x; e < NB * 16; e += blockDim.x) rinv_lds[e] = row_inv[e];
    if constexpr (TT > 0)
      for (int e = threadIdx.x; e < TT * 16; e += blockDim.x) {
        const double u2 = e < num_targets ? static_cast<double>(target_norm2[e]) : 0.0;
        iu2_lds[e] = u2 > 0.0 ? 1.0 / u2 : 0.0;
        tinv_lds[e] = e < num_targets ? target_inv[e] : 0.f;
      }
  }
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  constexpr int SPAN = 16 * TP;  // pixels per wave step
  std::size_t items = pixels;    // elements scored
  bool rows = false;             // candidate form: elements are spectra rows
  if constexpr (GATHER) {
    const unsigned c = *mg.count;
    if (c <= mg.cap) {
      items = c;
      rows = true;
    }
  }
  const std::size_t steps = div_up(items, SPAN);
  // The wave index through readfirstlane, so the step and every address
  // derived from it stay in scalar registers.
  const unsigned wave_in_wg = static_cast<unsigned>(__builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6)));
  const std::size_t wave_id = blockIdx.x * static_cast<std::size_t>(WAVES) + wave_in_wg;
  const std::size_t waves = static_cast<std::size_t>(gridDim.x) * WAVES;

  // Raw inputs of one K block of a step: band 32J + 8g + i of pixel
  // step*SPAN + n (TP = 1) or step*SPAN + 2n + t (TP = 2, tile t). Each K
  // block is read through its own buffer view, starting at band 32J of the
  // step's first pixel (a wave-uniform base built in scalar registers), so
  // a lane's eight offsets are fixed for the whole kernel. Bands past the
  // last read 0 (beyond the view); pixels past the last read padding or the
  // next plane, which only ever reaches their own (unwritten) output column.
  const std::size_t cube_bytes = static_cast<std::size_t>(bands) * stride * sizeof(In);
  std::uint32_t lane_off[8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
    lane_off[i] = static_cast<std::uint32_t>(
        (static_cast<std::size_t>(8 * g + i) * stride + (TP == 1 ? n : 2 * n)) * sizeof(In));
  typedef typename RawT<In, TP>::type Raw;
  // Candidate form: element t * 16 + n's spectra row (byte offset), pixel
  // and whether it is scored (a record outside the frame, or naming a row
  // outside [0, cap), is not).
  struct Elem {
    std::uint32_t off;
    std::size_t pix;
    bool ok;
  };
  const std::size_t frame_rows = GATHER ? pixels / static_cast<std::size_t>(mg.width > 0 ? mg.width : 1) : 0;
  auto elem_of = [&](std::size_t t) -> Elem {
    Elem e{0u, t * 16 + n, false};
    if (!rows) {
      e.ok = e.pix < items;
      return e;
    }
    const std::size_t i = t * 16 + n;
    if (i < items) {
      const hsi_detection d = mg.cand[i];
      if (d.x >= 0 && d.x < mg.width && d.y >= 0 && static_cast<std::size_t>(d.y) < frame_rows && d.target >= 0 &&
          static_cast<unsigned>(d.target) < mg.cap) {
        e.ok = true;
        e.pix = static_cast<std::size_t>(d.y) * static_cast<std::size_t>(mg.width) + static_cast<std::size_t>(d.x);
        e.off = static_cast<std::uint32_t>(static_cast<std::size_t>(d.target) * mg.spectra_stride * sizeof(float));
      }
    }
    return e;
  };
  const std::uint32_t spectra_bytes =
      GATHER ? static_cast<std::uint32_t>(static_cast<std::size_t>(mg.cap) * mg.spectra_stride * sizeof(float)) : 0u;
  auto load_kblock = [&](std::size_t t, int J, Raw (&xr)[8], std::uint32_t row_off) {
    if constexpr (GATHER) {
      if (rows) {
        // Eight contiguous bands of the element's row, two 16-byte loads;
        // past the buffer they read 0.
        const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float*>(mg.spectra), 0, static_cast<int>(spectra_bytes), 0x00020000);
        const std::uint32_t o = row_off + static_cast<std::uint32_t>((32 * J + 8 * g) * sizeof(float));
        const float4 a0 = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o, 0, 0));
        const float4 a1 = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, o + 16, 0, 0));
        xr[0] = a0.x; xr[1] = a0.y; xr[2] = a0.z; xr[3] = a0.w;
        xr[4] = a1.x; xr[5] = a1.y; xr[6] = a1.z; xr[7] = a1.w;
        return;
      }
    }
    (void)row_off;
    const std::size_t base = (t * SPAN + static_cast<std::size_t>(32 * J) * stride) * sizeof(In);
    const bool inside = base < cube_bytes;  // else an empty view: every load reads 0
    const std::uint32_t rest = inside ? static_cast<std::uint32_t>(cube_bytes - base) : 0u;
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<char*>(reinterpret_cast<const char*>(cube)) + (inside ? base : 0), 0, static_cast<int>(rest),
        0x00020000);
#pragma unroll
    for (int i = 0; i < 8; ++i) xr[i] = raw_load<In, TP>(r, lane_off[i]);
  };
  // One K block of the whitening: decode, centre and scale, split, and the
  // MFMAs of every row block that reaches it (row blocks below 2J do not).
  auto kblock = [&](int J, const Raw (&raw)[8], f4 (&z)[TP][NB]) {
    h8 yh[TP], yl[TP];
#pragma unroll
    for (int t = 0; t < TP; ++t) {
      float v[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int b = 32 * J + 8 * g + i;
        float x = raw_unpack<In, TP>(raw[i], t);
        // A row's bands past the last may hold anything (the next row, or
        // padding): drop them before the zero column scale meets them.
        if constexpr (GATHER)
          if (J == NKB - 1 && b >= bands) x = 0.f;
        v[i] = __builtin_fmaf(ya_lds[b], x, yc_lds[b]);
      }
      split8(v, yh[t], yl[t]);
    }
#pragma unroll
    for (int I0 = 2 * J; I0 < NB; I0 += G) {
      mixed_group_mfma<NB, G, TP>(w_lds, lane, J, I0, yh, yl, z);
      // one group's whitener fragments live at a time
      __builtin_amdgcn_sched_barrier(0);
    }
  };
  // Small whiteners at one tile per step leave registers for a whole tile in
  // flight (PD = NKB K blocks); otherwise the loads stream one K block ahead.
  Raw cur[PD][8], nxt[PD][8];
  Elem ecur{0u, 0, false}, enxt{0u, 0, false};
  if (wave_id < steps) {
    if constexpr (GATHER) ecur = elem_of(wave_id);
#pragma unroll
    for (int J = 0; J < PD; ++J) load_kblock(wave_id, J, cur[J], ecur.off);
  }

  for (std::size_t step = wave_id; step < steps; step += waves) {
    f4 z[TP][NB];
#pragma unroll
    for (int t = 0; t < TP; ++t)
#pragma unroll
      for (int I = 0; I < NB; ++I) z[t][I] = f4{0.f, 0.f, 0.f, 0.f};

    if constexpr (PD == NKB) {
      if (step + waves < steps) {
        if constexpr (GATHER) enxt = elem_of(step + waves);
#pragma unroll
        for (int J = 0; J < NKB; ++J) load_kblock(step + waves, J, nxt[J], enxt.off);
      }
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
        if (J + 1 < NKB) {
          load_kblock(step, J + 1, nxt[0], ecur.off);
        } else if (step + waves < steps) {
          if constexpr (GATHER) enxt = elem_of(step + waves);
          load_kblock(step + waves, 0, nxt[0], enxt.off);
        }
        kblock(J, cur[0], z);
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[0][i] = nxt[0][i];
        __builtin_amdgcn_sched_barrier(0);
      }
    }

    // ||z||^2: z_i = z''_i / r_i is exact (r_i a power of two); each lane
    // sums its 4 NB squares in FP32 (at most 56 terms, <= 4e-6 relative),
    // the four lanes of a pixel add in FP64.
    double zz[TP];
#pragma unroll
    for (int t = 0; t < TP; ++t) {
      float part = 0.f;
#pragma unroll
      for (int I = 0; I < NB; ++I)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float zi = z[t][I][r] * rinv_lds[16 * I + 4 * g + r];
          part = __builtin_fmaf(zi, zi, part);
        }
      zz[t] = static_cast<double>(part);
      zz[t] += __shfl_xor(zz[t], 16, 64);
      zz[t] += __shfl_xor(zz[t], 32, 64);
    }

    double best[TP];
    int best_t[TP];
    if constexpr (TT > 0) {
      f4 zu[TP][TT];
#pragma unroll
      for (int t = 0; t < TP; ++t)
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) zu[t][tt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int P = 0; P < NPAIR; ++P) {
        h8 zh[TP], zl[TP];
#pragma unroll
        for (int t = 0; t < TP; ++t) {
          float v[8];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            v[i] = z[t][2 * P][i];
            v[4 + i] = (2 * P + 1 < NB) ? z[t][2 * P + 1 < NB ? 2 * P + 1 : 0][i] : 0.f;
          }
          split8(v, zh[t], zl[t]);
        }
#pragma unroll
        for (int tt = 0; tt < TT; ++tt) {
          const h8* blk = u_lds + (tt * NPAIR + P) * 128;
          const h8 uh = blk[lane];
          const h8 ul = blk[64 + lane];
#pragma unroll
          for (int t = 0; t < TP; ++t) {
            zu[t][tt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(uh, zh[t], zu[t][tt], 0, 0, 0);
            zu[t][tt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(uh, zl[t], zu[t][tt], 0, 0, 0);
            zu[t][tt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ul, zh[t], zu[t][tt], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int t = 0; t < TP; ++t) {
        // The pixel's ||z||^2 is common to its targets, so they are ranked
        // by (z.u_t)^2 / ||u_t||^2 and only the best is divided by it.
        best[t] = -1.0;
        best_t[t] = -1;
        if (zz[t] > 0.0) {
#pragma unroll
          for (int tt = 0; tt < TT; ++tt)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
              const int tg = 16 * tt + 4 * g + r;
              if (tg < num_targets) {
                const double iu2 = iu2_lds[tg];
                if (iu2 > 0.0) {
                  const double zu_t = static_cast<double>(zu[t][tt][r]) * tinv_lds[tg];
                  const double q = (zu_t * zu_t) * iu2;
                  if (q == q) ace_merge(best[t], best_t[t], q, tg);
                }
              }
            }
        }
#pragma unroll
        for (int off = 16; off <= 32; off <<= 1) {
          const double s2 = __shfl_xor(best[t], off, 64);
          const int t2 = __shfl_xor(best_t[t], off, 64);
          ace_merge(best[t], best_t[t], s2, t2);
        }
        const double sc = best_t[t] >= 0 ? best[t] / zz[t] : 0.0;
        best[t] = sc < 0.0 ? 0.0 : (1.0 < sc ? 1.0 : sc);
      }
    }
    if constexpr (GATHER) {
      if (g == 0 && ecur.ok) {
        const std::size_t p = ecur.pix;
        if constexpr (TT > 0) {
          if (ace_score) ace_score[p] = static_cast<float>(best[0]);
          if (ace_target) ace_target[p] = best_t[0];
        }
        if (rx_out) rx_out[p] = static_cast<float>(zz[0]);
      }
      ecur = enxt;
    } else if (g == 0) {
      const std::size_t p0 = step * SPAN + (TP == 1 ? n : 2 * n);
#pragma unroll
      for (int t = 0; t < TP; ++t) {
        const std::size_t p = p0 + t;
        if (p < pixels) {
          if constexpr (TT > 0) {
            if (ace_score) ace_score[p] = static_cast<float>(best[t]);
            if (ace_target) ace_target[p] = best_t[t];
          }
          if (rx_out) rx_out[p] = static_cast<float>(zz[t]);
        }
      }
    }
  }
}

// Launch shape by row-block count, element size and ACE; HSI_MIXED_SHAPE=1|2
// overrides it (a tuning knob: the shapes differ only in FP32 rounding of
// the accumulation, at most 4e-7 relative at 1024 x 1024 x 224, and each is
// deterministic run to run).
inline int mixed_shape(int nb, std::size_t elem_bytes, bool ace) {
  static const int forced = [] {
    const char* e = std::getenv("HSI_MIXED_SHAPE");
    return e ? std::atoi(e) : 0;
  }();
  if (forced == 1 || forced == 2) return forced;
  return (nb < 10 && elem_bytes == 4 && ace) ? 1 : 2;
}

template <typename In, int NB, int TT, int S, bool GATHER = false>
hipError_t launch_mixed_tp(const RxAceArgs& a, hipStream_t stream, const MixedGather& mg = MixedGather{}) {
  if (!fits_buffer(a.bands, a.stride, sizeof(In))) return hipErrorInvalidValue;
  constexpr int WAVES = mixed_waves(NB, S);
  // A candidate list's length is known only on the device: size the grid for
  // its capacity (the kernel walks its elements, or the frame, grid-stride).
  const std::size_t steps = div_up(GATHER ? static_cast<std::size_t>(mg.cap) : a.pixels, 16 * mixed_tp(S));
  std::size_t wgs = div_up(steps, WAVES);
  const std::size_t cap = NB > 10 ? 256 : 1024;  // one resident workgroup per CU above 10 blocks
  if (wgs > cap) wgs = cap;
  if (wgs == 0) wgs = 1;
  hipLaunchKernelGGL((rxace_mixed_kernel<In, NB, TT, S, GATHER>), dim3(static_cast<unsigned>(wgs)),
                     dim3(WAVES * 64), 0, stream, static_cast<const In*>(a.cube), a.pixels,
                     a.bands, a.stride, a.num_targets, a.mean, a.white_hl, a.ace_hl, a.col_scale,
                     a.row_inv, a.target_inv, a.target_norm2, a.q_offset, a.q_scale, a.rx, a.ace_score,
                     a.ace_target, mg);
  return hipGetLastError();
}

template <typename In, int NB, int TT>
hipError_t launch_mixed_t(const RxAceArgs& a, hipStream_t stream) {
  return mixed_shape(NB, sizeof(In), TT > 0) == 2 ? launch_mixed_tp<In, NB, TT, 2>(a, stream)
                                                  : launch_mixed_tp<In, NB, TT, 1>(a, stream);
}

template <typename In, int NB>
hipError_t launch_mixed_nb(const RxAceArgs& a, hipStream_t stream) {
  if (a.num_targets == 0 || (!a.ace_score && !a.ace_target)) return launch_mixed_t<In, NB, 0>(a, stream);
  if (a.num_targets <= 16) return launch_mixed_t<In, NB, 1>(a, stream);
  return launch_mixed_t<In, NB, 2>(a, stream);
}

template <int NB>
hipError_t launch_mixed_gather_nb(const RxAceArgs& a, const MixedGather& mg, hipStream_t stream) {
  if (a.num_targets == 0 || (!a.ace_score && !a.ace_target))
    return launch_mixed_tp<float, NB, 0, 1, true>(a, stream, mg);
  if (a.num_targets <= 16) return launch_mixed_tp<float, NB, 1, 1, true>(a, stream, mg);
  return launch_mixed_tp<float, NB, 2, 1, true>(a, stream, mg);
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
// Layouts: the power-of-two scales in one workgroup (coalesced passes over
// the whitener, maxima reduced in registers, LDS and lane shuffles, no
// atomics), then the split blocks by a grid of workgroups, one element per
// thread.

__device__ __forceinline__ float pow2_for(float m) {
  // 2^(4 - e) where m = f * 2^e, f in [0.5, 1): scales m into [8, 16).
  if (!(m > 0.f)) return 1.f;
  int e = 0;
  (void)frexpf(m, &e);
  return ldexpf(1.f, 4 - e);
}

__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

constexpr int kScaleThreads = 1024;

// Column scales s_b = 1 / pow2_for(max_i |W[i][b]|), row scales r_i =
// pow2_for(max_b |W[i][b]| / s_b), target scales c_t = pow2_for(max_i
// |U[t][i]| / r_i); written as col_scale, row_inv = 1 / r, target_inv = 1 / c.
// Maxima of non-negative floats are exact in any order, so the result does
// not depend on the reduction's shape.
__global__ __launch_bounds__(kScaleThreads) void mixed_scales_kernel(MixedLayoutArgs a) {
  const int n = a.bands;
  const int T = a.num_targets;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  constexpr int kWaves = kScaleThreads / 64;
  __shared__ float part[kScaleThreads];
  __shared__ float cs[kMaxBands];
  __shared__ float rs[kMaxBands];
  // Columns: thread t takes column t % cols of rows t / cols, + phases, ...
  const int cols = n <= 128 ? 128 : 256;
  const int phases = kScaleThreads / cols;
  {
    const int b = tid % cols;
    float m = 0.f;
    // unconditional loads (b < n, i < n), so they are all in flight at once
    if (b < n)
#pragma unroll 8
      for (int i = tid / cols; i < n; i += phases) {
        const float v = fabsf(a.whitener[static_cast<std::size_t>(i) * n + b]);
        m = b <= i ? fmaxf(m, v) : m;
      }
    part[tid] = m;
  }
  __syncthreads();
  for (int b = tid; b < kMaxBands; b += kScaleThreads) {
    float m = 0.f;
    if (b < n)
      for (int ph = 0; ph < phases; ++ph) m = fmaxf(m, part[ph * cols + b]);
    cs[b] = b < n ? 1.f / pow2_for(m) : 0.f;
  }
  __syncthreads();
  // Rows: a wave per row, lanes along it.
  for (int i = wave; i < kMaxBands; i += kWaves) {
    float m = 0.f;
    if (i < n)
      for (int b = lane; b < n; b += 64) {
        const float v = fabsf(a.whitener[static_cast<std::size_t>(i) * n + b] / cs[b]);
        m = b <= i ? fmaxf(m, v) : m;
      }
    m = wave_max(m);
    if (lane == 0) rs[i] = pow2_for(m);
  }
  __syncthreads();
  for (int b = tid; b < kMaxBands; b += kScaleThreads) {
    a.col_scale[b] = cs[b];
    a.row_inv[b] = 1.f / rs[b];
  }
  for (int t = wave; t < kMaxTargets; t += kWaves) {
    float m = 0.f;
    if (t < T)
      for (int i = lane; i < n; i += 64) m = fmaxf(m, fabsf(a.target_white[static_cast<std::size_t>(t) * n + i] / rs[i]));
    m = wave_max(m);
    if (lane == 0) a.target_inv[t] = 1.f / pow2_for(m);
  }
}

// Whitener blocks: [block][hi|lo][lane][8], value W''[16I + l%16][32J + 8(l/16) + i]
// with W'' = W / s (columns) * r (rows). Target blocks: [tt][P][hi|lo][lane][8],
// value c_t * U[t][band] / r_band with band = 32P + 4(l/16) + i for i < 4 and
// 32P + 16 + 4(l/16) + (i-4) above. One element (its hi and lo halves) per
// thread over both tables.
__global__ __launch_bounds__(256) void mixed_blocks_kernel(MixedLayoutArgs a, int nwhite, int ntarget) {
  const int n = a.bands;
  const int T = a.num_targets;
  const int e0 = blockIdx.x * blockDim.x + threadIdx.x;
  if (e0 < nwhite) {
    const int e = e0;
    const int blk = e / 512;
    const int lane = (e % 512) / 8;
    const int i = e % 8;
    int I = 0;
    while (block_prefix(I + 1) <= blk) ++I;
    const int J = blk - block_prefix(I);
    const int row = 16 * I + (lane & 15);
    const int col = 32 * J + 8 * (lane >> 4) + i;
    float v = 0.f;
    if (row < n && col <= row)
      v = a.whitener[static_cast<std::size_t>(row) * n + col] / a.col_scale[col] / a.row_inv[row];
    const _Float16 h = static_cast<_Float16>(v);
    a.white_hl[blk * 1024 + lane * 8 + i] = h;
    a.white_hl[blk * 1024 + 512 + lane * 8 + i] = static_cast<_Float16>(v - static_cast<float>(h));
  } else if (e0 < nwhite + ntarget) {
    const int e = e0 - nwhite;
    const int nb = rx_blocks(n);
    const int npair = (nb + 1) / 2;
    const int tt = e / (npair * 512);
    const int P = (e / 512) % npair;
    const int lane = (e % 512) / 8;
    const int i = e % 8;
    const int t = 16 * tt + (lane & 15);
    const int band = 32 * P + (i < 4 ? 4 * (lane >> 4) + i : 16 + 4 * (lane >> 4) + (i - 4));
    float v = 0.f;
    if (t < T && band < n)
      v = a.target_white[static_cast<std::size_t>(t) * n + band] * a.row_inv[band] / a.target_inv[t];
    const _Float16 h = static_cast<_Float16>(v);
    const int base = (tt * npair + P) * 1024;
    a.ace_hl[base + lane * 8 + i] = h;
    a.ace_hl[base + 512 + lane * 8 + i] = static_cast<_Float16>(v - static_cast<float>(h));
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
  if (a.dtype == HSI_Q16) return launch_mixed_typed<q16_t>(a, stream);
  return launch_mixed_typed<float>(a, stream);
}

hipError_t launch_rxace_mixed_candidates(const RxAceArgs& a, const hsi_detection* cand, const unsigned* count,
                                         unsigned capacity, int width, const float* spectra,
                                         std::size_t spectra_stride, hipStream_t stream) {
  if (a.dtype != HSI_F32 || !cand || !count || !spectra || width <= 0) return hipErrorInvalidValue;
  if (spectra_stride < static_cast<std::size_t>(a.bands) ||
      static_cast<std::size_t>(capacity) * spectra_stride * sizeof(float) >= (static_cast<std::size_t>(1) << 31))
    return hipErrorInvalidValue;
  MixedGather mg;
  mg.cand = cand;
  mg.count = count;
  mg.cap = capacity;
  mg.width = width;
  mg.spectra = spectra;
  mg.spectra_stride = spectra_stride;
  switch (rx_blocks(a.bands)) {
#define HSI_NB(N) case N: return launch_mixed_gather_nb<N>(a, mg, stream);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8)
    HSI_NB(9) HSI_NB(10) HSI_NB(11) HSI_NB(12) HSI_NB(13) HSI_NB(14)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

hipError_t launch_mixed_layouts(const MixedLayoutArgs& a, hipStream_t stream) {
  if (a.bands < 1 || a.bands > kMaxRxBands) return hipErrorInvalidValue;
  const int nb = rx_blocks(a.bands);
  const int nwhite = block_prefix(nb) * 512;
  const int ntarget = (a.num_targets > 0 && a.ace_hl) ? (a.num_targets > 16 ? 2 : 1) * ((nb + 1) / 2) * 512 : 0;
  hipLaunchKernelGGL(mixed_scales_kernel, dim3(1), dim3(kScaleThreads), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mixed_blocks_kernel, dim3(static_cast<unsigned>(div_up(nwhite + ntarget, 256))), dim3(256), 0,
                     stream, a, nwhite, ntarget);
  return hipGetLastError();
}

}  // namespace hsi_amd
