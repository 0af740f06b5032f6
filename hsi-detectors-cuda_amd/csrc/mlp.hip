// mlp.hip - the learned pre-screen: a B -> 64 -> 32 -> 1 ReLU network that
// flags pixels worth scoring exactly (no counterpart in the reference).
//
// Parameters are FP32 masters packed as
//     W1[64][B] | b1[64] | W2[32][64] | b2[32] | w3[32] | b3
// Inference runs on the bf16 matrix cores (v_mfma_f32_16x16x32_bf16, FP32
// accumulate) over a BSQ cube: per 16-pixel tile, layer 1 is 4 row blocks x
// ceil(B/32) K blocks of MFMA, layer 2 takes layer 1's accumulators straight
// back in as its B operand (the accumulator of row block I holds rows
// 16I + 4g + r in lane group g, so the K order of layer 2 is permuted to
// match and W2 is stored in that order), and layer 3 is a dot product
// finished with two lane shuffles.
//
// Training is FP32 on the vector units and deterministic: one kernel runs
// forward and backward per sample and stores activations and deltas, fixed-
// order tiled GEMMs and row sums reduce them to the weight gradients, and an
// Adam kernel updates the masters. The loss is binary cross-entropy on the
// logit with a weight on positives (recall matters more than precision for a
// pre-screen).
#include <climits>
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

constexpr int kH1 = 64;
constexpr int kH2 = 32;
constexpr int kActRows = 2 * kH1 + 2 * kH2 + 2;  // h1, d1, h2, d2, dout, loss

struct Off {
  std::size_t w1, b1, w2, b2, w3, b3, total;
};
__host__ __device__ inline Off offsets(int bands) {
  Off o;
  o.w1 = 0;
  o.b1 = static_cast<std::size_t>(kH1) * bands;
  o.w2 = o.b1 + kH1;
  o.b2 = o.w2 + static_cast<std::size_t>(kH2) * kH1;
  o.w3 = o.b2 + kH2;
  o.b3 = o.w3 + kH2;
  o.total = o.b3 + 1;
  return o;
}

// Layer-2 K order: in K block kb, lane group g element i is hidden unit
// 32 kb + 4g + i (i < 4) or 32 kb + 16 + 4g + (i - 4).
__host__ __device__ inline int l2_k(int kb, int g, int i) {
  return 32 * kb + (i < 4 ? 4 * g + i : 16 + 4 * g + (i - 4));
}

// ---- inference -------------------------------------------------------------

// a1: [J][I][lane] bf8 = W1[16I + lane%16][32J + 8(lane/16) + i]
// a2: [kb][I2][lane] bf8 = W2[16 I2 + lane%16][l2_k(kb, lane/16, i)]
// tail: b1[64] | b2[32] | w3[32] | b3
// a1h, a2h (nullable): the same fragments in FP16, for the screen fused into
// the statistics pass (its inputs are the FP16 staging tile).
__global__ void mlp_layouts_kernel(const float* __restrict__ w, int bands, __bf16* __restrict__ a1,
                                   __bf16* __restrict__ a2, float* __restrict__ tail,
                                   _Float16* __restrict__ a1h, _Float16* __restrict__ a2h) {
  const Off o = offsets(bands);
  const int nkb = div_up(bands, 32);
  const int n1 = nkb * 4 * 64 * 8;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n1; e += gridDim.x * blockDim.x) {
    const int i = e % 8, lane = (e / 8) % 64, I = (e / 512) % 4, J = e / 2048;
    const int row = 16 * I + (lane & 15), col = 32 * J + 8 * (lane >> 4) + i;
    const float v = col < bands ? w[o.w1 + static_cast<std::size_t>(row) * bands + col] : 0.f;
    a1[e] = static_cast<__bf16>(v);
    if (a1h) a1h[e] = static_cast<_Float16>(v);
  }
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < 2 * 2 * 64 * 8; e += gridDim.x * blockDim.x) {
    const int i = e % 8, lane = (e / 8) % 64, I2 = (e / 512) % 2, kb = e / 1024;
    const int row = 16 * I2 + (lane & 15);
    const float v = w[o.w2 + static_cast<std::size_t>(row) * kH1 + l2_k(kb, lane >> 4, i)];
    a2[e] = static_cast<__bf16>(v);
    if (a2h) a2h[e] = static_cast<_Float16>(v);
  }
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < kH1 + 2 * kH2 + 1; e += gridDim.x * blockDim.x) {
    float v;
    if (e < kH1) v = w[o.b1 + e];
    else if (e < kH1 + kH2) v = w[o.b2 + (e - kH1)];
    else if (e < kH1 + 2 * kH2) v = w[o.w3 + (e - kH1 - kH2)];
    else v = w[o.b3];
    tail[e] = v;
  }
}

__device__ __forceinline__ float to_float(float v) { return v; }
__device__ __forceinline__ float to_float(_Float16 v) { return static_cast<float>(v); }

constexpr int kInferWaves = 4;

// Four consecutive pixels of one band through the cube's buffer view: one
// 16-byte (FP32) or 8-byte (FP16) load; a band past the last reads zeros.
__device__ __forceinline__ f4 load4(__amdgpu_buffer_rsrc_t r, std::uint32_t off, float) {
  return __builtin_bit_cast(f4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
}
__device__ __forceinline__ f4 load4(__amdgpu_buffer_rsrc_t r, std::uint32_t off, _Float16) {
  typedef _Float16 h4 __attribute__((ext_vector_type(4)));
  const h4 v = __builtin_bit_cast(h4, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
  return f4{static_cast<float>(v[0]), static_cast<float>(v[1]), static_cast<float>(v[2]), static_cast<float>(v[3])};
}

// A wave scores 64-pixel tiles: lane (g, n) loads pixels 4n..4n+3 of bands
// 32J + 8g + i (i < 8) as one wide load per band, which already is the MFMA
// B operand of four 16-pixel sub-tiles (sub-tile s = pixels 4n + s), so each
// weight fragment read from LDS serves four MFMAs and each load moves 1 KiB
// (FP32) per wave. The next K block's loads are in flight while the matrix
// cores work on this one.
// Screening outputs, per wave: wave w appends its candidates to its own
// region (slots [w * rc, w * rc + rc)) of the record scratch and of the
// caller's spectra rows, and leaves its count in counts[w] - no atomics, and
// the order is fixed (wave, then tile, then pixel). A scan and a compaction
// pass make the record list contiguous afterwards; the spectra stay where
// the pass wrote them, each record naming its row in `target`.
struct ScreenOut {
  float cut;
  hsi_detection* cand;       // region layout
  float* spectra;            // nullable: rows in region layout, pitch `pitch`
  std::size_t pitch;         // floats per spectra row (>= 32 * NKB, multiple of 4)
  unsigned* counts;          // [waves]
  unsigned rc;               // slots per wave
  int width;
};

// Screening of one 64-pixel tile: lane q takes pixel q's logit (held by lane
// q / 4, slot q % 4); flagged pixels get consecutive slots of the wave's
// region, their records written by lane q and their spectra - still in the
// lanes' registers (`keep`: band 32J + 8g + i of pixels 4n .. 4n + 3) - by
// the four lanes of each pixel column.
template <int NKB>
__device__ __forceinline__ void screen_tile(std::size_t pixels, std::size_t tile, const f4& res,
                                            const f4 (&keep)[NKB][8], const ScreenOut& so, std::size_t wave,
                                            unsigned& local) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  const int src = lane >> 2;
  const float r0 = __shfl(res[0], src, 64), r1 = __shfl(res[1], src, 64);
  const float r2 = __shfl(res[2], src, 64), r3 = __shfl(res[3], src, 64);
  const int sub = lane & 3;
  const float lg = sub == 0 ? r0 : (sub == 1 ? r1 : (sub == 2 ? r2 : r3));
  const std::size_t px = tile * 64 + static_cast<std::size_t>(lane);
  const bool flag = px < pixels && lg >= so.cut;
  const unsigned long long m = __ballot(flag);
  if (m == 0ull) return;
  const std::size_t region = wave * so.rc;
  if (flag) {
    const unsigned s = local + static_cast<unsigned>(__popcll(m & ((1ull << lane) - 1ull)));
    if (s < so.rc) {
      const std::size_t w = static_cast<std::size_t>(so.width);
      hsi_detection d;
      d.x = static_cast<int>(px % w);
      d.y = static_cast<int>(px / w);
      d.target = so.spectra ? static_cast<int>(region + s) : 0;   // the spectrum's row
      d.score = lg;
      so.cand[region + s] = d;
    }
  }
  if (so.spectra) {
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int q = 4 * n + u;
    if ((m >> q) & 1ull) {
      const unsigned s = local + static_cast<unsigned>(__popcll(m & ((1ull << q) - 1ull)));
      if (s < so.rc) {
        float* dst = so.spectra + (region + s) * so.pitch + 8 * g;
#pragma unroll
        for (int J = 0; J < NKB; ++J) {
          *reinterpret_cast<f4*>(dst + 32 * J) = f4{keep[J][0][u], keep[J][1][u], keep[J][2][u], keep[J][3][u]};
          *reinterpret_cast<f4*>(dst + 32 * J + 4) = f4{keep[J][4][u], keep[J][5][u], keep[J][6][u], keep[J][7][u]};
        }
      }
    }
  }
  }
  local += static_cast<unsigned>(__popcll(m));
}

template <typename In, int NKB, bool SCREEN>
__global__ __launch_bounds__(kInferWaves * 64) void mlp_infer_kernel(
    const In* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
    const float* __restrict__ center, const bf8* __restrict__ a1, const bf8* __restrict__ a2,
    const float* __restrict__ tail, float* __restrict__ logit, ScreenOut so) {
  __shared__ bf8 a1_lds[NKB * 4 * 64];
  __shared__ bf8 a2_lds[4 * 64];
  __shared__ float c_lds[NKB * 32];
  __shared__ float t_lds[kH1 + 2 * kH2 + 1];
  for (int e = threadIdx.x; e < NKB * 256; e += blockDim.x) a1_lds[e] = a1[e];
  for (int e = threadIdx.x; e < 256; e += blockDim.x) a2_lds[e] = a2[e];
  for (int e = threadIdx.x; e < NKB * 32; e += blockDim.x) c_lds[e] = (center && e < bands) ? center[e] : 0.f;
  for (int e = threadIdx.x; e < kH1 + 2 * kH2 + 1; e += blockDim.x) t_lds[e] = tail[e];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  const std::size_t tiles = div_up(pixels, 64);
  const std::size_t wave_id = blockIdx.x * static_cast<std::size_t>(kInferWaves) + (threadIdx.x >> 6);
  const std::size_t waves = static_cast<std::size_t>(gridDim.x) * kInferWaves;
  const __amdgpu_buffer_rsrc_t rsrc = cube_rsrc(cube, static_cast<std::uint32_t>(bands * stride * sizeof(In)));
  const std::uint32_t lane_off = static_cast<std::uint32_t>((static_cast<std::size_t>(8 * g) * stride + 4 * n) * sizeof(In));

  auto load_kblock = [&](std::size_t t, int J, f4 (&xr)[8]) {
    const std::uint32_t base = lane_off + static_cast<std::uint32_t>((t * 64 + static_cast<std::size_t>(32 * J) * stride) * sizeof(In));
#pragma unroll
    for (int i = 0; i < 8; ++i)
      xr[i] = load4(rsrc, base + static_cast<std::uint32_t>(static_cast<std::size_t>(i) * stride * sizeof(In)), In{});
  };
  // Layer 1 for one 32-band block: z1 += A1[J] x (x centred and rounded to
  // bf16 here; padded bands read 0 and have weight 0).
  auto layer1 = [&](int J, const f4 (&x)[8], f4 (&z1)[4][4]) {
    bf8 xb[4];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const float c = c_lds[32 * J + 8 * g + i];
#pragma unroll
      for (int sub = 0; sub < 4; ++sub) xb[sub][i] = static_cast<__bf16>(x[i][sub] - c);
    }
#pragma unroll
    for (int I = 0; I < 4; ++I) {
      const bf8 a = a1_lds[(J * 4 + I) * 64 + lane];
#pragma unroll
      for (int sub = 0; sub < 4; ++sub)
        z1[sub][I] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, xb[sub], z1[sub][I], 0, 0, 0);
    }
  };
  // Layers 2 and 3: lane (g, n) gets the logits of pixels 4n .. 4n + 3.
  auto layers23 = [&](const f4 (&z1)[4][4]) {
    f4 res = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int sub = 0; sub < 4; ++sub) {
      f4 z2[2] = {f4{0.f, 0.f, 0.f, 0.f}, f4{0.f, 0.f, 0.f, 0.f}};
#pragma unroll
      for (int kb = 0; kb < 2; ++kb) {
        bf8 hb;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int I = 2 * kb + (i < 4 ? 0 : 1);
          const int r = i & 3;
          const float h = z1[sub][I][r] + t_lds[16 * I + 4 * g + r];
          hb[i] = static_cast<__bf16>(h > 0.f ? h : 0.f);
        }
#pragma unroll
        for (int I2 = 0; I2 < 2; ++I2)
          z2[I2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2_lds[(kb * 2 + I2) * 64 + lane], hb, z2[I2], 0, 0, 0);
      }
      float out = 0.f;
#pragma unroll
      for (int I2 = 0; I2 < 2; ++I2)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = 16 * I2 + 4 * g + r;
          const float h = z2[I2][r] + t_lds[kH1 + k];
          out += t_lds[kH1 + kH2 + k] * (h > 0.f ? h : 0.f);
        }
      out += __shfl_xor(out, 16, 64);
      out += __shfl_xor(out, 32, 64);
      res[sub] = out + t_lds[kH1 + 2 * kH2];
    }
    return res;
  };
  auto store_logits = [&](std::size_t tile, const f4& res) {
    const std::size_t p = tile * 64 + 4 * static_cast<std::size_t>(n);
    if (logit && g == 0) {
      if (p + 4 <= pixels) {
        *reinterpret_cast<f4*>(logit + p) = res;
      } else {
#pragma unroll
        for (int sub = 0; sub < 4; ++sub)
          if (p + sub < pixels) logit[p + sub] = res[sub];
      }
    }
  };
  unsigned local = 0;

  if constexpr (SCREEN) {
    // Screening keeps each tile's raw inputs (`keep`, all of its bands) until
    // its logits are known. Loads run two 32-band blocks ahead: blocks 2.. of
    // a tile are issued into `keep` as the tile starts, and blocks 0 and 1 of
    // the next tile into `nx` once block 1 is scored.
    constexpr int AHEAD = NKB < 2 ? NKB : 2;
    f4 keep[NKB][8], nx[AHEAD][8];
    if (wave_id < tiles) {
#pragma unroll
      for (int J = 0; J < AHEAD; ++J) load_kblock(wave_id, J, nx[J]);
    }
    for (std::size_t tile = wave_id; tile < tiles; tile += waves) {
      __asm__ volatile("" ::: "memory");
#pragma unroll
      for (int J = 0; J < AHEAD; ++J)
#pragma unroll
        for (int i = 0; i < 8; ++i) keep[J][i] = nx[J][i];
#pragma unroll
      for (int J = AHEAD; J < NKB; ++J) load_kblock(tile, J, keep[J]);
      f4 z1[4][4];
#pragma unroll
      for (int sub = 0; sub < 4; ++sub)
#pragma unroll
        for (int I = 0; I < 4; ++I) z1[sub][I] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int J = 0; J < NKB; ++J) {
        layer1(J, keep[J], z1);
        if (J == AHEAD - 1 && tile + waves < tiles) {
#pragma unroll
          for (int K = 0; K < AHEAD; ++K) load_kblock(tile + waves, K, nx[K]);
        }
      }
      const f4 res = layers23(z1);
      store_logits(tile, res);
      screen_tile<NKB>(pixels, tile, res, keep, so, wave_id, local);
    }
  } else {
    f4 cur[8], nxt[8];
    if (wave_id < tiles) load_kblock(wave_id, 0, cur);
    for (std::size_t tile = wave_id; tile < tiles; tile += waves) {
      // Re-read the weight fragments from LDS every tile rather than hoisting
      // all of them into registers.
      __asm__ volatile("" ::: "memory");
      f4 z1[4][4];
#pragma unroll
      for (int sub = 0; sub < 4; ++sub)
#pragma unroll
        for (int I = 0; I < 4; ++I) z1[sub][I] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int J = 0; J < NKB; ++J) {
        if (J + 1 < NKB) load_kblock(tile, J + 1, nxt);
        else if (tile + waves < tiles) load_kblock(tile + waves, 0, nxt);
        layer1(J, cur, z1);
#pragma unroll
        for (int i = 0; i < 8; ++i) cur[i] = nxt[i];
      }
      store_logits(tile, layers23(z1));
    }
  }
  if constexpr (SCREEN) {
    if ((threadIdx.x & 63) == 0) so.counts[wave_id] = local;
  }
}

// Exclusive scan of the per-wave counts (one workgroup), the total, and an
// overflow mark: a wave whose candidates did not fit its region lost some, so
// the count is raised past the capacity and the rescoring scores the frame.
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void screen_scan_kernel(const unsigned* __restrict__ counts, int waves,
                                                                   unsigned rc, unsigned cap,
                                                                   unsigned* __restrict__ prefix,
                                                                   unsigned* __restrict__ total_out) {
  __shared__ unsigned part[kScanThreads];
  __shared__ unsigned over;
  const int per = (waves + kScanThreads - 1) / kScanThreads;
  const int b = threadIdx.x * per;
  unsigned s = 0;
  bool lost = false;
  for (int k = 0; k < per; ++k)
    if (b + k < waves) {
      const unsigned c = counts[b + k];
      lost |= c > rc;
      s += c < rc ? c : rc;
    }
  if (threadIdx.x == 0) over = 0u;
  part[threadIdx.x] = s;
  __syncthreads();
  if (lost) atomicOr(&over, 1u);
  for (int off = 1; off < kScanThreads; off <<= 1) {
    const unsigned v = threadIdx.x >= static_cast<unsigned>(off) ? part[threadIdx.x - off] : 0u;
    __syncthreads();
    part[threadIdx.x] += v;
    __syncthreads();
  }
  unsigned run = part[threadIdx.x] - s;   // exclusive
  for (int k = 0; k < per; ++k)
    if (b + k < waves) {
      prefix[b + k] = run;
      const unsigned c = counts[b + k];
      run += c < rc ? c : rc;
    }
  if (threadIdx.x == kScanThreads - 1) {
    const unsigned total = part[kScanThreads - 1];
    *total_out = (over != 0u && total <= cap) ? cap + 1u : total;
  }
}

// Region layout -> contiguous list: one wave per region copies its records
// to their final slots (the spectra rows they name stay in place).
__global__ __launch_bounds__(256) void screen_compact_kernel(const unsigned* __restrict__ counts,
                                                             const unsigned* __restrict__ prefix, int waves,
                                                             unsigned rc, unsigned cap,
                                                             const hsi_detection* __restrict__ rec_in,
                                                             hsi_detection* __restrict__ rec_out) {
  const int w = static_cast<int>(blockIdx.x) * 4 + static_cast<int>(threadIdx.x >> 6);
  if (w >= waves) return;
  const int lane = threadIdx.x & 63;
  const unsigned c = min(counts[w], rc);
  const unsigned base = prefix[w];
  const std::size_t region = static_cast<std::size_t>(w) * rc;
  for (unsigned j = lane; j < c; j += 64)
    if (base + j < cap) rec_out[base + j] = rec_in[region + j];
}

template <typename In>
hipError_t launch_infer_typed(const MlpInferArgs& a, hipStream_t stream) {
  if (!fits_buffer(a.bands, a.stride, sizeof(In))) return hipErrorInvalidValue;
  const std::size_t tiles = div_up(a.pixels, 64);
  const In* cube = static_cast<const In*>(a.cube);
  const bf8* a1 = reinterpret_cast<const bf8*>(a.a1);
  const bf8* a2 = reinterpret_cast<const bf8*>(a.a2);
  const int nkb = div_up(a.bands, 32);
  if (!a.cand) {
    if (!a.logit) return hipErrorInvalidValue;
    std::size_t wgs = div_up(tiles, kInferWaves);
    if (wgs > 2048) wgs = 2048;
    const ScreenOut none{};
    switch (nkb) {
#define HSI_K(N)                                                                                   \
  case N:                                                                                          \
    hipLaunchKernelGGL((mlp_infer_kernel<In, N, false>), dim3(static_cast<unsigned>(wgs)),         \
                       dim3(kInferWaves * 64), 0, stream, cube, a.pixels, a.bands, a.stride,        \
                       a.center, a1, a2, a.tail, a.logit, none);                                    \
    break;
      HSI_K(1) HSI_K(2) HSI_K(3) HSI_K(4) HSI_K(5) HSI_K(6) HSI_K(7) HSI_K(8)
#undef HSI_K
      default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
  }
  // Screening: at least four tiles per wave, at most 1024 workgroups.
  if constexpr (!std::is_same_v<In, float>) {
    return hipErrorInvalidValue;
  } else {
    const int pb = 32 * nkb;
    if (!a.cand_count || a.width <= 0 || !a.scratch_rec || !a.scratch_counts ||
        (a.spectra && (a.spectra_stride < static_cast<std::size_t>(pb) || a.spectra_stride % 4 != 0 ||
                       a.cand_cap > static_cast<unsigned>(INT_MAX))))
      return hipErrorInvalidValue;
    std::size_t wgs = div_up(tiles, static_cast<std::size_t>(4 * kInferWaves));
    if (wgs > kScreenMaxWaves / kInferWaves) wgs = kScreenMaxWaves / kInferWaves;
    if (wgs == 0) wgs = 1;
    const int waves = static_cast<int>(wgs) * kInferWaves;
    const unsigned rc = a.cand_cap / static_cast<unsigned>(waves);
    ScreenOut so{a.cut, a.scratch_rec, a.spectra, a.spectra_stride, a.scratch_counts, rc, a.width};
    switch (nkb) {
#define HSI_K(N)                                                                                   \
  case N:                                                                                          \
    hipLaunchKernelGGL((mlp_infer_kernel<float, N, true>), dim3(static_cast<unsigned>(wgs)),       \
                       dim3(kInferWaves * 64), 0, stream, cube, a.pixels, a.bands, a.stride,        \
                       a.center, a1, a2, a.tail, a.logit, so);                                      \
    break;
      HSI_K(1) HSI_K(2) HSI_K(3) HSI_K(4) HSI_K(5) HSI_K(6) HSI_K(7) HSI_K(8)
#undef HSI_K
      default:
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_screen_compact(a.scratch_counts, waves, rc, a.cand_cap, a.scratch_counts + kScreenMaxWaves,
                                 a.cand_count, a.scratch_rec, a.cand, stream);
  }
}

// ---- training ----------------------------------------------------------------

// Forward and backward for one sample per thread. X is band-major
// (X[b * x_stride + s]), already centred. act rows (each n long): h1[64],
// d1[64], h2[32], d2[32], dout, loss; dout and loss carry the 1/n of the mean.
constexpr int kTrainThreads = 256;

__global__ __launch_bounds__(kTrainThreads) void mlp_fwd_bwd_kernel(
    const float* __restrict__ x, std::size_t x_stride, int n, int bands, const float* __restrict__ labels,
    float pos_weight, const float* __restrict__ w, float* __restrict__ act) {
  __shared__ float wl[kH1 * kMaxBands + kH1 + kH2 * kH1 + 2 * kH2 + 1];  // W1 | b1 | W2 | b2 | w3 | b3
  const Off o = offsets(bands);
  for (std::size_t e = threadIdx.x; e < o.total; e += blockDim.x) wl[e] = w[e];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const std::size_t N = static_cast<std::size_t>(n);
  float h1[kH1];
#pragma unroll
  for (int j = 0; j < kH1; ++j) h1[j] = wl[o.b1 + j];
  for (int b = 0; b < bands; ++b) {
    const float xv = x[static_cast<std::size_t>(b) * x_stride + s];
#pragma unroll
    for (int j = 0; j < kH1; ++j) h1[j] = fmaf(wl[o.w1 + static_cast<std::size_t>(j) * bands + b], xv, h1[j]);
  }
#pragma unroll
  for (int j = 0; j < kH1; ++j) h1[j] = h1[j] > 0.f ? h1[j] : 0.f;
  float h2[kH2];
#pragma unroll
  for (int k = 0; k < kH2; ++k) {
    float acc = wl[o.b2 + k];
#pragma unroll
    for (int j = 0; j < kH1; ++j) acc = fmaf(wl[o.w2 + k * kH1 + j], h1[j], acc);
    h2[k] = acc > 0.f ? acc : 0.f;
  }
  float out = wl[o.b3];
#pragma unroll
  for (int k = 0; k < kH2; ++k) out = fmaf(wl[o.w3 + k], h2[k], out);
  const float y = labels[s];
  // weighted BCE on the logit: pw*y*softplus(-out) + (1-y)*softplus(out)
  const float sp_pos = out > 0.f ? log1pf(expf(-out)) : -out + log1pf(expf(out));  // softplus(-out)
  const float sp_neg = sp_pos + out;                                                 // softplus(out)
  const float prob = 1.f / (1.f + expf(-out));
  const float inv_n = 1.f / static_cast<float>(n);
  const float loss = (pos_weight * y * sp_pos + (1.f - y) * sp_neg) * inv_n;
  const float dout = (pos_weight * y * (prob - 1.f) + (1.f - y) * prob) * inv_n;
  float d2[kH2];
#pragma unroll
  for (int k = 0; k < kH2; ++k) d2[k] = h2[k] > 0.f ? dout * wl[o.w3 + k] : 0.f;
#pragma unroll
  for (int j = 0; j < kH1; ++j) {
    float acc = 0.f;
#pragma unroll
    for (int k = 0; k < kH2; ++k) acc = fmaf(wl[o.w2 + k * kH1 + j], d2[k], acc);
    act[static_cast<std::size_t>(j) * N + s] = h1[j];
    act[static_cast<std::size_t>(kH1 + j) * N + s] = h1[j] > 0.f ? acc : 0.f;
  }
#pragma unroll
  for (int k = 0; k < kH2; ++k) {
    act[static_cast<std::size_t>(2 * kH1 + k) * N + s] = h2[k];
    act[static_cast<std::size_t>(2 * kH1 + kH2 + k) * N + s] = d2[k];
  }
  act[static_cast<std::size_t>(2 * kH1 + 2 * kH2) * N + s] = dout;
  act[static_cast<std::size_t>(2 * kH1 + 2 * kH2 + 1) * N + s] = loss;
}

// C[m * ldc + c] = sum_s A[m * lda + s] * B[c * ldb + s], s in [0, n): one
// 16x16 output tile per workgroup, s in chunks of 64 through LDS, each
// output summed in a fixed order.
__global__ __launch_bounds__(256) void grad_gemm_kernel(const float* __restrict__ A, std::size_t lda,
                                                        const float* __restrict__ B, std::size_t ldb,
                                                        int M, int Ncols, int n, float* __restrict__ C,
                                                        int ldc) {
  __shared__ float as[16][65];
  __shared__ float bs[16][65];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 16, c0 = blockIdx.x * 16;
  float acc = 0.f;
  for (int s0 = 0; s0 < n; s0 += 64) {
    for (int e = threadIdx.x; e < 16 * 64; e += 256) {
      const int r = e / 64, c = e % 64;
      const int s = s0 + c;
      as[r][c] = (m0 + r < M && s < n) ? A[static_cast<std::size_t>(m0 + r) * lda + s] : 0.f;
      bs[r][c] = (c0 + r < Ncols && s < n) ? B[static_cast<std::size_t>(c0 + r) * ldb + s] : 0.f;
    }
    __syncthreads();
#pragma unroll 8
    for (int k = 0; k < 64; ++k) acc = fmaf(as[ty][k], bs[tx][k], acc);
    __syncthreads();
  }
  if (m0 + ty < M && c0 + tx < Ncols) C[(m0 + ty) * ldc + c0 + tx] = acc;
}

// out[m] = sum_s A[m * lda + s]: one wave per row, lane-strided partials and
// a fixed-order tree.
__global__ __launch_bounds__(64) void row_sum_kernel(const float* __restrict__ A, std::size_t lda, int n,
                                                     float* __restrict__ out, std::size_t out_stride) {
  const int m = blockIdx.x;
  float acc = 0.f;
  for (int s = threadIdx.x; s < n; s += 64) acc += A[static_cast<std::size_t>(m) * lda + s];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (threadIdx.x == 0) out[static_cast<std::size_t>(m) * out_stride] = acc;
}

__global__ void adam_kernel(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                            const float* __restrict__ g, std::size_t count, float lr, float beta1,
                            float beta2, float eps, float c1, float c2) {
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// Gather (and centre) pixels into a band-major sample matrix.
// An index outside [0, pixels) yields a zero sample.
__global__ void gather_kernel(const float* __restrict__ cube, std::size_t stride, std::size_t pixels, int bands,
                              const int* __restrict__ index, int n, const float* __restrict__ center,
                              float* __restrict__ out, std::size_t out_stride) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int q = index[s];
  const bool ok = q >= 0 && static_cast<std::size_t>(q) < pixels;
  const std::size_t p = ok ? static_cast<std::size_t>(q) : 0;
  for (int b = 0; b < bands; ++b)
    out[static_cast<std::size_t>(b) * out_stride + s] =
        ok ? cube[static_cast<std::size_t>(b) * stride + p] - (center ? center[b] : 0.f) : 0.f;
}

}  // namespace

std::size_t mlp_param_count(int bands) { return offsets(bands).total; }
std::size_t mlp_a1_halves(int bands) { return static_cast<std::size_t>(div_up(bands, 32)) * 4 * 64 * 8; }
std::size_t mlp_a2_halves() { return 2 * 2 * 64 * 8; }
std::size_t mlp_tail_floats() { return kH1 + 2 * kH2 + 1; }
std::size_t mlp_act_floats(int n) { return static_cast<std::size_t>(kActRows) * n; }

hipError_t launch_screen_compact(const unsigned* counts, int waves, unsigned rc, unsigned cap, unsigned* prefix,
                                 unsigned* total, const hsi_detection* rec_in, hsi_detection* rec_out,
                                 hipStream_t stream) {
  if (waves <= 0 || waves > kScreenMaxWaves) return hipErrorInvalidValue;
  hipLaunchKernelGGL(screen_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, counts, waves, rc, cap, prefix,
                     total);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(screen_compact_kernel, dim3(static_cast<unsigned>(div_up(waves, 4))), dim3(256), 0, stream,
                     counts, prefix, waves, rc, cap, rec_in, rec_out);
  return hipGetLastError();
}

hipError_t launch_mlp_layouts(const float* w, int bands, void* a1, void* a2, float* tail, void* a1h, void* a2h,
                              hipStream_t stream) {
  hipLaunchKernelGGL(mlp_layouts_kernel, dim3(16), dim3(256), 0, stream, w, bands,
                     static_cast<__bf16*>(a1), static_cast<__bf16*>(a2), tail, static_cast<_Float16*>(a1h),
                     static_cast<_Float16*>(a2h));
  return hipGetLastError();
}

hipError_t launch_mlp_infer(const MlpInferArgs& a, hipStream_t stream) {
  if (a.dtype == HSI_F16) return launch_infer_typed<_Float16>(a, stream);
  return launch_infer_typed<float>(a, stream);
}

hipError_t launch_mlp_gradients(const float* x, std::size_t x_stride, int n, int bands, const float* labels,
                                float pos_weight, const float* w, float* act, float* grad, float* loss,
                                hipStream_t stream) {
  const Off o = offsets(bands);
  const std::size_t N = static_cast<std::size_t>(n);
  hipLaunchKernelGGL(mlp_fwd_bwd_kernel, dim3(static_cast<unsigned>(div_up(N, kTrainThreads))),
                     dim3(kTrainThreads), 0, stream, x, x_stride, n, bands, labels, pos_weight, w, act);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  const float* h1 = act;
  const float* d1 = act + kH1 * N;
  const float* h2 = act + 2 * kH1 * N;
  const float* d2 = act + (2 * kH1 + kH2) * N;
  const float* dout = act + (2 * kH1 + 2 * kH2) * N;
  const float* lrow = dout + N;
  auto gemm = [&](const float* A, int M, const float* Bm, std::size_t ldb, int Nc, float* C, int ldc) {
    hipLaunchKernelGGL(grad_gemm_kernel, dim3(static_cast<unsigned>(div_up(Nc, 16)), static_cast<unsigned>(div_up(M, 16))),
                       dim3(256), 0, stream, A, N, Bm, ldb, M, Nc, n, C, ldc);
    return hipGetLastError();
  };
  auto rows = [&](const float* A, int M, float* out) {
    hipLaunchKernelGGL(row_sum_kernel, dim3(static_cast<unsigned>(M)), dim3(64), 0, stream, A, N, n, out,
                       static_cast<std::size_t>(1));
    return hipGetLastError();
  };
  if ((err = gemm(d1, kH1, x, x_stride, bands, grad + o.w1, bands)) != hipSuccess) return err;
  if ((err = rows(d1, kH1, grad + o.b1)) != hipSuccess) return err;
  if ((err = gemm(d2, kH2, h1, N, kH1, grad + o.w2, kH1)) != hipSuccess) return err;
  if ((err = rows(d2, kH2, grad + o.b2)) != hipSuccess) return err;
  if ((err = gemm(dout, 1, h2, N, kH2, grad + o.w3, kH2)) != hipSuccess) return err;
  if ((err = rows(dout, 1, grad + o.b3)) != hipSuccess) return err;
  if (loss && (err = rows(lrow, 1, loss)) != hipSuccess) return err;
  return hipSuccess;
}

hipError_t launch_adam(float* w, float* m, float* v, const float* g, std::size_t count, float lr, float beta1,
                       float beta2, float eps, float c1, float c2, hipStream_t stream) {
  hipLaunchKernelGGL(adam_kernel, dim3(static_cast<unsigned>(div_up(count, 256))), dim3(256), 0, stream, w, m, v,
                     g, count, lr, beta1, beta2, eps, c1, c2);
  return hipGetLastError();
}

hipError_t launch_gather(const float* cube, std::size_t stride, std::size_t pixels, int bands, const int* index,
                         int n, const float* center, float* out, std::size_t out_stride, hipStream_t stream) {
  hipLaunchKernelGGL(gather_kernel, dim3(static_cast<unsigned>(div_up(n, 256))), dim3(256), 0, stream, cube, stride,
                     pixels, bands, index, n, center, out, out_stride);
  return hipGetLastError();
}

}  // namespace hsi_amd
