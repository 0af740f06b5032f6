// stats.hip - scene moments on the device (gfx950).
//
// sum(x) and the lower triangle of sum(x x^T) over every pixel, as
// SpectralMoments::add accumulates them (linalg.cpp:18-29,
// detectors_cpu.cpp:77-93). This is a rank-P symmetric update of a BxB matrix
// - a SYRK with K = pixels - so it runs on the matrix cores over 16x16 band
// blocks (lower blocks only). An extra "ones" band appended after the last
// band makes the same products yield sum(x) (its row) and the pixel count
// (its diagonal). Each workgroup owns a contiguous pixel range, stages
// 32-pixel chunks of the cube in LDS (double-buffered: the next chunk's global
// loads are in flight while the matrix cores work on the current one), and
// its waves accumulate their share of the blocks in registers. Per-workgroup
// partials are reduced in a fixed order by a second kernel, so results are
// bit-reproducible run to run.
//
// Two arithmetic forms:
//  PLAIN  v_mfma_f64_16x16x4 on the widened floats: every product of two
//         floats is exact in double, so the only difference from the CPU
//         oracle is the order of the FP64 sums.
//  MIXED  the shifted moments of d = x - r (r: the caller's per-band
//         reference, e.g. the previous frame's mean, or the mean of a sparse
//         sample of the frame) over an F32, F16 or Q16 cube on v_mfma_f32_16x16x32_f16
//         with d split into FP16 hi + lo parts (hi*hi + hi*lo + lo*hi), FP32
//         accumulation over 128-pixel tiles, and FP64 across tiles - the
//         hierarchy the reference designed for its device accumulator
//         (linalg.hpp:53-90, stats_cuda.hpp:121-135). The raw moments are then
//         rebuilt in FP64 as moments_from_shifted does (linalg.cpp:81-103).
#include <type_traits>

#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

constexpr int kChunk = 64;       // pixel granularity of the workgroup ranges
constexpr int kMaxMomWgs = 512;  // most workgroups of any launch (sizes the scratch)

__host__ __device__ constexpr int pair_row(int q) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= q) ++i;
  return i;
}
__host__ __device__ constexpr int pair_col(int q) { return q - pair_row(q) * (pair_row(q) + 1) / 2; }

// Block pairs are dealt to owners (one wave of one of the workgroups that
// share a pixel range) in contiguous runs of the row-major pair order, so an
// owner's pairs touch few distinct 16-band blocks and it loads only those
// operands. Owner position k holds pairs [first(k), first(k + 1)); the wave at
// position k is chosen so that the two waves sharing a SIMD (w and w + 4) hold
// adjacent runs, which balances the matrix-core work per SIMD.
__host__ __device__ constexpr int run_first(int k, int np, int owners) { return (k * np + owners - 1) / owners; }
__host__ __device__ constexpr int run_len(int np, int owners) {
  int m = 0;
  for (int k = 0; k < owners; ++k) {
    const int c = run_first(k + 1, np, owners) - run_first(k, np, owners);
    if (c > m) m = c;
  }
  return m;
}
__host__ __device__ constexpr int wave_position(int wave) { return 2 * (wave % 4) + wave / 4; }

template <int Q, int NP, int NB, int OWNERS, int POS>
__device__ __forceinline__ void mfma_run_f64(const double (&op)[NB], d4 (&acc)[run_len(NP, OWNERS)]) {
  constexpr int F = run_first(POS, NP, OWNERS);
  if constexpr (Q < run_first(POS + 1, NP, OWNERS)) {
    constexpr int I = pair_row(Q), J = pair_col(Q);
    acc[Q - F] = __builtin_amdgcn_mfma_f64_16x16x4f64(op[I], op[J], acc[Q - F], 0, 0, 0);
    mfma_run_f64<Q + 1, NP, NB, OWNERS, POS>(op, acc);
  }
}

template <int Q, int NP, int NB, int OWNERS, int POS>
__device__ __forceinline__ void mfma_run_f16(const h8 (&hi)[NB], const h8 (&lo)[NB],
                                             f4 (&acc)[run_len(NP, OWNERS)]) {
  constexpr int F = run_first(POS, NP, OWNERS);
  if constexpr (Q < run_first(POS + 1, NP, OWNERS)) {
    constexpr int I = pair_row(Q), J = pair_col(Q), M = Q - F;
    acc[M] = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi[I], hi[J], acc[M], 0, 0, 0);
    acc[M] = __builtin_amdgcn_mfma_f32_16x16x32_f16(hi[I], lo[J], acc[M], 0, 0, 0);
    acc[M] = __builtin_amdgcn_mfma_f32_16x16x32_f16(lo[I], hi[J], acc[M], 0, 0, 0);
    mfma_run_f16<Q + 1, NP, NB, OWNERS, POS>(hi, lo, acc);
  }
}

// Bit b set: block b is an operand of some pair in owner POS's run. Used
// through a constexpr variable so it is folded at compile time.
template <int NP, int OWNERS, int POS>
__host__ __device__ constexpr unsigned long long run_blocks() {
  unsigned long long m = 0;
  for (int q = run_first(POS, NP, OWNERS); q < run_first(POS + 1, NP, OWNERS); ++q)
    m |= (1ull << pair_row(q)) | (1ull << pair_col(q));
  return m;
}

// Staging loads, in two steps. stage_raw issues the float4 load of band b at
// pixel px (a multiple of 4) from an address clamped into the cube - band
// below `bands`, pixel inside [p_begin, p_end) - with no branch around it, so
// every load of a chunk is in flight at once and the wait for them comes
// where the values are first used. (A load inside a bounds branch, with its
// value edited in the same branch, makes the compiler wait for each load
// before issuing the next.) stage_fix then substitutes what lies outside:
// the ones band, or zeros past the band count and the pixel range.
__device__ __forceinline__ float4 stage_raw(const float* __restrict__ cube, std::size_t stride, int bands,
                                            int b, std::size_t px, std::size_t p_begin, std::size_t p_end) {
  const int bc = b < bands ? b : bands - 1;
  const std::size_t pc = px < p_end ? px : p_begin;
  return *reinterpret_cast<const float4*>(cube + static_cast<std::size_t>(bc) * stride + pc);
}

__device__ __forceinline__ float4 stage_fix(float4 v, int bands, int b, std::size_t px, std::size_t p_end,
                                            bool ones_band) {
  if (b >= bands) {
    const float o = ones_band && b == bands ? 1.f : 0.f;
    v = make_float4(o, o, o, o);
  }
  if (px >= p_end) v.x = 0.f;
  if (px + 1 >= p_end) v.y = 0.f;
  if (px + 2 >= p_end) v.z = 0.f;
  if (px + 3 >= p_end) v.w = 0.f;
  return v;
}

// The mixed form also stages 16-bit cubes (F16, Q16): four elements per
// 8-byte load, widened to floats (Q16 decoded with its band's code) before
// the shift.
template <typename In> struct Stage4 { typedef uint2 type; };
template <> struct Stage4<float> { typedef float4 type; };

// Four consecutive elements of the cube at byte offset `off` of its buffer
// view (zeros at or past the view's end).
template <typename In>
__device__ __forceinline__ typename Stage4<In>::type stage_buffer(__amdgpu_buffer_rsrc_t r, std::uint32_t off) {
  if constexpr (sizeof(In) == 4)
    return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0));
  else
    return __builtin_bit_cast(uint2, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// Raw staged elements as floats: F32 as read, F16 widened, Q16 as their
// codes (the caller folds each band's decode into its per-row affine).
template <typename In>
__device__ __forceinline__ float4 stage_widen(typename Stage4<In>::type r) {
  if constexpr (sizeof(In) == 4) {
    return r;
  } else {
    const std::uint16_t e[4] = {static_cast<std::uint16_t>(r.x & 0xffffu), static_cast<std::uint16_t>(r.x >> 16),
                                static_cast<std::uint16_t>(r.y & 0xffffu), static_cast<std::uint16_t>(r.y >> 16)};
    float v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
      v[i] = is_q16<In>::value ? static_cast<float>(e[i]) : static_cast<float>(__builtin_bit_cast(_Float16, e[i]));
    return make_float4(v[0], v[1], v[2], v[3]);
  }
}

// Launch shape of each form, per block count NB. Eight waves per workgroup
// let one wave's LDS traffic overlap another's MFMAs. The plain form stages
// 32-pixel chunks, so up to 9 staged blocks two workgroups share a CU (four
// waves per SIMD; 468 -> 381 us at 113 bands against one workgroup with
// 64-pixel chunks). The mixed form stages 64-pixel chunks; up to 8 blocks
// its registers are held to 128 (amdgpu_waves_per_eu(4)), so two workgroups
// share a CU (1024x1024x113: 143 -> 136 us against one workgroup with
// 128-pixel chunks and 197 VGPRs).
// Beyond 12 blocks the mixed form splits a pixel range's block pairs over two
// workgroups (eight ids apart, so the dispatcher puts both on one XCD and the
// second read of each chunk hits that XCD's L2) to fit its accumulators.
constexpr int kWavesPerWg = 8;
constexpr int kXcds = 8;
template <int NB> constexpr int f64_chunk() { return 32; }
template <int NBS> constexpr int f64_max_wgs() { return NBS <= 9 ? 512 : 256; }
constexpr int kMixedMaxWgs = 512;  // most mixed-form workgroups (sizes the scratch)
template <int NB> constexpr int f16_max_wgs() { return NB <= 8 ? 512 : 256; }
template <int NB> constexpr int f16_chunk() { return 64; }
template <int NB> constexpr int f16_splits() { return NB > 12 ? 2 : 1; }

// Plain-form split of the work. The matrix cores take the block pairs of the
// first NBM 16-band blocks; a last block that would hold only one or two real
// bands (113 = 7 x 16 + 1) is left to the vector ALUs instead of spending a
// whole 16-row block pair per column block on it, and so are the sums: lane j
// (j < bands) adds x_j and, per leftover band r, x_r * x_j over the
// workgroup's pixels, reading the staged tile like the matrix-core operands.
// The pixel count needs no accumulation.
constexpr int kMaxVecRows = 2;

struct F64Split {
  int nbm;   // 16-band blocks on the matrix cores
  int rows;  // leftover bands on the vector ALUs (0..kMaxVecRows)
};
__host__ __device__ inline F64Split f64_split(int bands) {
  const int full = bands / 16, rem = bands % 16;
  if (rem == 0) return {full, 0};
  if (rem <= kMaxVecRows && full > 0) return {full, rem};
  return {full + 1, 0};
}
// Per-workgroup partial: the block pairs, then the vector part ([1 + rows]
// rows of vc = 16 * (NBM + 1) doubles: sums, then one row per leftover band).
__host__ __device__ inline std::size_t f64_partial_doubles(int nbm) {
  return static_cast<std::size_t>(nbm) * (nbm + 1) / 2 * 256 + (1 + kMaxVecRows) * 16 * (nbm + 1);
}

// Column blocks of the vector part, one per (round k, wave w) slot
// k * kWavesPerWg + w (-1: none). A wave preferably takes a block that its
// own matrix-core run already loads, so the value is in a register; the
// blocks left over (always the last, which holds the leftover bands) are
// read from LDS.
constexpr unsigned long long run_blocks_of(int np, int owners, int pos) {
  unsigned long long m = 0;
  for (int q = run_first(pos, np, owners); q < run_first(pos + 1, np, owners); ++q)
    m |= (1ull << pair_row(q)) | (1ull << pair_col(q));
  return m;
}
template <int KC>
struct VecBlocks {
  int blk[kWavesPerWg * KC];
};
template <int NBM, int KC>
__host__ __device__ constexpr VecBlocks<KC> vec_blocks() {
  constexpr int NBS = NBM + 1, NP = NBM * (NBM + 1) / 2;
  VecBlocks<KC> r{};
  bool taken[NBS] = {};
  for (int i = 0; i < kWavesPerWg * KC; ++i) r.blk[i] = -1;
  for (int k = 0; k < KC; ++k)
    for (int w = 0; w < kWavesPerWg; ++w) {
      const unsigned long long used = run_blocks_of(NP, kWavesPerWg, wave_position(w));
      for (int b = 0; b < NBM; ++b)
        if (((used >> b) & 1) && !taken[b]) {
          r.blk[k * kWavesPerWg + w] = b;
          taken[b] = true;
          break;
        }
    }
  for (int b = 0; b < NBS; ++b)
    for (int i = 0; i < kWavesPerWg * KC && !taken[b]; ++i)
      if (r.blk[i] < 0) {
        r.blk[i] = b;
        taken[b] = true;
      }
  return r;
}

// Every staged block 0 .. NBM is the vector part's block of exactly one slot.
template <int NBM, int KC>
__host__ __device__ constexpr bool vec_blocks_complete() {
  constexpr VecBlocks<KC> r = vec_blocks<NBM, KC>();
  for (int b = 0; b <= NBM; ++b) {
    int n = 0;
    for (int i = 0; i < kWavesPerWg * KC; ++i) n += r.blk[i] == b ? 1 : 0;
    if (n != 1) return false;
  }
  for (int i = 0; i < kWavesPerWg * KC; ++i)
    if (r.blk[i] < -1 || r.blk[i] > NBM) return false;
  return true;
}

template <int NBM, int W, int POS, bool BUF>
__device__ __forceinline__ void moments_f64_wave(const float* __restrict__ cube, std::size_t p_begin,
                                                 std::size_t p_end, int bands, std::size_t stride,
                                                 int vec_rows, double* tiles, double* __restrict__ out) {
  constexpr int NBS = NBM + 1;                        // staged blocks (leftover bands in the last)
  constexpr int CH = f64_chunk<NBS>();
  constexpr int THREADS = kWavesPerWg * 64;
  constexpr int NP = NBM * (NBM + 1) / 2;
  constexpr int PW = run_len(NP, kWavesPerWg);
  constexpr int LDB = NBS * 16 + 1;                   // odd pitch spreads LDS banks
  constexpr int LOADS = NBS * 16 * (CH / 4);          // float4 per chunk
  constexpr int PER = (LOADS + THREADS - 1) / THREADS;
  constexpr int VC = 16 * NBS;
  constexpr int TILE = CH * LDB;

  const int lane = threadIdx.x & 63;
  d4 acc[PW];
#pragma unroll
  for (int m = 0; m < PW; ++m) acc[m] = d4{0.0, 0.0, 0.0, 0.0};
  // Vector part, spread over all eight waves so none holds the barrier
  // back: lane (g, c) of wave W takes column c of block vbs.blk[8k + W]
  // (k < KC) and, at each 4-pixel step, the pixel whose matrix-core operands
  // it has just read (4s + g), so the column value is mostly already in a
  // register; the four pixel phases are summed by shuffles at the end.
  constexpr int KC = (16 * NBS + 127) / 128;        // column rounds (bands + ones <= 16 NBS)
  constexpr VecBlocks<KC> vbs = vec_blocks<NBM, KC>();
  static_assert(vec_blocks_complete<NBM, KC>(), "vector-part column blocks must cover every staged block once");
  const int vg = lane >> 4;
  double vsum[KC], vrow[KC][kMaxVecRows];
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    vsum[k] = 0.0;
    vrow[k][0] = vrow[k][1] = 0.0;
  }
  const int chunks = static_cast<int>(div_up(p_end - p_begin, CH));

  // BUF: the reads go through a buffer view of the cube, as in the mixed
  // form (thread t takes band rows t / (CH / 4) + k * RSTEP at one pixel
  // column of every chunk: one register plus a per-row constant, no clamp;
  // rows past the band count read zeros from past the view's end, which the
  // launcher guarantees), and only a range's last chunk masks pixels past
  // p_end. Otherwise every read is clamped into the cube and fixed up.
  constexpr int RSTEP = THREADS / (CH / 4);
  static_assert(THREADS % (CH / 4) == 0, "a thread stages one pixel column of every row it takes");
  const __amdgpu_buffer_rsrc_t view =
      cube_rsrc(cube, BUF ? static_cast<std::uint32_t>(static_cast<std::size_t>(bands) * stride * 4) : 0u);
  const std::uint32_t row0 = static_cast<std::uint32_t>(
      (static_cast<std::size_t>(threadIdx.x / (CH / 4)) * stride + (threadIdx.x % (CH / 4)) * 4) * 4);
  const std::uint32_t kstep = static_cast<std::uint32_t>(static_cast<std::size_t>(RSTEP) * stride * 4);
  float4 pre[PER];
  auto load = [&](std::size_t p) {
    if constexpr (BUF) {
      const std::uint32_t at = row0 + static_cast<std::uint32_t>(p * 4);
#pragma unroll
      for (int k = 0; k < PER; ++k)
        pre[k] = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(
                                                view, at + static_cast<std::uint32_t>(k) * kstep, 0, 0));
    } else {
#pragma unroll
      for (int k = 0; k < PER; ++k) {
        const int e = threadIdx.x + THREADS * k;
        const int b = e / (CH / 4);
        pre[k] = stage_raw(cube, stride, bands, b, p + (e % (CH / 4)) * 4, p_begin, p_end);
      }
    }
  };
  auto store_rows = [&](double* t, std::size_t p, auto partial) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + THREADS * k;
      if (e < LOADS) {
        const int b = e / (CH / 4);
        const int k4 = (e % (CH / 4)) * 4;
        float4 v = pre[k];
        if constexpr (!BUF) {
          v = stage_fix(v, bands, b, p + k4, p_end, false);
        } else if constexpr (decltype(partial)::value) {
          const std::size_t px = p + k4;
          if (px >= p_end) v.x = 0.f;
          if (px + 1 >= p_end) v.y = 0.f;
          if (px + 2 >= p_end) v.z = 0.f;
          if (px + 3 >= p_end) v.w = 0.f;
        }
        t[(k4 + 0) * LDB + b] = v.x;
        t[(k4 + 1) * LDB + b] = v.y;
        t[(k4 + 2) * LDB + b] = v.z;
        t[(k4 + 3) * LDB + b] = v.w;
      }
    }
  };
  auto store = [&](double* t, std::size_t p) {
    if (!BUF || p + CH <= p_end)
      store_rows(t, p, std::false_type{});
    else
      store_rows(t, p, std::true_type{});
  };

  load(p_begin);
  store(tiles, p_begin);
  __syncthreads();
  for (int c = 0; c < chunks; ++c) {
    const bool more = c + 1 < chunks;
    const std::size_t pn = p_begin + static_cast<std::size_t>(c + 1) * CH;
    if (more) load(pn);
    const double* t = tiles + (c & 1) * TILE;
    // Fully unrolled up to 7 blocks; beyond that two steps at a time keep the
    // registers within the occupancy the launch needs.
    constexpr int kStepUnroll = NBM <= 7 ? CH / 4 : 2;
#pragma unroll kStepUnroll
    for (int s = 0; s < CH / 4; ++s) {
      // v_mfma_f64_16x16x4 operands: lane l supplies row (l % 16) of the
      // 16-band block at pixel (l / 16) of this 4-pixel step, for A and B.
      double op[NBM];
      const double* row = t + (4 * s + (lane >> 4)) * LDB + (lane & 15);
      constexpr unsigned long long used = run_blocks<NP, kWavesPerWg, POS>();
#pragma unroll
      for (int i = 0; i < NBM; ++i) op[i] = (used >> i) & 1 ? row[16 * i] : 0.0;
      mfma_run_f64<run_first(POS, NP, kWavesPerWg), NP, NBM, kWavesPerWg, POS>(op, acc);
      // Vector part of this pixel while the matrix cores work.
#pragma unroll
      for (int k = 0; k < KC; ++k) {
        const int vb = vbs.blk[k * kWavesPerWg + W];
        if (vb >= 0) {
          const double xj = (vb < NBM && ((used >> vb) & 1)) ? op[vb < NBM ? vb : 0] : row[16 * vb];
          vsum[k] += xj;
          if (vec_rows > 0) {
            const double* lr = row - (lane & 15) + 16 * NBM;  // leftover bands of this pixel
            vrow[k][0] = __builtin_fma(lr[0], xj, vrow[k][0]);
            if (vec_rows > 1) vrow[k][1] = __builtin_fma(lr[1], xj, vrow[k][1]);
          }
        }
      }
    }
    if (more) store(tiles + ((c + 1) & 1) * TILE, pn);
    __syncthreads();
  }

  // Partial layout: [wg][pair][lane][4], the accumulator registers verbatim,
  // then the vector part.
#pragma unroll
  for (int m = 0; m < PW; ++m) {
    const int q = run_first(POS, NP, kWavesPerWg) + m;
    if (q < run_first(POS + 1, NP, kWavesPerWg)) {
      double* dst = out + (static_cast<std::size_t>(q) * 64 + lane) * 4;
      dst[0] = acc[m][0];
      dst[1] = acc[m][1];
      dst[2] = acc[m][2];
      dst[3] = acc[m][3];
    }
  }
  double* v = out + NP * 256;
#pragma unroll
  for (int k = 0; k < KC; ++k) {
    double a = vsum[k], r0 = vrow[k][0], r1 = vrow[k][1];
    a += __shfl_xor(a, 16, 64);
    a += __shfl_xor(a, 32, 64);
    r0 += __shfl_xor(r0, 16, 64);
    r0 += __shfl_xor(r0, 32, 64);
    r1 += __shfl_xor(r1, 16, 64);
    r1 += __shfl_xor(r1, 32, 64);
    const int vb = vbs.blk[k * kWavesPerWg + W];
    const int col = 16 * vb + (lane & 15);
    if (vb >= 0 && vg == 0 && col < bands) {
      v[col] = a;
      if (vec_rows > 0) v[VC + col] = r0;
      if (vec_rows > 1) v[2 * VC + col] = r1;
    }
  }
}

template <int NBM, bool BUF>
__global__ __launch_bounds__(kWavesPerWg * 64) void moments_f64_kernel(const float* __restrict__ cube,
                                                                      std::size_t pixels, int bands,
                                                                      std::size_t stride,
                                                                      std::size_t px_per_wg, int vec_rows,
                                                                      double* __restrict__ partial) {
  constexpr int CH = f64_chunk<NBM + 1>();
  constexpr int LDB = (NBM + 1) * 16 + 1;
  __shared__ double tiles[2 * CH * LDB];
  const std::size_t p_begin = blockIdx.x * px_per_wg;
  std::size_t p_end = p_begin + px_per_wg;
  if (p_end > pixels) p_end = pixels;
  if (p_begin >= p_end) return;
  double* out = partial + static_cast<std::size_t>(blockIdx.x) * f64_partial_doubles(NBM);
  switch (threadIdx.x >> 6) {
#define HSI_W(W)                                                                                   \
  case W:                                                                                          \
    moments_f64_wave<NBM, W, wave_position(W), BUF>(cube, p_begin, p_end, bands, stride, vec_rows, tiles, \
                                               out);                                             \
    break;
    HSI_W(0) HSI_W(1) HSI_W(2) HSI_W(3) HSI_W(4) HSI_W(5) HSI_W(6) HSI_W(7)
#undef HSI_W
  }
}

// Mixed form: d = x - shift as FP16 hi + lo in LDS, [band][pixel] with a
// pitch of CH + 8 halves (16 lanes reading 16 bytes each at rows m = 0..15
// land on 16 disjoint bank quads). Each owner position is its own code path,
// holding only its run's accumulators and operand blocks.
// The screen fused into the statistics pass: its outputs. Candidates go to
// per-wave regions (no atomics): screening wave v of the grid owns rows
// [v * rc, v * rc + rc) of the record scratch and of the spectra, keeps its
// count in a register and leaves it in counts[v]; a scan and a compaction
// pass make the record list contiguous afterwards.
struct FusedScreenDev {
  float cut;
  int nkb;                  // 32-band K blocks of layer 1: ceil(bands / 32)
  int width;
  unsigned rc;              // rows per wave
  unsigned* counts;         // [waves]
  hsi_detection* rec;       // region layout
  float* spectra;
  std::size_t spectra_stride;
};
constexpr int kFusedChunk = 128;

// The network held in one screening wave's registers: layer 1 and 2 FP16
// fragments in the matrix-core operand layout (lane l = 16g + m), and the
// biases and output weights each lane needs (b1 and b2 rows 4g .. 4g + 3 of
// each 16-row block, w3 likewise).
template <int NKBM>
struct ScreenRegs {
  h8 w1[NKBM * 4];
  h8 w2[4];
  float b1[16];
  float b2[8];
  float w3[8];
  float b3;
};

template <int NKBM>
__device__ __forceinline__ void screen_regs_load(ScreenRegs<NKBM>& r, const h8* __restrict__ w1h,
                                                 const h8* __restrict__ w2h, const float* __restrict__ tail,
                                                 int nkb) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
#pragma unroll
  for (int J = 0; J < NKBM; ++J)
#pragma unroll
    for (int I = 0; I < 4; ++I) r.w1[J * 4 + I] = J < nkb ? w1h[(J * 4 + I) * 64 + lane] : h8{};
#pragma unroll
  for (int e = 0; e < 4; ++e) r.w2[e] = w2h[e * 64 + lane];
#pragma unroll
  for (int I = 0; I < 4; ++I)
#pragma unroll
    for (int q = 0; q < 4; ++q) r.b1[4 * I + q] = tail[16 * I + 4 * g + q];
#pragma unroll
  for (int I2 = 0; I2 < 2; ++I2)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      r.b2[4 * I2 + q] = tail[64 + 16 * I2 + 4 * g + q];
      r.w3[4 * I2 + q] = tail[96 + 16 * I2 + 4 * g + q];
    }
  r.b3 = tail[128];
}

// The logits of the 16 pixels at column col0 of a staged chunk (hi: the
// chunk's FP16 parts of x - shift, [row][pixel] with pitch PITCH), in every
// lane group for pixel col0 + lane % 16. Layer 1 is v_mfma_f32_16x16x32_f16
// with the FP16 weights as A and the hi parts as B (lane (g, n): bands
// 32J + 8g .. +7 of pixel col0 + n); layers 2 and 3 as in the standalone
// inference kernel (layer 1's accumulators feed layer 2 directly, its K order
// permuted to match). K blocks past ceil(bands / 32), and weight rows past
// the band count, are zero.
template <int NB, int PITCH>
__device__ __forceinline__ float screen_logit_regs(const _Float16* tile, const ScreenRegs<(NB + 1) / 2>& r,
                                                   int col0) {
  constexpr int NKBM = (NB + 1) / 2;
  constexpr int ROWS = NB * 16;
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  // Accumulators start at the biases.
  f4 z1[4];
#pragma unroll
  for (int I = 0; I < 4; ++I) z1[I] = f4{r.b1[4 * I], r.b1[4 * I + 1], r.b1[4 * I + 2], r.b1[4 * I + 3]};
  // B operand by transposed LDS reads (ds_read_b64_tr_b16): within each
  // 16-lane group, lane c supplies the address of the four pixels
  // col0 + 4 (c % 4) .. +3 of row k0 + c / 4, and lane i receives rows
  // k0 .. k0 + 3 of pixel col0 + i - the operand layout, two reads per K
  // block. A row past the tile (odd NB) is read as the tile's last row, a
  // padding row of zeros or ones whose weights are zero.
  const int c = lane & 15;
#pragma unroll
  for (int J = 0; J < NKBM; ++J) {
    h8 x;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      int row = 32 * J + 8 * g + 4 * h + (c >> 2);
      if constexpr (32 * NKBM > ROWS) row = row < ROWS ? row : ROWS - 1;
      typedef short v4s __attribute__((ext_vector_type(4)));
      const h4 t = __builtin_bit_cast(h4, __builtin_amdgcn_ds_read_tr16_b64_v4i16(
                                              (__attribute__((address_space(3))) v4s*)(tile + row * PITCH + col0 +
                                                                                       4 * (c & 3))));
#pragma unroll
      for (int e = 0; e < 4; ++e) x[4 * h + e] = t[e];
    }
#pragma unroll
    for (int I = 0; I < 4; ++I) z1[I] = __builtin_amdgcn_mfma_f32_16x16x32_f16(r.w1[J * 4 + I], x, z1[I], 0, 0, 0);
  }
  f4 z2[2] = {f4{r.b2[0], r.b2[1], r.b2[2], r.b2[3]}, f4{r.b2[4], r.b2[5], r.b2[6], r.b2[7]}};
#pragma unroll
  for (int kb = 0; kb < 2; ++kb) {
    h8 hb;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int I = 2 * kb + (i < 4 ? 0 : 1);
      const int q = i & 3;
      const float h = z1[I][q];
      hb[i] = static_cast<_Float16>(h > 0.f ? h : 0.f);
    }
#pragma unroll
    for (int I2 = 0; I2 < 2; ++I2)
      z2[I2] = __builtin_amdgcn_mfma_f32_16x16x32_f16(r.w2[kb * 2 + I2], hb, z2[I2], 0, 0, 0);
  }
  float out = 0.f;
#pragma unroll
  for (int I2 = 0; I2 < 2; ++I2)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const float h = z2[I2][q];
      out += r.w3[4 * I2 + q] * (h > 0.f ? h : 0.f);
    }
  out += __shfl_xor(out, 16, 64);
  out += __shfl_xor(out, 32, 64);
  return out + r.b3;
}

// Candidates among the 16 pixels at column col0 of the chunk starting at
// pixel p0: a record and a spectra row each, in region `region_id`. A
// candidate's spectrum is rebuilt from the staged parts as shift + (hi + lo)
// in FP32 - within 2^-21 |x - shift| + 2^-23 |x| + 2^-24 of the pixel (the
// subtraction, lo's rounding, absolute where lo is subnormal, and the add) -
// so nothing waits on memory inside the chunk loop (the stores are not
// waited for).
template <int PITCH>
__device__ __forceinline__ void screen_emit_tile(float logit, const _Float16* hi, const _Float16* lo,
                                                 const float* sh, const FusedScreenDev& fs, int col0,
                                                 std::size_t region_id, std::size_t p0, std::size_t p_end,
                                                 int bands, unsigned& local) {
  const int lane = threadIdx.x & 63;
  const int g = lane >> 4;
  const int n = lane & 15;
  const std::size_t px = p0 + static_cast<std::size_t>(col0 + n);
  const bool flag = g == 0 && px < p_end && logit >= fs.cut;
  const unsigned long long m = __ballot(flag);
  if (m == 0ull) return;
  const int pad = 32 * fs.nkb;
  const std::size_t region = region_id * fs.rc;
  const std::size_t w = static_cast<std::size_t>(fs.width);
  // Four candidates a round (the mask and the candidates' columns are
  // wave-uniform, so the bit walk is scalar): all of a round's LDS reads
  // are issued before any store.
  constexpr int kRound = 4;
  unsigned slot = local;
  for (unsigned long long mm = m; mm != 0ull;) {
    int js[kRound];
#pragma unroll
    for (int u = 0; u < kRound; ++u) {
      js[u] = mm != 0ull ? __ffsll(static_cast<long long>(mm)) - 1 : -1;
      if (mm != 0ull) mm &= mm - 1ull;
    }
    float v0[kRound], v1[kRound], lg[kRound];
#pragma unroll
    for (int u = 0; u < kRound; ++u) {
      v0[u] = v1[u] = 0.f;
      lg[u] = __shfl(logit, js[u] < 0 ? 0 : js[u], 64);
      if (js[u] >= 0 && fs.spectra) {
        const int cj = col0 + js[u];
        const int b0 = lane, b1 = lane + 64;
        if (b0 < bands)
          v0[u] = sh[b0] + (static_cast<float>(hi[b0 * PITCH + cj]) + static_cast<float>(lo[b0 * PITCH + cj]));
        if (b1 < bands)
          v1[u] = sh[b1] + (static_cast<float>(hi[b1 * PITCH + cj]) + static_cast<float>(lo[b1 * PITCH + cj]));
      }
    }
#pragma unroll
    for (int u = 0; u < kRound; ++u) {
      if (js[u] < 0) continue;
      const unsigned su = slot + static_cast<unsigned>(u);
      if (su >= fs.rc) continue;
      const std::size_t row = region + su;
      if (lane == 0) {
        const std::size_t pj = p0 + static_cast<std::size_t>(col0 + js[u]);
        hsi_detection d;
        d.x = static_cast<int>(pj % w);
        d.y = static_cast<int>(pj / w);
        d.target = static_cast<int>(row);
        d.score = lg[u];
        fs.rec[row] = d;
      }
      if (fs.spectra) {
        float* dst = fs.spectra + row * fs.spectra_stride;
        if (lane < pad) dst[lane] = v0[u];
        if (lane + 64 < pad) dst[lane + 64] = v1[u];
      }
    }
    slot += kRound;
  }
  local += static_cast<unsigned>(__popcll(m));
}

// Whether every staging read offset of the mixed-form owners (ch-pixel
// chunks, nb blocks) stays below 2^32 bytes for planes of `stride` elements
// of `eb` bytes: then reads of rows past the band count land at or past the
// cube's buffer view and return zeros.
inline bool stage_offsets_fit(int nb, int ch, std::size_t stride, std::size_t eb) {
  const std::size_t threads = kWavesPerWg * 64, per_row = static_cast<std::size_t>(ch) / 4;
  const std::size_t per = div_up(static_cast<std::size_t>(nb) * 16 * per_row, threads);
  const std::size_t rows = per * (threads / per_row);  // rows a workgroup forms offsets for
  return rows * stride * eb + 16 < (static_cast<std::size_t>(1) << 32);
}

template <typename In, int NB, int POS, int CH = f16_chunk<NB>()>
__device__ __forceinline__ void moments_mixed_owner(const void* __restrict__ cube, std::size_t p_begin,
                                                    std::size_t p_end, int bands, std::size_t stride,
                                                    const float* sh, const float* qo, const float* qs,
                                                    _Float16* hi_t, _Float16* lo_t, double* __restrict__ out) {
  constexpr int NP = NB * (NB + 1) / 2;
  constexpr int OWNERS = kWavesPerWg * f16_splits<NB>();
  constexpr int THREADS = kWavesPerWg * 64;
  constexpr int PW = run_len(NP, OWNERS);
  constexpr int PITCH = CH + 8;
  constexpr int ROWS = NB * 16;
  constexpr int LOADS = ROWS * (CH / 4);
  constexpr int PER = (LOADS + THREADS - 1) / THREADS;
  constexpr int KSTEPS = CH / 32;
  constexpr int TILE = ROWS * PITCH;

  const int lane = threadIdx.x & 63;
  f4 acc[PW];
  d4 tot[PW];
#pragma unroll
  for (int m = 0; m < PW; ++m) {
    acc[m] = f4{0.f, 0.f, 0.f, 0.f};
    tot[m] = d4{0.0, 0.0, 0.0, 0.0};
  }
  const int chunks = static_cast<int>(div_up(p_end - p_begin, CH));

  // Staging reads go through a buffer view of the whole cube (the launchers
  // check that it fits one). Thread t takes band rows t / (CH / 4) + k * RSTEP
  // at pixel column 4 * (t % (CH / 4)) of every chunk, so a read's byte
  // offset is one register plus a per-k constant, with no clamp: rows past
  // the band count and pixels past the cube's end read as zeros, and neither
  // they nor pixels past p_end reach the tiles (see the store).
  static_assert(THREADS % (CH / 4) == 0, "a thread stages one pixel column of every row it takes");
  constexpr int RSTEP = THREADS / (CH / 4);
  constexpr std::size_t EB = sizeof(In);
  const __amdgpu_buffer_rsrc_t view =
      cube_rsrc(cube, static_cast<std::uint32_t>(static_cast<std::size_t>(bands) * stride * EB));
  const std::uint32_t row0 = static_cast<std::uint32_t>(
      (static_cast<std::size_t>(threadIdx.x / (CH / 4)) * stride + (threadIdx.x % (CH / 4)) * 4) * EB);
  const std::uint32_t kstep = static_cast<std::uint32_t>(static_cast<std::size_t>(RSTEP) * stride * EB);
  typename Stage4<In>::type pre[PER];
  auto load = [&](std::size_t p) {
    const std::uint32_t at = row0 + static_cast<std::uint32_t>(p * EB);
#pragma unroll
    for (int k = 0; k < PER; ++k) pre[k] = stage_buffer<In>(view, at + static_cast<std::uint32_t>(k) * kstep);
  };
  // Each thread stages the same rows of every chunk, so the per-row affine
  // d = raw * rm + ra lives in registers: x - shift on a real band (for F32
  // and F16 an exact product by 1, so d is the rounded difference; for Q16
  // code * scale + (offset - shift), the decode and the shift in one
  // rounding), 1 on the ones band, 0 on the padding rows past it.
  float rm[PER], ra[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int b = (static_cast<int>(threadIdx.x) + THREADS * k) / (CH / 4);
    const int bc = b < ROWS ? b : 0;
    const bool real = b < bands;
    rm[k] = real ? (is_q16<In>::value ? qs[bc] : 1.f) : 0.f;
    ra[k] = real ? (is_q16<In>::value ? qo[bc] - sh[bc] : -sh[bc]) : (b == bands ? 1.f : 0.f);
  }
  // A chunk is partial only at the end of a range; its pixels past p_end are
  // zeroed. The two forms are separate code, so full chunks carry no masks.
  auto store_rows = [&](int buf, std::size_t p, auto partial) {
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + THREADS * k;
      if (e < LOADS) {
        const int b = e / (CH / 4);
        const int k4 = (e % (CH / 4)) * 4;
        const float4 f = stage_widen<In>(pre[k]);
        float d0 = __builtin_fmaf(f.x, rm[k], ra[k]);
        float d1 = __builtin_fmaf(f.y, rm[k], ra[k]);
        float d2 = __builtin_fmaf(f.z, rm[k], ra[k]);
        float d3 = __builtin_fmaf(f.w, rm[k], ra[k]);
        if constexpr (decltype(partial)::value) {
          const std::size_t px = p + k4;
          const std::size_t left = p_end > px ? p_end - px : 0;
          d0 = left > 0 ? d0 : 0.f;
          d1 = left > 1 ? d1 : 0.f;
          d2 = left > 2 ? d2 : 0.f;
          d3 = left > 3 ? d3 : 0.f;
        }
        const uint2 s01 = split_pair(d0, d1), s23 = split_pair(d2, d3);
        *reinterpret_cast<uint2*>(hi_t + buf * TILE + b * PITCH + k4) = make_uint2(s01.x, s23.x);
        *reinterpret_cast<uint2*>(lo_t + buf * TILE + b * PITCH + k4) = make_uint2(s01.y, s23.y);
      }
    }
  };
  auto store = [&](int buf, std::size_t p) {
    if (p + CH <= p_end)
      store_rows(buf, p, std::false_type{});
    else
      store_rows(buf, p, std::true_type{});
  };

  load(p_begin);
  store(0, p_begin);
  __syncthreads();
  for (int c = 0; c < chunks; ++c) {
    const bool more = c + 1 < chunks;
    const std::size_t pn = p_begin + static_cast<std::size_t>(c + 1) * CH;
    if (more) load(pn);
    const int buf = c & 1;
#pragma unroll
    for (int ks = 0; ks < KSTEPS; ++ks) {
      // Operands: lane l = 16g + m takes eight pixels 32ks + 8g .. +7 of band
      // row m of every 16-band block (A and B alike, so the pixel order
      // cancels out).
      h8 hi[NB], lo[NB];
      constexpr unsigned long long used = run_blocks<NP, OWNERS, POS>();
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        const int off = buf * TILE + (16 * i + (lane & 15)) * PITCH + 32 * ks + 8 * (lane >> 4);
        if ((used >> i) & 1) {
          hi[i] = *reinterpret_cast<const h8*>(hi_t + off);
          lo[i] = *reinterpret_cast<const h8*>(lo_t + off);
        } else {
          hi[i] = h8{};
          lo[i] = h8{};
        }
      }
      mfma_run_f16<run_first(POS, NP, OWNERS), NP, NB, OWNERS, POS>(hi, lo, acc);
    }

    // FP32 tiles of 128 pixels, then FP64.
    if ((c + 1) % (128 / CH) == 0 || !more) {
#pragma unroll
      for (int m = 0; m < PW; ++m) {
#pragma unroll
        for (int r = 0; r < 4; ++r) tot[m][r] += static_cast<double>(acc[m][r]);
        acc[m] = f4{0.f, 0.f, 0.f, 0.f};
      }
    }
    if (more) store(buf ^ 1, pn);
    __syncthreads();
  }

#pragma unroll
  for (int m = 0; m < PW; ++m) {
    const int q = run_first(POS, NP, OWNERS) + m;
    if (q < run_first(POS + 1, NP, OWNERS)) {
      double* dst = out + (static_cast<std::size_t>(q) * 64 + lane) * 4;
      dst[0] = tot[m][0];
      dst[1] = tot[m][1];
      dst[2] = tot[m][2];
      dst[3] = tot[m][3];
    }
  }
}

template <typename In, int NB>
__global__ __launch_bounds__(kWavesPerWg * 64) __attribute__((amdgpu_waves_per_eu(4))) void moments_mixed_kernel(const void* __restrict__ cube,
                                                                        std::size_t pixels, int bands,
                                                                        std::size_t stride,
                                                                        std::size_t px_per_wg,
                                                                        const float* __restrict__ shift,
                                                                        const float* __restrict__ q_offset,
                                                                        const float* __restrict__ q_scale,
                                                                        double* __restrict__ partial) {
  constexpr int CH = f16_chunk<NB>();
  constexpr int NP = NB * (NB + 1) / 2;
  constexpr int SPLITS = f16_splits<NB>();
  constexpr int THREADS = kWavesPerWg * 64;
  constexpr int PITCH = CH + 8;
  constexpr int ROWS = NB * 16;
  __shared__ _Float16 hi_t[2 * ROWS * PITCH];
  __shared__ _Float16 lo_t[2 * ROWS * PITCH];
  __shared__ float sh[ROWS];
  __shared__ float qo[is_q16<In>::value ? ROWS : 1];
  __shared__ float qs[is_q16<In>::value ? ROWS : 1];

  for (int b = threadIdx.x; b < ROWS; b += THREADS) {
    sh[b] = b < bands ? shift[b] : 0.f;
    if constexpr (is_q16<In>::value) {
      qo[b] = b < bands ? q_offset[b] : 0.f;
      qs[b] = b < bands ? q_scale[b] : 0.f;
    }
  }
  if constexpr (!is_q16<In>::value)
    if (threadIdx.x == 0) qo[0] = qs[0] = 0.f;
  __syncthreads();

  // Workgroup j: pixel range (j / (8 * SPLITS)) * 8 + j % 8, pair half (j / 8) % SPLITS.
  const unsigned range = (blockIdx.x / (kXcds * SPLITS)) * kXcds + blockIdx.x % kXcds;
  const int split = static_cast<int>((blockIdx.x / kXcds) % SPLITS);
  const std::size_t p_begin = range * px_per_wg;
  std::size_t p_end = p_begin + px_per_wg;
  if (p_end > pixels) p_end = pixels;
  if (p_begin >= p_end) return;
  double* out = partial + static_cast<std::size_t>(range) * NP * 256;
  const int pos = split * kWavesPerWg + wave_position(threadIdx.x >> 6);
  switch (pos) {
#define HSI_POS(K)                                                                                 \
  case K:                                                                                          \
    if constexpr (K < kWavesPerWg * SPLITS)                                                        \
      moments_mixed_owner<In, NB, K>(cube, p_begin, p_end, bands, stride, sh, qo, qs, hi_t, lo_t, out); \
    break;
    HSI_POS(0) HSI_POS(1) HSI_POS(2) HSI_POS(3) HSI_POS(4) HSI_POS(5) HSI_POS(6) HSI_POS(7)
    HSI_POS(8) HSI_POS(9) HSI_POS(10) HSI_POS(11) HSI_POS(12) HSI_POS(13) HSI_POS(14) HSI_POS(15)
#undef HSI_POS
  }
}

// Mixed-form moments with the learned pre-screen on the same staged tiles,
// F32 cubes up to 8 moment blocks, one workgroup per CU. The roles are split
// over the waves: waves 0-7 stage 128-pixel chunks and accumulate the
// moments exactly as the mixed-form kernel does, while kScreenWaves more waves
// (one per SIMD) screen the staged chunk with the network in their
// registers - wave 8 + s takes pixels 32s .. 32s + 31 of each chunk as two
// 16-pixel tiles - and write the candidates to their own regions. Both roles
// meet at the same workgroup barrier once per chunk, so a screening wave's
// matrix-core and gather work issues beside the moment waves' staging.
constexpr int kScreenWaves = 4;
template <int NB>
__global__ __launch_bounds__((kWavesPerWg + kScreenWaves) * 64) __attribute__((amdgpu_waves_per_eu(3))) void
moments_screen_kernel(const float* __restrict__ cube, std::size_t pixels, int bands, std::size_t stride,
                         std::size_t px_per_wg, const float* __restrict__ shift, double* __restrict__ partial,
                         const h8* __restrict__ w1h, const h8* __restrict__ w2h, const float* __restrict__ tail,
                         const FusedScreenDev fs) {
  static_assert(NB <= 8 && f16_splits<NB>() == 1, "the fused screen covers up to 8 moment blocks");
  constexpr int CH = kFusedChunk;
  constexpr int PITCH = CH + 8;
  constexpr int ROWS = NB * 16;
  constexpr int NKBM = (NB + 1) / 2;
  constexpr int TILE = ROWS * PITCH;
  static_assert(kScreenWaves * 32 == CH, "two 16-pixel screening tiles per screening wave");
  __shared__ _Float16 hi_t[2 * TILE];
  __shared__ _Float16 lo_t[2 * TILE];
  __shared__ float sh[ROWS];
  __shared__ float qz[1];
  for (int b = threadIdx.x; b < ROWS; b += (kWavesPerWg + kScreenWaves) * 64) sh[b] = b < bands ? shift[b] : 0.f;
  if (threadIdx.x == 0) qz[0] = 0.f;
  __syncthreads();
  const std::size_t p_begin = blockIdx.x * px_per_wg;
  std::size_t p_end = p_begin + px_per_wg;
  if (p_end > pixels) p_end = pixels;
  if (p_begin >= p_end) return;
  const int wave = static_cast<int>(threadIdx.x >> 6);
  if (wave < kWavesPerWg) {
    constexpr int NP = NB * (NB + 1) / 2;
    double* out = partial + static_cast<std::size_t>(blockIdx.x) * NP * 256;
    switch (wave_position(wave)) {
#define HSI_POS(K)                                                                                        \
  case K:                                                                                                 \
    moments_mixed_owner<float, NB, K, CH>(cube, p_begin, p_end, bands, stride, sh, qz, qz, hi_t, lo_t, out); \
    break;
      HSI_POS(0) HSI_POS(1) HSI_POS(2) HSI_POS(3) HSI_POS(4) HSI_POS(5) HSI_POS(6) HSI_POS(7)
#undef HSI_POS
    }
    return;
  }
  // Screening wave s: the same barrier sequence as the moment waves (one
  // after the first chunk is staged, then one per chunk).
  const int s = wave - kWavesPerWg;
  const std::size_t region_id = static_cast<std::size_t>(blockIdx.x) * kScreenWaves + s;
  ScreenRegs<NKBM> r;
  screen_regs_load<NKBM>(r, w1h, w2h, tail, fs.nkb);
  unsigned local = 0u;
  const int chunks = static_cast<int>(div_up(p_end - p_begin, CH));
  __syncthreads();
  for (int c = 0; c < chunks; ++c) {
    const int buf = c & 1;
    const std::size_t p0 = p_begin + static_cast<std::size_t>(c) * CH;
    const _Float16* hi = hi_t + buf * TILE;
    const _Float16* lo = lo_t + buf * TILE;
    const float l0 = screen_logit_regs<NB, PITCH>(hi, r, 32 * s);
    const float l1 = screen_logit_regs<NB, PITCH>(hi, r, 32 * s + 16);
    screen_emit_tile<PITCH>(l0, hi, lo, sh, fs, 32 * s, region_id, p0, p_end, bands, local);
    screen_emit_tile<PITCH>(l1, hi, lo, sh, fs, 32 * s + 16, region_id, p0, p_end, bands, local);
    __syncthreads();
  }
  if ((threadIdx.x & 63) == 0) fs.counts[region_id] = local;
}

// Fixed-order sum of the per-workgroup partials, then scatter of each
// accumulator element to its matrix entry. A block of 256 threads owns 16
// consecutive elements: thread t sums partials w = t/16, t/16 + 16, ... of
// element t%16 (16 consecutive doubles per group, so loads coalesce), and the
// 16 group sums are added in group order - deterministic for a given
// workgroup count. v_mfma_f64_16x16x4 leaves (row l/16 + 4r, column l%16) of
// a block in lane l, register r (rows interleaved); v_mfma_f32_16x16x32_f16
// leaves row 4*(l/16) + r (blocked).
//
// Mixed form (ones band inside the blocks): row == bands carries the sums and
// the count. Plain form: after the nbm*(nbm+1)/2 block pairs come the vector
// part's rows (sums, then the leftover bands 16*nbm + r), and the count is
// the pixel count.
constexpr int kRedElems = 16;
constexpr int kRedGroups = 16;

struct ReduceArgs {
  const double* partial;
  int wgs;
  std::size_t per_wg;   // doubles per workgroup partial
  int nb;               // blocks on the matrix cores
  int bands;
  int interleaved;      // 1: f64 matrix-core layout, 0: f16 layout
  int ones_band;        // 1: mixed form (ones band in the blocks)
  int vec_rows;         // plain form: leftover bands after the blocks
  int vc;               // plain form: pitch of the vector part
  double pixels;        // plain form: the count
  double* count;
  double* sum;
  double* sum_outer;
};

__global__ __launch_bounds__(kRedElems * kRedGroups) void moments_reduce_kernel(ReduceArgs a) {
  __shared__ double part[kRedGroups][kRedElems];
  const std::size_t np = static_cast<std::size_t>(a.nb) * (a.nb + 1) / 2;
  const int local = threadIdx.x % kRedElems;
  const int group = threadIdx.x / kRedElems;
  const std::size_t e = static_cast<std::size_t>(blockIdx.x) * kRedElems + local;
  double s = 0.0;
  if (e < a.per_wg) {
    // Eight loads in flight per round, summed in the same workgroup order.
    for (int w0 = group; w0 < a.wgs; w0 += 8 * kRedGroups) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int w = w0 + u * kRedGroups;
        v[u] = w < a.wgs ? a.partial[static_cast<std::size_t>(w) * a.per_wg + e] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) s += v[u];
    }
  }
  part[group][local] = s;
  __syncthreads();
  if (group != 0 || e >= a.per_wg) return;
  double total = 0.0;
  for (int g = 0; g < kRedGroups; ++g) total += part[g][local];
  const int bands = a.bands;
  if (!a.ones_band && e == 0) *a.count = a.pixels;
  if (e >= np * 256) {                    // plain form, vector part
    const int idx = static_cast<int>(e - np * 256);
    const int which = idx / a.vc, col = idx % a.vc;
    if (col >= bands || which > a.vec_rows) return;
    if (which == 0) {
      a.sum[col] = total;
    } else {
      const int row = 16 * a.nb + which - 1;
      if (col <= row) a.sum_outer[static_cast<std::size_t>(row) * bands + col] = total;
    }
    return;
  }
  const int q = static_cast<int>(e / 256);
  const int lane = static_cast<int>(e % 256) / 4;
  const int r = static_cast<int>(e % 4);
  const int row = 16 * pair_row(q) + (a.interleaved ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r);
  const int col = 16 * pair_col(q) + (lane & 15);
  if (row < bands) {
    if (col <= row) a.sum_outer[static_cast<std::size_t>(row) * bands + col] = total;
  } else if (a.ones_band && row == bands) {
    if (col < bands) a.sum[col] = total;
    else if (col == bands) *a.count = total;
  }
}

// Reference point for the mixed form, one block per band: the caller's shift
// when `given` is non-null, else the FP64 mean of a sparse sample of the
// frame (at most 64 runs of 64 contiguous pixels). Any point near the mean
// protects the covariance from cancellation; the exact moments are rebuilt
// from it afterwards.
template <typename In>
__global__ __launch_bounds__(256) void shift_kernel(const void* __restrict__ cube, std::size_t pixels,
                                                    std::size_t stride, std::size_t run_stride,
                                                    const float* __restrict__ given,
                                                    const float* __restrict__ q_offset,
                                                    const float* __restrict__ q_scale,
                                                    float* __restrict__ shift,
                                                    double* __restrict__ shift_d) {
  __shared__ double red[256];
  __shared__ double cnt[256];
  const int b = blockIdx.x;
  if (given) {
    if (threadIdx.x == 0) {
      shift[b] = given[b];
      shift_d[b] = static_cast<double>(given[b]);
    }
    return;
  }
  const In* plane = static_cast<const In*>(cube) + static_cast<std::size_t>(b) * stride;
  const std::size_t runs = div_up(pixels, run_stride);
  const std::size_t samples = runs * 64;
  double s = 0.0, c = 0.0;
  // At most 64 runs of 64 samples: 16 per thread, loaded first from a
  // clamped index with no branch around them so that all are in flight at
  // once, then added in order; samples past the frame add 0.
  constexpr int kPer = 64 * 64 / 256;
  float v[kPer];
  bool in[kPer];
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    const std::size_t i = threadIdx.x + 256 * static_cast<std::size_t>(k);
    const std::size_t p = (i / 64) * run_stride + (i % 64);
    in[k] = i < samples && p < pixels;
    const std::size_t pc = in[k] ? p : 0;
    if constexpr (is_q16<In>::value) v[k] = q16_decode(static_cast<float>(plane[pc].code), q_scale[b], q_offset[b]);
    else v[k] = static_cast<float>(plane[pc]);
  }
#pragma unroll
  for (int k = 0; k < kPer; ++k) {
    s += in[k] ? static_cast<double>(v[k]) : 0.0;
    c += in[k] ? 1.0 : 0.0;
  }
  red[threadIdx.x] = s;
  cnt[threadIdx.x] = c;
  __syncthreads();
  for (int span = 128; span > 0; span >>= 1) {
    if (threadIdx.x < span) {
      red[threadIdx.x] += red[threadIdx.x + span];
      cnt[threadIdx.x] += cnt[threadIdx.x + span];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float r = cnt[0] > 0.0 ? static_cast<float>(red[0] / cnt[0]) : 0.f;
    shift[b] = r;
    shift_d[b] = static_cast<double>(r);
  }
}

// Raw moments from shifted ones (moments_from_shifted, linalg.cpp:81-103):
// M1 = S1 + N r, M2 = S2 + r S1^T + S1 r^T + N r r^T, lower triangle, in one
// launch: the shifted sums S1 come in their own array, so M2 reads them
// while M1 is written.
__global__ void unshift_kernel(int bands, const double* __restrict__ r, const double* __restrict__ count,
                               const double* __restrict__ shifted_sum, double* __restrict__ sum,
                               double* __restrict__ sum_outer) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  const int tri = bands * (bands + 1) / 2;
  if (e >= tri) return;
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= e) ++i;
  const int j = e - i * (i + 1) / 2;
  const double n = *count;
  double* m2 = sum_outer + static_cast<std::size_t>(i) * bands + j;
  *m2 = *m2 + r[i] * shifted_sum[j] + shifted_sum[i] * r[j] + n * r[i] * r[j];
  if (e < bands) sum[e] = shifted_sum[e] + n * r[e];
}

std::size_t mom_wgs(std::size_t pixels, int max_wgs = kMixedMaxWgs) {
  const std::size_t chunks = div_up(pixels, kChunk);
  return chunks < static_cast<std::size_t>(max_wgs) ? chunks : static_cast<std::size_t>(max_wgs);
}

// fs non-null: the fused screening form (F32, up to 8 blocks; one workgroup
// per CU).
template <typename In, int NB>
hipError_t launch_mixed_nb(const MomentArgs& m, double* scratch, hipStream_t stream,
                           const FusedScreenArgs* fs = nullptr) {
  constexpr bool kCanScreen = std::is_same_v<In, float> && NB <= 8;
  if (fs && !kCanScreen) return hipErrorInvalidValue;
  if (!stage_offsets_fit(NB, fs ? kFusedChunk : f16_chunk<NB>(), m.stride, sizeof(In))) return hipErrorInvalidValue;
  const std::size_t pixels = m.pixels;
  const int bands = m.bands;
  const std::size_t chunks = div_up(pixels, kChunk);
  const std::size_t wgs = mom_wgs(pixels, fs ? 256 : f16_max_wgs<NB>());
  const std::size_t px_per_wg = div_up(chunks, wgs) * kChunk;
  const std::size_t used = div_up(pixels, px_per_wg);
  const int np = NB * (NB + 1) / 2;
  double* partial = scratch;
  hipError_t err;
  // Shift vectors live after the partials (FP64, then FP32), then the
  // shifted sums.
  double* shift_d = scratch + wgs * np * 256;
  float* shift = reinterpret_cast<float*>(shift_d + kMaxBands);
  const std::size_t run_stride = pixels > 64 * 64 ? div_up(pixels, 64) : 64;
  hipLaunchKernelGGL((shift_kernel<In>), dim3(bands), dim3(256), 0, stream, m.cube, pixels, m.stride, run_stride,
                     m.shift, m.q_offset, m.q_scale, shift, shift_d);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  if constexpr (kCanScreen) {
    if (fs) {
      const int waves = static_cast<int>(used) * kScreenWaves;
      const unsigned rc = fs->cap / static_cast<unsigned>(waves);
      FusedScreenDev d{fs->cut, static_cast<int>(div_up(static_cast<std::size_t>(bands), 32)), fs->width, rc,
                       fs->counts, fs->scratch_rec, fs->spectra, fs->spectra_stride};
      hipLaunchKernelGGL(moments_screen_kernel<NB>, dim3(static_cast<unsigned>(used)),
                         dim3((kWavesPerWg + kScreenWaves) * 64), 0, stream, static_cast<const float*>(m.cube),
                         pixels, bands, m.stride, px_per_wg, shift, partial, static_cast<const h8*>(fs->w1h),
                         static_cast<const h8*>(fs->w2h), fs->tail, d);
      if ((err = hipGetLastError()) != hipSuccess) return err;
      if ((err = launch_screen_compact(fs->counts, waves, rc, fs->cap, fs->counts + kScreenMaxWaves, fs->count,
                                       fs->scratch_rec, fs->cand, stream)) != hipSuccess)
        return err;
    }
  }
  if (!fs) {
    // Ranges rounded up to whole groups of 8 so the range/split decode is
    // exact; ranges past the pixel count return at once and write nothing.
    const std::size_t ranges = f16_splits<NB>() > 1 ? round_up(used, kXcds) : used;
    hipLaunchKernelGGL((moments_mixed_kernel<In, NB>), dim3(static_cast<unsigned>(ranges * f16_splits<NB>())),
                       dim3(kWavesPerWg * 64), 0, stream, m.cube, pixels, bands, m.stride, px_per_wg, shift,
                       m.q_offset, m.q_scale, partial);
  }
  if ((err = hipGetLastError()) != hipSuccess) return err;
  double* shifted_sum = shift_d + 2 * kMaxBands;
  ReduceArgs ra{partial, static_cast<int>(used), static_cast<std::size_t>(np) * 256, NB, bands, 0, 1,
                0, 0, 0.0, m.count, shifted_sum, m.sum_outer};
  hipLaunchKernelGGL(moments_reduce_kernel, dim3(static_cast<unsigned>(np * 256 / kRedElems)),
                     dim3(kRedElems * kRedGroups), 0, stream, ra);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  const int tri = bands * (bands + 1) / 2;
  hipLaunchKernelGGL(unshift_kernel, dim3(static_cast<unsigned>(div_up(tri, 256))), dim3(256), 0,
                     stream, bands, shift_d, m.count, shifted_sum, m.sum, m.sum_outer);
  return hipGetLastError();
}

template <typename In>
hipError_t launch_mixed_typed(const MomentArgs& m, double* scratch, hipStream_t stream) {
  if (!fits_buffer(m.bands, m.stride, sizeof(In))) return hipErrorInvalidValue;
  switch (moment_blocks(m.bands)) {
#define HSI_NB(N) \
  case N:         \
    return launch_mixed_nb<In, N>(m, scratch, stream);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8) HSI_NB(9)
    HSI_NB(10) HSI_NB(11) HSI_NB(12) HSI_NB(13) HSI_NB(14) HSI_NB(15) HSI_NB(16) HSI_NB(17)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

hipError_t launch_fused_screen(const MomentArgs& m, const FusedScreenArgs& fs, double* scratch, hipStream_t stream) {
  switch (moment_blocks(m.bands)) {
#define HSI_NB(N) case N: return launch_mixed_nb<float, N>(m, scratch, stream, &fs);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

template <int NBM>
hipError_t launch_f64_nbm(const float* cube, std::size_t pixels, int bands, std::size_t stride,
                          int vec_rows, double* partial, double* count, double* sum, double* sum_outer,
                          hipStream_t stream) {
  const std::size_t chunks = div_up(pixels, kChunk);
  const std::size_t wgs = mom_wgs(pixels, f64_max_wgs<NBM + 1>());
  const std::size_t px_per_wg = div_up(chunks, wgs) * kChunk;
  const std::size_t used = div_up(pixels, px_per_wg);
  if (fits_buffer(bands, stride, sizeof(float)) && stage_offsets_fit(NBM + 1, f64_chunk<NBM + 1>(), stride, sizeof(float)))
    hipLaunchKernelGGL((moments_f64_kernel<NBM, true>), dim3(static_cast<unsigned>(used)), dim3(kWavesPerWg * 64),
                       0, stream, cube, pixels, bands, stride, px_per_wg, vec_rows, partial);
  else
    hipLaunchKernelGGL((moments_f64_kernel<NBM, false>), dim3(static_cast<unsigned>(used)), dim3(kWavesPerWg * 64),
                       0, stream, cube, pixels, bands, stride, px_per_wg, vec_rows, partial);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  const std::size_t per_wg = f64_partial_doubles(NBM);
  ReduceArgs ra{partial, static_cast<int>(used), per_wg, NBM, bands, 1, 0, vec_rows, 16 * (NBM + 1),
                static_cast<double>(pixels), count, sum, sum_outer};
  hipLaunchKernelGGL(moments_reduce_kernel, dim3(static_cast<unsigned>(div_up(per_wg, kRedElems))),
                     dim3(kRedElems * kRedGroups), 0, stream, ra);
  return hipGetLastError();
}

hipError_t launch_moments_f64(const float* cube, std::size_t pixels, int bands, std::size_t stride,
                              double* scratch, double* count, double* sum, double* sum_outer,
                              hipStream_t stream) {
  const F64Split sp = f64_split(bands);
  switch (sp.nbm) {
#define HSI_NBM(N)                                                                                 \
  case N:                                                                                          \
    return launch_f64_nbm<N>(cube, pixels, bands, stride, sp.rows, scratch, count, sum, sum_outer, \
                             stream);
    HSI_NBM(1) HSI_NBM(2) HSI_NBM(3) HSI_NBM(4) HSI_NBM(5) HSI_NBM(6) HSI_NBM(7) HSI_NBM(8)
    HSI_NBM(9) HSI_NBM(10) HSI_NBM(11) HSI_NBM(12) HSI_NBM(13) HSI_NBM(14) HSI_NBM(15) HSI_NBM(16)
#undef HSI_NBM
  }
  return hipErrorInvalidValue;
}

}  // namespace

std::size_t moments_scratch_doubles(std::size_t pixels, int bands) {
  // mixed form: partials, then a double and a float shift vector and the
  // shifted sums; plain form: partials of block pairs and the vector part
  const std::size_t mixed = mom_wgs(pixels) * static_cast<std::size_t>(moment_pairs(bands)) * 256 + 3 * kMaxBands;
  const std::size_t plain = mom_wgs(pixels, kMaxMomWgs) * f64_partial_doubles(f64_split(bands).nbm);
  return mixed > plain ? mixed : plain;
}

bool mixed_moments_fit(int bands, std::size_t stride, std::size_t elem_bytes, bool screen) {
  return fits_buffer(bands, stride, elem_bytes) &&
         stage_offsets_fit(moment_blocks(bands), screen ? kFusedChunk : f16_chunk<1>(), stride, elem_bytes);
}

hipError_t launch_moments(const MomentArgs& m, double* scratch, hipStream_t stream) {
  if (m.precision == HSI_PLAIN) {
    if (m.dtype != HSI_F32) return hipErrorInvalidValue;
    return launch_moments_f64(static_cast<const float*>(m.cube), m.pixels, m.bands, m.stride, scratch, m.count,
                              m.sum, m.sum_outer, stream);
  }
  if (m.precision != HSI_MIXED) return hipErrorInvalidValue;
  switch (m.dtype) {
    case HSI_F32: return launch_mixed_typed<float>(m, scratch, stream);
    case HSI_F16: return launch_mixed_typed<_Float16>(m, scratch, stream);
    case HSI_Q16: return launch_mixed_typed<q16_t>(m, scratch, stream);
  }
  return hipErrorInvalidValue;
}

hipError_t launch_moments_screen(const MomentArgs& m, const FusedScreenArgs& fs, double* scratch,
                                 hipStream_t stream) {
  if (m.precision != HSI_MIXED || m.dtype != HSI_F32 || m.bands < 1 || m.bands > kFusedScreenMaxBands ||
      !fs.w1h || !fs.w2h || !fs.tail || !fs.count || !fs.cand || !fs.counts || !fs.scratch_rec || fs.width <= 0 ||
      !fits_buffer(m.bands, m.stride, sizeof(float)))
    return hipErrorInvalidValue;
  // one workgroup per CU at most: 256 x kScreenWaves screening regions
  static_assert(256 * kScreenWaves <= kScreenMaxWaves, "screening counters");
  if (fs.spectra && fs.spectra_stride < div_up(static_cast<std::size_t>(m.bands), 32) * 32)
    return hipErrorInvalidValue;
  return launch_fused_screen(m, fs, scratch, stream);
}

}  // namespace hsi_amd
