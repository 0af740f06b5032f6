// detect.hip - threshold, non-maximum suppression and raster-order compaction.
//
// The reference kernel (detect.cu:43-87) reserves output slots with one
// atomic per warp, so its list order changes run to run, and it ignores the
// score polarity. This version follows detect_cpu (sam_cpu.cpp:86-138)
// exactly, including order:
//   1. flag:  every pixel tests the threshold and, if it passes, its
//             (2r+1)^2 neighbourhood; each wave64 ballots its survivors into
//             one 64-bit mask word and each workgroup writes its count.
//   2. scan:  one workgroup turns the per-workgroup counts into exclusive
//             offsets and writes the total found.
//   3. write: each survivor's slot is its workgroup offset plus the
//             popcount of the survivors before it, so detections land in
//             raster order; slots past capacity are counted, not written.
// For radii 1-3 the flag pass takes four blocks per workgroup and counts per
// group of four, and on frames of up to 2M pixels the write pass finds each
// group's offset itself (the sum of the counts before it), so a detection is
// two launches: on this hardware a launch costs ~4 us even when it does
// nothing, which is most of a detection's time.
#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

constexpr int kDetBlock = 256;
constexpr int kScanThreads = 1024;

__device__ __forceinline__ bool better(float a, float b, bool higher) { return higher ? a > b : a < b; }

// Suppression test of detect_cpu's local_best (sam_cpu.cpp:96-112): a
// neighbour strictly better, or equal with a smaller linear index, suppresses.
// R > 0 is a compile-time radius: every neighbour load is issued before any
// comparison, so a candidate pays one memory latency, not (2R+1)^2. R == 0
// handles any radius with a loop.
template <int R>
__device__ __forceinline__ bool local_best(const float* __restrict__ score, std::size_t p, int width,
                                           int height, int radius, float self, bool higher) {
  const int x = static_cast<int>(p % static_cast<std::size_t>(width));
  const int y = static_cast<int>(p / static_cast<std::size_t>(width));
  if constexpr (R > 0) {
    constexpr int D = 2 * R + 1;
    float nb[D][D];
#pragma unroll
    for (int dy = -R; dy <= R; ++dy)
#pragma unroll
      for (int dx = -R; dx <= R; ++dx) {
        const int ny = y + dy, nx = x + dx;
        const bool inside = ny >= 0 && ny < height && nx >= 0 && nx < width;
        nb[dy + R][dx + R] = inside ? score[static_cast<std::size_t>(ny) * width + nx] : self;
      }
    bool ok = true;
#pragma unroll
    for (int dy = -R; dy <= R; ++dy)
#pragma unroll
      for (int dx = -R; dx <= R; ++dx) {
        if (dx == 0 && dy == 0) continue;
        const int ny = y + dy, nx = x + dx;
        const bool inside = ny >= 0 && ny < height && nx >= 0 && nx < width;
        const float other = nb[dy + R][dx + R];
        // in raster order a neighbour precedes p iff dy < 0, or dy == 0 and dx < 0
        const bool before = dy < 0 || (dy == 0 && dx < 0);
        if (inside && (better(other, self, higher) || (other == self && before))) ok = false;
      }
    return ok;
  } else {
    for (int dy = -radius; dy <= radius; ++dy) {
      const int ny = y + dy;
      if (ny < 0 || ny >= height) continue;
      for (int dx = -radius; dx <= radius; ++dx) {
        if (dx == 0 && dy == 0) continue;
        const int nx = x + dx;
        if (nx < 0 || nx >= width) continue;
        const std::size_t q = static_cast<std::size_t>(ny) * width + nx;
        const float other = score[q];
        if (better(other, self, higher)) return false;
        if (other == self && q < p) return false;
      }
    }
    return true;
  }
}

template <int R>
__device__ __forceinline__ bool passes(const float* __restrict__ score, std::size_t p, int width,
                                       int height, float threshold, int radius, bool higher) {
  const float self = score[p];
  if (!(higher ? self >= threshold : self <= threshold)) return false;
  if (radius <= 0) return true;
  return local_best<R>(score, p, width, height, radius, self, higher);
}

// Suppression test against a workgroup's neighbourhood staged in LDS: row
// dy + R of `nb` holds the linear run p0 + dy * width - R .. p0 + dy *
// width + SPAN - 1 + R, so the pixel at span position t finds neighbour
// (dy, dx) at nb[dy + R][t + R + dx]. A run may wrap across image rows; the
// coordinate test rejects every neighbour that is not really inside the
// image. A neighbour suppresses when it is strictly better, or equal and
// earlier in raster order (before: dy < 0, or dy == 0 and dx < 0), so each
// neighbour is one comparison: >= / <= before p, > / < after it (a NaN
// neighbour compares false either way and never suppresses, as in
// detect_cpu).
template <int R, bool HI, int RUN>
__device__ __forceinline__ bool local_best_lds(const float (*nb)[RUN], int t, int x, int y, int width, int height,
                                               float self) {
  bool row_in[2 * R + 1], col_in[2 * R + 1];
#pragma unroll
  for (int d = -R; d <= R; ++d) {
    row_in[d + R] = static_cast<unsigned>(y + d) < static_cast<unsigned>(height);
    col_in[d + R] = static_cast<unsigned>(x + d) < static_cast<unsigned>(width);
  }
  bool suppressed = false;
#pragma unroll
  for (int dy = -R; dy <= R; ++dy)
#pragma unroll
    for (int dx = -R; dx <= R; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const float o = nb[dy + R][t + R + dx];
      const bool before = dy < 0 || (dy == 0 && dx < 0);
      const bool beats = before ? (HI ? o >= self : o <= self) : (HI ? o > self : o < self);
      suppressed = suppressed || (row_in[dy + R] && col_in[dx + R] && beats);
    }
  return !suppressed;
}

// Radii 1-3 (compile-time R) on frames whose linear indices, with a
// neighbourhood's reach, fit in 32 bits. One workgroup takes kFlagGroup
// consecutive detect blocks (1024 pixels), so the grid is one round of
// resident workgroups, and one trip to memory: the (2R + 1) runs of its
// neighbourhood are loaded coalesced with every load in flight at once
// (each pixel's own score is in the middle run), then the lanes over the
// threshold take the suppression test. The masks are written per detect
// block, as one workgroup per block would write them, and the count once per
// workgroup (group_counts[group]), for detect_write_group_kernel.
constexpr int kFlagGroup = 4;
// Frames of up to this many flag groups (2M pixels) rank without a separate
// scan pass: each write workgroup sums the counts of the groups before it.
constexpr int kMaxRankGroups = 2048;

template <int R, bool HI>
__global__ __launch_bounds__(kDetBlock) void detect_flag_fast_kernel(
    const float* __restrict__ score, int width, int height, float threshold, int blocks,
    unsigned long long* __restrict__ masks, unsigned* __restrict__ group_counts) {
  constexpr int SPAN = kDetBlock * kFlagGroup;
  constexpr int RUN = SPAN + 2 * R;
  constexpr int STAGED = (2 * R + 1) * RUN;
  constexpr int PER = (STAGED + kDetBlock - 1) / kDetBlock;
  constexpr int WAVES = kDetBlock / kWave;
  __shared__ float nb[2 * R + 1][RUN];
  __shared__ unsigned wave_counts[kFlagGroup][WAVES];
  const int pixels = width * height;
  const int p0 = static_cast<int>(blockIdx.x) * SPAN;
  const int t = threadIdx.x;
  {
    float v[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = t + k * kDetBlock;
      const int row = e / RUN;
      const int q = p0 + (row - R) * width + (e - row * RUN - R);
      const bool ok = e < STAGED && q >= 0 && q < pixels;
      v[k] = score[ok ? q : 0];
      if (!ok) v[k] = 0.f;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = t + k * kDetBlock;
      if (e < STAGED) nb[e / RUN][e - (e / RUN) * RUN] = v[k];
    }
  }
  __syncthreads();
  const int lane = t & 63;
  const int wave = t >> 6;
#pragma unroll
  for (int s = 0; s < kFlagGroup; ++s) {
    const int tt = s * kDetBlock + t;
    const int p = p0 + tt;
    const float self = nb[R][tt + R];
    bool pass = false;
    if (p < pixels && (HI ? self >= threshold : self <= threshold)) {
      const int y = p / width;
      pass = local_best_lds<R, HI, RUN>(nb, tt, p - y * width, y, width, height, self);
    }
    const unsigned long long ballot = __ballot(pass);
    if (lane == 0) {
      const int blk = static_cast<int>(blockIdx.x) * kFlagGroup + s;
      if (blk < blocks) masks[static_cast<std::size_t>(blk) * WAVES + wave] = ballot;
      wave_counts[s][wave] = static_cast<unsigned>(__popcll(ballot));
    }
  }
  __syncthreads();
  if (t == 0) {
    unsigned total = 0;
#pragma unroll
    for (int s = 0; s < kFlagGroup; ++s)
#pragma unroll
      for (int w = 0; w < WAVES; ++w) total += wave_counts[s][w];
    group_counts[blockIdx.x] = total;
  }
}

// Raster-order write for the flag groups above. Without a scan pass
// (group_offsets null, at most kMaxRankGroups groups) the workgroup of group
// g finds its first slot as the sum of the counts of groups 0 .. g-1, all
// loaded at once, and the last group writes the total found; otherwise the
// scan pass has written each group's first slot and the total. Each
// survivor's slot adds the popcounts of the group's mask words before it. A
// group with no survivor stops after reading its masks.
__global__ __launch_bounds__(kDetBlock) void detect_write_group_kernel(
    const float* __restrict__ score, const int* __restrict__ target, int width, int height, int blocks,
    const unsigned long long* __restrict__ masks, const unsigned* __restrict__ group_counts,
    const unsigned* __restrict__ group_offsets, int capacity, hsi_detection* __restrict__ out,
    unsigned* __restrict__ total_out) {
  constexpr int WAVES = kDetBlock / kWave;
  constexpr int WORDS = kFlagGroup * WAVES;             // mask words of a group
  constexpr int PER = kMaxRankGroups / kDetBlock;       // group counts per thread
  __shared__ unsigned long long words[WORDS];
  __shared__ unsigned wave_sum[WAVES];
  const int groups = static_cast<int>(gridDim.x);
  const int g = static_cast<int>(blockIdx.x);
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  if (t < WORDS) {
    const int blk = g * kFlagGroup + t / WAVES;
    words[t] = blk < blocks ? masks[static_cast<std::size_t>(blk) * WAVES + t % WAVES] : 0ull;
  }
  __syncthreads();
  bool any = false;
#pragma unroll
  for (int w = 0; w < WORDS; ++w) any = any || words[w] != 0ull;
  const bool last = g == groups - 1 && !group_offsets;
  if (!any && !last) return;
  unsigned first = 0;
  if (group_offsets) {
    first = group_offsets[g];
  } else {
    // first slot of the group: sum of the counts before it
    unsigned c[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int i = t + k * kDetBlock;
      const bool ok = i < g;
      c[k] = group_counts[ok ? i : 0];
      if (!ok) c[k] = 0u;
    }
    unsigned local = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) local += c[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) local += __shfl_xor(local, off, kWave);
    if (lane == 0) wave_sum[wave] = local;
    __syncthreads();
#pragma unroll
    for (int w = 0; w < WAVES; ++w) first += wave_sum[w];
    if (last && t == 0) *total_out = first + group_counts[g];
  }
  if (!any) return;
  const std::size_t pixels = static_cast<std::size_t>(width) * height;
  unsigned before = first;  // survivors of the words before word w
#pragma unroll
  for (int w = 0; w < WORDS; ++w) {
    const unsigned long long m = words[w];
    if (w % WAVES == wave && ((m >> lane) & 1ull)) {
      const unsigned rank = before + static_cast<unsigned>(__popcll(m & ((1ull << lane) - 1ull)));
      const std::size_t p = static_cast<std::size_t>(g) * (kDetBlock * kFlagGroup) + static_cast<std::size_t>(w) * kWave + lane;
      if (rank < static_cast<unsigned>(capacity) && p < pixels) {
        hsi_detection d;
        d.x = static_cast<int>(p % static_cast<std::size_t>(width));
        d.y = static_cast<int>(p / static_cast<std::size_t>(width));
        d.target = target ? target[p] : 0;
        d.score = score[p];
        out[rank] = d;
      }
    }
    before += static_cast<unsigned>(__popcll(m));
  }
}

// Any radius, any frame size: 64-bit indices, neighbours read from global
// memory (every load of a compile-time radius issued before any compare).
template <int R>
__global__ __launch_bounds__(kDetBlock) void detect_flag_kernel(
    const float* __restrict__ score, int width, int height, float threshold, int radius,
    int higher, unsigned long long* __restrict__ masks, unsigned* __restrict__ block_counts) {
  __shared__ unsigned wave_counts[kDetBlock / kWave];
  const std::size_t pixels = static_cast<std::size_t>(width) * height;
  const std::size_t p = blockIdx.x * static_cast<std::size_t>(kDetBlock) + threadIdx.x;
  const bool pass = p < pixels && passes<R>(score, p, width, height, threshold, radius, higher != 0);
  const unsigned long long ballot = __ballot(pass);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) {
    masks[blockIdx.x * static_cast<std::size_t>(kDetBlock / kWave) + wave] = ballot;
    wave_counts[wave] = static_cast<unsigned>(__popcll(ballot));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned total = 0;
    for (int w = 0; w < kDetBlock / kWave; ++w) total += wave_counts[w];
    block_counts[blockIdx.x] = total;
  }
}

// Exclusive scan of the per-block counts in one workgroup.
__global__ __launch_bounds__(kScanThreads) void detect_scan_kernel(
    const unsigned* __restrict__ block_counts, unsigned* __restrict__ block_offsets,
    unsigned blocks, unsigned* __restrict__ total_out) {
  __shared__ unsigned partial[kScanThreads];
  const unsigned per = (blocks + kScanThreads - 1) / kScanThreads;
  const unsigned begin = threadIdx.x * per;
  // Up to kScanHeld counts per thread (8192 blocks: frames of up to 2M
  // pixels) stay in registers, all loaded at once; beyond that the thread
  // walks its run twice.
  constexpr unsigned kScanHeld = 8;
  unsigned held[kScanHeld];
  unsigned local = 0;
  if (per <= kScanHeld) {
#pragma unroll
    for (unsigned i = 0; i < kScanHeld; ++i) {
      const unsigned b = begin + i;
      const bool ok = i < per && b < blocks;
      held[i] = block_counts[ok ? b : 0];
      if (!ok) held[i] = 0;
    }
#pragma unroll
    for (unsigned i = 0; i < kScanHeld; ++i) local += held[i];
  } else {
    for (unsigned i = 0; i < per; ++i) {
      const unsigned b = begin + i;
      if (b < blocks) local += block_counts[b];
    }
  }
  // Inclusive scan of the thread sums: within each wave by shuffles, then
  // over the wave totals by the first wave (two barriers in all).
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned v = local;
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const unsigned u = __shfl_up(v, off, kWave);
    if (lane >= off) v += u;
  }
  if (lane == kWave - 1) partial[wave] = v;
  __syncthreads();
  if (wave == 0) {
    unsigned w = lane < kScanThreads / kWave ? partial[lane] : 0u;
#pragma unroll
    for (int off = 1; off < kScanThreads / kWave; off <<= 1) {
      const unsigned u = __shfl_up(w, off, kWave);
      if (lane >= off) w += u;
    }
    if (lane < kScanThreads / kWave) partial[lane] = w;
  }
  __syncthreads();
  const unsigned inclusive = v + (wave > 0 ? partial[wave - 1] : 0u);
  unsigned run = inclusive - local;
  if (per <= kScanHeld) {
#pragma unroll
    for (unsigned i = 0; i < kScanHeld; ++i) {
      const unsigned b = begin + i;
      if (i < per && b < blocks) block_offsets[b] = run;
      run += held[i];
    }
  } else {
    for (unsigned i = 0; i < per; ++i) {
      const unsigned b = begin + i;
      if (b < blocks) {
        block_offsets[b] = run;
        run += block_counts[b];
      }
    }
  }
  if (threadIdx.x == kScanThreads - 1) *total_out = inclusive;
}

__global__ __launch_bounds__(kDetBlock) void detect_write_kernel(
    const float* __restrict__ score, const int* __restrict__ target, int width, int height,
    const unsigned long long* __restrict__ masks, const unsigned* __restrict__ block_offsets,
    int capacity, hsi_detection* __restrict__ out) {
  const std::size_t pixels = static_cast<std::size_t>(width) * height;
  const std::size_t p = blockIdx.x * static_cast<std::size_t>(kDetBlock) + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const unsigned long long* block_masks = masks + blockIdx.x * static_cast<std::size_t>(kDetBlock / kWave);
  const unsigned long long mine = block_masks[wave];
  if (p >= pixels || !((mine >> lane) & 1ull)) return;
  unsigned rank = block_offsets[blockIdx.x];
  for (int w = 0; w < wave; ++w) rank += static_cast<unsigned>(__popcll(block_masks[w]));
  rank += static_cast<unsigned>(__popcll(mine & ((1ull << lane) - 1ull)));
  if (rank < static_cast<unsigned>(capacity)) {
    hsi_detection d;
    d.x = static_cast<int>(p % static_cast<std::size_t>(width));
    d.y = static_cast<int>(p / static_cast<std::size_t>(width));
    d.target = target ? target[p] : 0;
    d.score = score[p];
    out[rank] = d;
  }
}

}  // namespace

// Scratch of hsi_detect_no_alloc: one static array per GPU, sized for frames
// of up to kDetectNoAllocMaxPixels. Host code never names it directly (its
// host-side name is not a device address): detect_no_alloc_scratch looks up
// its address on the current device with hipGetSymbolAddress.
constexpr std::size_t kDetectNoAllocWords =
    (kDetectNoAllocMaxPixels / kDetBlock) * (kDetBlock / kWave) * 2 + 2 * (kDetectNoAllocMaxPixels / kDetBlock);
static_assert(kDetectNoAllocMaxPixels % kDetBlock == 0, "the cap must be whole detect blocks");
static __device__ __align__(256) unsigned scratch[kDetectNoAllocWords];

hipError_t detect_no_alloc_scratch(unsigned** out) {
  void* ptr = nullptr;
  const hipError_t err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  *out = err == hipSuccess ? static_cast<unsigned*>(ptr) : nullptr;
  return err;
}

std::size_t detect_scratch_words(std::size_t pixels) {
  const std::size_t blocks = div_up(pixels, kDetBlock);
  // masks (2 words each), counts, offsets
  return blocks * (kDetBlock / kWave) * 2 + 2 * blocks;
}

hipError_t launch_detect(const float* score, const int* target, int height, int width,
                         hsi_detection_params params, hsi_detection* out, unsigned* count,
                         unsigned* scratch, hipStream_t stream) {
  const std::size_t pixels = static_cast<std::size_t>(width) * height;
  const std::size_t blocks = div_up(pixels, kDetBlock);
  auto* masks = reinterpret_cast<unsigned long long*>(scratch);
  unsigned* counts = scratch + blocks * (kDetBlock / kWave) * 2;
  unsigned* offsets = counts + blocks;
  const int r = params.nms_radius;
  const dim3 grid(static_cast<unsigned>(blocks)), block(kDetBlock);
  // The 32-bit form holds every index it forms (the last workgroup's reach
  // included) below 2^31.
  const std::size_t fast_groups = div_up(blocks, kFlagGroup);
  const bool fast = r >= 1 && r <= 3 &&
                    fast_groups * kFlagGroup * kDetBlock +
                            static_cast<std::size_t>(r + 1) * static_cast<std::size_t>(width) + r <
                        (std::size_t{1} << 31);
  const bool hi = params.higher_is_better != 0;
  if (fast) {
    // flag groups of four blocks, then (past kMaxRankGroups groups) a scan of
    // the group counts, then the write
#define HSI_FLAG_FAST(R, HI)                                                                           \
  hipLaunchKernelGGL((detect_flag_fast_kernel<R, HI>), dim3(static_cast<unsigned>(fast_groups)), block, 0, \
                     stream, score, width, height, params.threshold, static_cast<int>(blocks), masks, counts)
    switch (r * 2 + (hi ? 1 : 0)) {
      case 2: HSI_FLAG_FAST(1, false); break;
      case 3: HSI_FLAG_FAST(1, true); break;
      case 4: HSI_FLAG_FAST(2, false); break;
      case 5: HSI_FLAG_FAST(2, true); break;
      case 6: HSI_FLAG_FAST(3, false); break;
      default: HSI_FLAG_FAST(3, true); break;
    }
#undef HSI_FLAG_FAST
    hipError_t err = hipGetLastError();
    if (err != hipSuccess) return err;
    const bool scan = fast_groups > static_cast<std::size_t>(kMaxRankGroups);
    if (scan) {
      hipLaunchKernelGGL(detect_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, counts, offsets,
                         static_cast<unsigned>(fast_groups), count);
      err = hipGetLastError();
      if (err != hipSuccess) return err;
    }
    hipLaunchKernelGGL(detect_write_group_kernel, dim3(static_cast<unsigned>(fast_groups)), block, 0, stream, score,
                       target, width, height, static_cast<int>(blocks), masks, counts, scan ? offsets : nullptr,
                       params.max_detections, out, count);
    return hipGetLastError();
  }
  switch (r) {
    case 1:
      hipLaunchKernelGGL(detect_flag_kernel<1>, grid, block, 0, stream, score, width, height,
                         params.threshold, r, params.higher_is_better, masks, counts);
      break;
    case 2:
      hipLaunchKernelGGL(detect_flag_kernel<2>, grid, block, 0, stream, score, width, height,
                         params.threshold, r, params.higher_is_better, masks, counts);
      break;
    case 3:
      hipLaunchKernelGGL(detect_flag_kernel<3>, grid, block, 0, stream, score, width, height,
                         params.threshold, r, params.higher_is_better, masks, counts);
      break;
    default:
      hipLaunchKernelGGL(detect_flag_kernel<0>, grid, block, 0, stream, score, width, height,
                         params.threshold, r, params.higher_is_better, masks, counts);
  }
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(detect_scan_kernel, dim3(1), dim3(kScanThreads), 0, stream, counts, offsets,
                     static_cast<unsigned>(blocks), count);
  err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(detect_write_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kDetBlock), 0,
                     stream, score, target, width, height, masks, offsets, params.max_detections,
                     out);
  return hipGetLastError();
}

}  // namespace hsi_amd
