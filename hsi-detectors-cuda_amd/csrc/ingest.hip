// ingest.hip - layout and precision conversions on the way into the scorers.
#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

// FP32 -> FP16, round to nearest even; eight elements per lane per step.
__global__ void narrow_kernel(const float* __restrict__ src, _Float16* __restrict__ dst,
                              std::size_t count) {
  const std::size_t i = (blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x) * 8;
  if (i + 8 <= count) {
    const float4 a = *reinterpret_cast<const float4*>(src + i);
    const float4 b = *reinterpret_cast<const float4*>(src + i + 4);
    h8 h = {static_cast<_Float16>(a.x), static_cast<_Float16>(a.y), static_cast<_Float16>(a.z),
            static_cast<_Float16>(a.w), static_cast<_Float16>(b.x), static_cast<_Float16>(b.y),
            static_cast<_Float16>(b.z), static_cast<_Float16>(b.w)};
    *reinterpret_cast<h8*>(dst + i) = h;
  } else {
    for (std::size_t k = i; k < count; ++k) dst[k] = static_cast<_Float16>(src[k]);
  }
}

// BIP [pixel][band] -> BSQ [band][stride], through a 64x64 LDS tile so both
// the read (along bands) and the write (along pixels) are contiguous.
constexpr int kTile = 64;
__global__ __launch_bounds__(256) void transpose_kernel(const float* __restrict__ src,
                                                        float* __restrict__ dst,
                                                        std::size_t pixels, int bands,
                                                        std::size_t stride) {
  __shared__ float tile[kTile][kTile + 1];
  const std::size_t p0 = static_cast<std::size_t>(blockIdx.x) * kTile;
  const int b0 = blockIdx.y * kTile;
  const int tx = threadIdx.x & 63;
  const int ty = threadIdx.x >> 6;
  for (int r = ty; r < kTile; r += 4) {
    const std::size_t p = p0 + r;
    const int b = b0 + tx;
    tile[r][tx] = (p < pixels && b < bands) ? src[p * bands + b] : 0.f;
  }
  __syncthreads();
  for (int r = ty; r < kTile; r += 4) {
    const int b = b0 + r;
    const std::size_t p = p0 + tx;
    if (b < bands && p < pixels) dst[static_cast<std::size_t>(b) * stride + p] = tile[tx][r];
  }
}

// Snapshot-mosaic demosaic: band b of output pixel (y, x) is raw sample
// (y * period + pos / period, x * period + pos % period), pos = the band's
// mosaic position; integer samples scaled to [0, 1] in float as the
// reference's PNG loader does (value * (1 / full scale)). One thread per
// output pixel per band; reads of one band plane stride by `period`.
template <typename Raw>
__global__ void demosaic_kernel(const Raw* __restrict__ raw, int raw_width, int out_h, int out_w, int period,
                                int bands, const int* __restrict__ position, float scale, float* __restrict__ cube,
                                std::size_t stride) {
  const std::size_t pixels = static_cast<std::size_t>(out_h) * out_w;
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(blockDim.x) + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= pixels || b >= bands) return;
  const int y = static_cast<int>(i / out_w), x = static_cast<int>(i % out_w);
  const int pos = position[b];
  const std::size_t r = static_cast<std::size_t>(y * period + pos / period) * raw_width +
                        static_cast<std::size_t>(x) * period + pos % period;
  cube[static_cast<std::size_t>(b) * stride + i] = static_cast<float>(raw[r]) * scale;
}

// ---- Q16 encoding -------------------------------------------------------------
//
// Per band: offset = min x, scale = (max x - min x) / 65535, code
// q = rint((x - offset) * (65535 / (max - min))) clamped to [0, 65535]. The
// band extrema are found with atomics on an order-preserving unsigned image
// of the floats, so the result does not depend on the order of the updates.
__device__ __forceinline__ unsigned ordered_bits(float f) {
  const unsigned b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float from_ordered(unsigned u) {
  return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

__global__ void q16_range_init_kernel(unsigned* __restrict__ extrema, int bands) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b < bands) {
    extrema[2 * b] = 0xffffffffu;  // running minimum
    extrema[2 * b + 1] = 0u;       // running maximum
  }
}

constexpr int kRangeThreads = 256;
constexpr std::size_t kRangeChunk = 65536;  // pixels per workgroup: 64 float4 loads per thread

// Block (x, b): pixels [x * kRangeChunk, +kRangeChunk) of band b, reduced in
// the workgroup (shuffles, then LDS) to one pair of atomics, so a band's
// extrema see a few atomics per megapixel rather than one per wave.
__global__ __launch_bounds__(kRangeThreads) void q16_range_kernel(const float* __restrict__ src, std::size_t pixels,
                                                                  std::size_t stride,
                                                                  unsigned* __restrict__ extrema) {
  __shared__ float wlo[kRangeThreads / 64], whi[kRangeThreads / 64];
  const int b = blockIdx.y;
  const float* plane = src + static_cast<std::size_t>(b) * stride;
  const std::size_t p0 = static_cast<std::size_t>(blockIdx.x) * kRangeChunk;
  std::size_t p1 = p0 + kRangeChunk;
  if (p1 > pixels) p1 = pixels;
  float lo = INFINITY, hi = -INFINITY;
  for (std::size_t p = p0 + static_cast<std::size_t>(threadIdx.x) * 4; p < p1; p += kRangeThreads * 4) {
    const float4 v = *reinterpret_cast<const float4*>(plane + p);  // p + 3 < stride: stride % 8 == 0
    const float e[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i)
      if (p + i < p1) {
        lo = fminf(lo, e[i]);
        hi = fmaxf(hi, e[i]);
      }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = fminf(lo, __shfl_xor(lo, off, 64));
    hi = fmaxf(hi, __shfl_xor(hi, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    wlo[threadIdx.x >> 6] = lo;
    whi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kRangeThreads / 64; ++w) {
      lo = fminf(lo, wlo[w]);
      hi = fmaxf(hi, whi[w]);
    }
    if (lo <= hi) {
      atomicMin(&extrema[2 * b], ordered_bits(lo));
      atomicMax(&extrema[2 * b + 1], ordered_bits(hi));
    }
  }
}

// Eight elements per thread of plane b (a multiple of 8 never straddles the
// padded stride); elements past the pixel count are coded 0.
__global__ __launch_bounds__(256) void q16_encode_kernel(const float* __restrict__ src, std::size_t pixels,
                                                         std::size_t stride, const unsigned* __restrict__ extrema,
                                                         std::uint16_t* __restrict__ dst, float* __restrict__ q_offset,
                                                         float* __restrict__ q_scale) {
  const int b = blockIdx.y;
  const float lo = from_ordered(extrema[2 * b]);
  const float hi = from_ordered(extrema[2 * b + 1]);
  const float span = hi - lo;
  const float scale = span > 0.f ? span / 65535.f : 0.f;
  const float inv = span > 0.f ? 65535.f / span : 0.f;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    q_offset[b] = lo;
    q_scale[b] = scale;
  }
  const std::size_t i = (static_cast<std::size_t>(blockIdx.x) * 256 + threadIdx.x) * 8;
  if (i >= stride) return;
  const float* in = src + static_cast<std::size_t>(b) * stride + i;
  const float4 a = *reinterpret_cast<const float4*>(in);
  const float4 c = *reinterpret_cast<const float4*>(in + 4);
  const float v[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
  unsigned w[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    unsigned q2[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int e = 2 * k + h;
      float q = rintf((v[e] - lo) * inv);
      q = fminf(fmaxf(q, 0.f), 65535.f);
      q2[h] = i + e < pixels ? static_cast<unsigned>(q) : 0u;
    }
    w[k] = q2[0] | (q2[1] << 16);
  }
  *reinterpret_cast<uint4*>(dst + static_cast<std::size_t>(b) * stride + i) = make_uint4(w[0], w[1], w[2], w[3]);
}

}  // namespace

std::size_t q16_extrema_words(int bands) { return 2 * static_cast<std::size_t>(bands); }

hipError_t launch_encode_q16(const float* src, std::size_t pixels, int bands, std::size_t stride,
                             unsigned* extrema, std::uint16_t* dst, float* q_offset, float* q_scale,
                             hipStream_t stream) {
  hipLaunchKernelGGL(q16_range_init_kernel, dim3(static_cast<unsigned>(div_up(bands, 256))), dim3(256), 0, stream,
                     extrema, bands);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  hipLaunchKernelGGL(q16_range_kernel, dim3(static_cast<unsigned>(div_up(pixels, kRangeChunk)), bands),
                     dim3(kRangeThreads), 0, stream, src, pixels, stride, extrema);
  if ((err = hipGetLastError()) != hipSuccess) return err;
  hipLaunchKernelGGL(q16_encode_kernel, dim3(static_cast<unsigned>(div_up(stride, 256 * 8)), bands), dim3(256), 0,
                     stream, src, pixels, stride, extrema, dst, q_offset, q_scale);
  return hipGetLastError();
}

hipError_t launch_demosaic(const void* raw, int raw_type, int raw_height, int raw_width, int period, int bands,
                           const int* d_position, float* cube, std::size_t stride, hipStream_t stream) {
  (void)raw_height;
  const int out_h = raw_height / period, out_w = raw_width / period;
  const std::size_t pixels = static_cast<std::size_t>(out_h) * out_w;
  const dim3 grid(static_cast<unsigned>(div_up(pixels, 256)), static_cast<unsigned>(bands));
  switch (raw_type) {
    case 0:
      hipLaunchKernelGGL((demosaic_kernel<unsigned char>), grid, dim3(256), 0, stream,
                         static_cast<const unsigned char*>(raw), raw_width, out_h, out_w, period, bands, d_position,
                         1.0f / 255.0f, cube, stride);
      break;
    case 1:
      hipLaunchKernelGGL((demosaic_kernel<unsigned short>), grid, dim3(256), 0, stream,
                         static_cast<const unsigned short*>(raw), raw_width, out_h, out_w, period, bands, d_position,
                         1.0f / 65535.0f, cube, stride);
      break;
    default:
      hipLaunchKernelGGL((demosaic_kernel<float>), grid, dim3(256), 0, stream, static_cast<const float*>(raw),
                         raw_width, out_h, out_w, period, bands, d_position, 1.0f, cube, stride);
  }
  return hipGetLastError();
}

hipError_t launch_narrow_to_half(const float* src, _Float16* dst, std::size_t count,
                                 hipStream_t stream) {
  const std::size_t threads = div_up(count, 8);
  hipLaunchKernelGGL(narrow_kernel, dim3(static_cast<unsigned>(div_up(threads, 256))), dim3(256), 0,
                     stream, src, dst, count);
  return hipGetLastError();
}

hipError_t launch_transpose_bip_to_bsq(const float* src, float* dst, int height, int width,
                                       int bands, std::size_t stride, hipStream_t stream) {
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  const dim3 grid(static_cast<unsigned>(div_up(pixels, kTile)),
                  static_cast<unsigned>(div_up(static_cast<std::size_t>(bands), kTile)));
  hipLaunchKernelGGL(transpose_kernel, grid, dim3(256), 0, stream, src, dst, pixels, bands, stride);
  return hipGetLastError();
}

}  // namespace hsi_amd
