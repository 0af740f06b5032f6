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

}  // namespace

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
