// stats.hip - scene statistics on the device (gfx950).
//
// Two stages per frame, both FP64 where the reference is FP64:
//
// 1. Moments. sum(x) and the lower triangle of sum(x x^T) over every pixel,
//    as SpectralMoments::add accumulates them (linalg.cpp:18-29,
//    detectors_cpu.cpp:77-93). This is a rank-P symmetric update of a BxB
//    matrix - a SYRK with K = pixels - so it runs on the matrix cores:
//    v_mfma_f64_16x16x4 over 16x16 band blocks (lower blocks only). Each
//    workgroup stages 32-pixel chunks of the cube in LDS (widened to double,
//    one extra "ones" band appended so that the same products give sum(x) and
//    the pixel count), and its waves accumulate their share of the blocks in
//    registers over a contiguous pixel range. Per-workgroup partials are
//    reduced in a fixed order by a second kernel, so the result is
//    bit-reproducible run to run. Every product of two floats is exact in
//    double, so the only difference from the CPU oracle is summation order.
//
// 2. Build. One workgroup turns the moments into DetectorStatistics the way
//    build_statistics does (detector.cpp:47-152): mean; covariance and
//    autocorrelation with diagonal loading (linalg.cpp:45-79); Cholesky
//    (linalg.cpp:105-122) with a device status word for a matrix that is not
//    positive definite; the triangular inverse (linalg.cpp:150-162); the
//    whitened centred targets and their squared norms; and the CEM filters
//    R^-1 d / (d^T R^-1 d), with R^-1 d formed as L_R^-T (L_R^-1 d). Matrices
//    live packed-lower in LDS (up to 128 bands) or in global scratch beyond.
#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

constexpr int kChunk = 32;       // pixels staged per LDS chunk
constexpr int kMomWaves = 4;
constexpr int kMaxMomWgs = 256;

__host__ __device__ constexpr int pair_row(int q) {
  int i = 0;
  while ((i + 1) * (i + 2) / 2 <= q) ++i;
  return i;
}
__host__ __device__ constexpr int pair_col(int q) { return q - pair_row(q) * (pair_row(q) + 1) / 2; }

template <int Q, int NP, int NB>
__device__ __forceinline__ void mfma_pairs(int wave, const double (&op)[NB],
                                           d4 (&acc)[(NP + kMomWaves - 1) / kMomWaves]) {
  if constexpr (Q < NP) {
    if (wave == Q % kMomWaves) {
      acc[Q / kMomWaves] = __builtin_amdgcn_mfma_f64_16x16x4f64(op[pair_row(Q)], op[pair_col(Q)],
                                                               acc[Q / kMomWaves], 0, 0, 0);
    }
    mfma_pairs<Q + 1, NP, NB>(wave, op, acc);
  }
}

template <int NB>
__global__ __launch_bounds__(256) void moments_f64_kernel(const float* __restrict__ cube,
                                                          std::size_t pixels, int bands,
                                                          std::size_t stride,
                                                          std::size_t px_per_wg,
                                                          double* __restrict__ partial) {
  constexpr int NP = NB * (NB + 1) / 2;
  constexpr int PW = (NP + kMomWaves - 1) / kMomWaves;
  constexpr int LDB = NB * 16 + 1;  // odd row pitch in doubles spreads LDS banks
  __shared__ double tile[kChunk * LDB];

  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  d4 acc[PW];
#pragma unroll
  for (int m = 0; m < PW; ++m) acc[m] = d4{0.0, 0.0, 0.0, 0.0};

  const std::size_t p_begin = blockIdx.x * px_per_wg;
  std::size_t p_end = p_begin + px_per_wg;
  if (p_end > pixels) p_end = pixels;

  for (std::size_t p = p_begin; p < p_end; p += kChunk) {
    __syncthreads();
    // Stage NB*16 band rows x 32 pixels: float4 loads along the band plane
    // (8 lanes cover one 128-byte row), widened to double and stored
    // transposed so a 16-band slice of one pixel is contiguous.
    for (int e = threadIdx.x; e < NB * 16 * (kChunk / 4); e += 256) {
      const int b = e / (kChunk / 4);
      const int k4 = (e % (kChunk / 4)) * 4;
      const std::size_t px = p + k4;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (px < p_end) {
        if (b < bands) {
          v = *reinterpret_cast<const float4*>(cube + static_cast<std::size_t>(b) * stride + px);
        } else if (b == bands) {
          v = make_float4(1.f, 1.f, 1.f, 1.f);
        }
        if (px + 1 >= p_end) v.y = 0.f;
        if (px + 2 >= p_end) v.z = 0.f;
        if (px + 3 >= p_end) v.w = 0.f;
      }
      tile[(k4 + 0) * LDB + b] = v.x;
      tile[(k4 + 1) * LDB + b] = v.y;
      tile[(k4 + 2) * LDB + b] = v.z;
      tile[(k4 + 3) * LDB + b] = v.w;
    }
    __syncthreads();
#pragma unroll 2
    for (int s = 0; s < kChunk / 4; ++s) {
      // v_mfma_f64_16x16x4 operands: lane l supplies row (l % 16) of the
      // 16-band block at pixel (l / 16) of this 4-pixel step, for A and B.
      double op[NB];
      const double* row = tile + (4 * s + (lane >> 4)) * LDB + (lane & 15);
#pragma unroll
      for (int i = 0; i < NB; ++i) op[i] = row[16 * i];
      mfma_pairs<0, NP, NB>(wave, op, acc);
    }
  }

  // Partial layout: [wg][pair][lane][4], the accumulator registers verbatim.
  double* out = partial + static_cast<std::size_t>(blockIdx.x) * NP * 256;
#pragma unroll
  for (int m = 0; m < PW; ++m) {
    const int q = wave + kMomWaves * m;
    if (q < NP) {
      double* dst = out + (static_cast<std::size_t>(q) * 64 + lane) * 4;
      dst[0] = acc[m][0];
      dst[1] = acc[m][1];
      dst[2] = acc[m][2];
      dst[3] = acc[m][3];
    }
  }
}

// Fixed-order sum of the per-workgroup partials, then scatter of each
// accumulator element to its matrix entry. v_mfma_f64_16x16x4 leaves element
// (row l/16 + 4r, column l%16) of the 16x16 block in lane l, register r
// (rows interleaved across the four lane groups).
__global__ void moments_reduce_kernel(const double* __restrict__ partial, int wgs, int nb,
                                      int bands, double* __restrict__ count,
                                      double* __restrict__ sum, double* __restrict__ sum_outer) {
  const int np = nb * (nb + 1) / 2;
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= np * 256) return;
  double s = 0.0;
  for (int w = 0; w < wgs; ++w) s += partial[static_cast<std::size_t>(w) * np * 256 + e];
  const int q = e / 256;
  const int lane = (e % 256) / 4;
  const int r = e % 4;
  const int bi = pair_row(q);
  const int bj = pair_col(q);
  const int row = 16 * bi + (lane >> 4) + 4 * r;
  const int col = 16 * bj + (lane & 15);
  if (row < bands) {
    if (col <= row) sum_outer[static_cast<std::size_t>(row) * bands + col] = s;
  } else if (row == bands) {
    if (col < bands) sum[col] = s;
    else if (col == bands) *count = s;
  }
}

template <int NB>
hipError_t launch_moments_nb(const float* cube, std::size_t pixels, int bands, std::size_t stride,
                             double* scratch, double* count, double* sum, double* sum_outer,
                             hipStream_t stream) {
  const std::size_t chunks = div_up(pixels, kChunk);
  const std::size_t wgs = chunks < kMaxMomWgs ? chunks : kMaxMomWgs;
  const std::size_t px_per_wg = div_up(chunks, wgs) * kChunk;
  const std::size_t used = div_up(pixels, px_per_wg);
  hipLaunchKernelGGL(moments_f64_kernel<NB>, dim3(static_cast<unsigned>(used)), dim3(256), 0,
                     stream, cube, pixels, bands, stride, px_per_wg, scratch);
  hipError_t err = hipGetLastError();
  if (err != hipSuccess) return err;
  const int np = NB * (NB + 1) / 2;
  const unsigned rblocks = static_cast<unsigned>(div_up(static_cast<std::size_t>(np) * 256, 256));
  hipLaunchKernelGGL(moments_reduce_kernel, dim3(rblocks), dim3(256), 0, stream, scratch,
                     static_cast<int>(used), NB, bands, count, sum, sum_outer);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Build.

constexpr int kBuildThreads = 1024;
constexpr int kLdsBands = 128;
constexpr int kLdsPacked = kLdsBands * (kLdsBands + 1) / 2;

__device__ __forceinline__ int packed(int i, int k) { return i * (i + 1) / 2 + k; }

// Row of the packed index e within a lower triangle (largest r with
// r(r+1)/2 <= e), exact after the float estimate's correction.
__device__ __forceinline__ int tri_row(int e) {
  int r = static_cast<int>((sqrtf(8.0f * static_cast<float>(e) + 1.0f) - 1.0f) * 0.5f);
  while ((r + 1) * (r + 2) / 2 <= e) ++r;
  while (r > 0 && r * (r + 1) / 2 > e) --r;
  return r;
}

// Right-looking Cholesky of a packed lower matrix, in place. Column j is
// finished from the diagonal, then the trailing triangle takes its rank-1
// update. Returns false (uniformly across the block) if a pivot is not
// positive - the test cholesky() makes (linalg.cpp:113), which also catches
// NaN - or is below 4*n*eps of that diagonal's starting value: an exactly
// singular matrix cancels to a pivot of +-1 ulp, whose sign depends on the
// summation order of the moments, and a whitener built on such a pivot is
// meaningless once narrowed to float.
__device__ bool block_cholesky(double* a, int n, double* col, double* diag0) {
  const int tid = threadIdx.x;
  const double tol = 4.0 * n * 2.220446049250313e-16;
  for (int i = tid; i < n; i += kBuildThreads) diag0[i] = a[packed(i, i)];
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double pivot = a[packed(j, j)];
    if (!(pivot > 0.0) || !(pivot > tol * diag0[j])) return false;
    const double d = sqrt(pivot);
    for (int i = j + 1 + tid; i < n; i += kBuildThreads) {
      const double v = a[packed(i, j)] / d;
      a[packed(i, j)] = v;
      col[i] = v;
    }
    __syncthreads();
    if (tid == 0) a[packed(j, j)] = d;
    const int s = n - j - 1;
    const int work = s * (s + 1) / 2;
    for (int e = tid; e < work; e += kBuildThreads) {
      const int r = tri_row(e);
      const int c = e - r * (r + 1) / 2;
      const int i = j + 1 + r;
      const int k = j + 1 + c;
      a[packed(i, k)] -= col[i] * col[k];
    }
    __syncthreads();
  }
  return true;
}

// X = L^-1 for packed lower L, right-looking over rows: once row j of X is
// final, every later row subtracts L[i][j] * X[j][k].
__device__ void block_tri_inverse(const double* l, double* x, int n) {
  const int tid = threadIdx.x;
  const int total = n * (n + 1) / 2;
  for (int e = tid; e < total; e += kBuildThreads) x[e] = 0.0;
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    const double d = l[packed(j, j)];
    for (int k = tid; k <= j; k += kBuildThreads) {
      x[packed(j, k)] = (k == j) ? 1.0 / d : x[packed(j, k)] / d;
    }
    __syncthreads();
    const int rows = n - j - 1;
    const int cols = j + 1;
    const int work = rows * cols;
    for (int e = tid; e < work; e += kBuildThreads) {
      const int i = j + 1 + e / cols;
      const int k = e % cols;
      x[packed(i, k)] -= l[packed(i, j)] * x[packed(j, k)];
    }
    __syncthreads();
  }
}

template <bool IN_LDS>
__global__ __launch_bounds__(kBuildThreads) void build_kernel(BuildArgs a) {
  __shared__ double lds_mat[IN_LDS ? 2 * kLdsPacked : 1];
  __shared__ double mu[kMaxBands];
  __shared__ double col[kMaxBands];
  __shared__ double diag0[kMaxBands];
  __shared__ double red[kBuildThreads / kWave];
  const int n = a.bands;
  const int tid = threadIdx.x;
  const int tri = n * (n + 1) / 2;
  double* m1 = IN_LDS ? lds_mat : a.work;
  double* m2 = IN_LDS ? lds_mat + kLdsPacked : a.work + tri;
  double* vec = a.work + 2 * tri;  // [T][n] scratch for the target passes

  const double count = *a.count;
  if (!(count > 0.0)) {
    if (tid == 0) *a.status = HSI_STATUS_COV_NOT_PD;
    return;
  }
  for (int i = tid; i < n; i += kBuildThreads) mu[i] = a.sum[i] / count;
  __syncthreads();

  // covariance(loading): correlation(0) - mu mu^T, loading on the diagonal.
  for (int e = tid; e < tri; e += kBuildThreads) {
    const int i = tri_row(e);
    const int j = e - i * (i + 1) / 2;
    double v = a.sum_outer[static_cast<std::size_t>(i) * n + j] / count;
    v = v - mu[i] * mu[j];
    if (i == j) v += a.loading;
    m1[e] = v;
  }
  __syncthreads();
  if (!block_cholesky(m1, n, col, diag0)) {
    if (tid == 0) *a.status = HSI_STATUS_COV_NOT_PD;
    return;
  }
  block_tri_inverse(m1, m2, n);

  for (int i = tid; i < n; i += kBuildThreads) a.mean[i] = static_cast<float>(mu[i]);
  for (int e = tid; e < n * n; e += kBuildThreads) {
    const int i = e / n;
    const int k = e % n;
    a.whitener[e] = (k <= i) ? static_cast<float>(m2[packed(i, k)]) : 0.0f;
  }

  const int T = a.num_targets;
  if (T > 0) {
    // ACE: u_t = L^-1 (d_t - mu), ||u_t||^2.
    for (int e = tid; e < T * n; e += kBuildThreads) {
      const int t = e / n;
      const int i = e % n;
      const float* d = a.targets + static_cast<std::size_t>(t) * n;
      double s = 0.0;
      for (int j = 0; j <= i; ++j) s += m2[packed(i, j)] * (static_cast<double>(d[j]) - mu[j]);
      vec[e] = s;
      a.target_white[e] = static_cast<float>(s);
    }
    __syncthreads();
    for (int t = 0; t < T; ++t) {
      double s = 0.0;
      for (int i = tid; i < n; i += kBuildThreads) s += vec[t * n + i] * vec[t * n + i];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, kWave);
      if ((tid & 63) == 0) red[tid >> 6] = s;
      __syncthreads();
      if (tid == 0) {
        double tot = 0.0;
        for (int w = 0; w < kBuildThreads / kWave; ++w) tot += red[w];
        a.target_norm2[t] = static_cast<float>(tot);
      }
      __syncthreads();
    }

    // CEM: correlation(loading) -> Cholesky -> L_R^-1.
    for (int e = tid; e < tri; e += kBuildThreads) {
      const int i = tri_row(e);
      const int j = e - i * (i + 1) / 2;
      double v = a.sum_outer[static_cast<std::size_t>(i) * n + j] / count;
      if (i == j) v += a.loading;
      m1[e] = v;
    }
    __syncthreads();
    if (!block_cholesky(m1, n, col, diag0)) {
      if (tid == 0) *a.status = HSI_STATUS_CORR_NOT_PD;
      return;
    }
    block_tri_inverse(m1, m2, n);
    // y_t = L_R^-1 d_t
    for (int e = tid; e < T * n; e += kBuildThreads) {
      const int t = e / n;
      const int i = e % n;
      const float* d = a.targets + static_cast<std::size_t>(t) * n;
      double s = 0.0;
      for (int j = 0; j <= i; ++j) s += m2[packed(i, j)] * static_cast<double>(d[j]);
      vec[e] = s;
    }
    __syncthreads();
    // R^-1 d_t = L_R^-T y_t, kept in the (now free) first matrix slot.
    double* solved = m1;
    for (int e = tid; e < T * n; e += kBuildThreads) {
      const int t = e / n;
      const int i = e % n;
      double s = 0.0;
      for (int k = i; k < n; ++k) s += m2[packed(k, i)] * vec[t * n + k];
      solved[e] = s;
    }
    __syncthreads();
    for (int t = 0; t < T; ++t) {
      const float* d = a.targets + static_cast<std::size_t>(t) * n;
      double s = 0.0;
      for (int i = tid; i < n; i += kBuildThreads) s += static_cast<double>(d[i]) * solved[t * n + i];
      for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, kWave);
      if ((tid & 63) == 0) red[tid >> 6] = s;
      __syncthreads();
      double gain = 0.0;
      for (int w = 0; w < kBuildThreads / kWave; ++w) gain += red[w];
      for (int i = tid; i < n; i += kBuildThreads) {
        a.cem_weight[static_cast<std::size_t>(t) * n + i] =
            gain > 0.0 ? static_cast<float>(solved[t * n + i] / gain) : 0.0f;
      }
      __syncthreads();
    }
  }
  if (tid == 0) *a.status = 0u;
}

}  // namespace

std::size_t moments_scratch_doubles(std::size_t pixels, int bands) {
  const std::size_t chunks = div_up(pixels, kChunk);
  const std::size_t wgs = chunks < kMaxMomWgs ? chunks : kMaxMomWgs;
  return wgs * static_cast<std::size_t>(moment_pairs(bands)) * 256;
}

hipError_t launch_moments(const float* cube, std::size_t pixels, int bands, std::size_t stride,
                          hsi_precision precision, double* scratch, double* count, double* sum,
                          double* sum_outer, hipStream_t stream) {
  (void)precision;
  switch (moment_blocks(bands)) {
#define HSI_NB(N) \
  case N: return launch_moments_nb<N>(cube, pixels, bands, stride, scratch, count, sum, sum_outer, stream);
    HSI_NB(1) HSI_NB(2) HSI_NB(3) HSI_NB(4) HSI_NB(5) HSI_NB(6) HSI_NB(7) HSI_NB(8) HSI_NB(9)
    HSI_NB(10) HSI_NB(11) HSI_NB(12) HSI_NB(13) HSI_NB(14) HSI_NB(15) HSI_NB(16) HSI_NB(17)
#undef HSI_NB
  }
  return hipErrorInvalidValue;
}

std::size_t build_work_doubles(int bands, int num_targets) {
  const std::size_t tri = static_cast<std::size_t>(bands) * (bands + 1) / 2;
  return 2 * tri + static_cast<std::size_t>(num_targets > 0 ? num_targets : 1) * bands;
}

hipError_t launch_build(const BuildArgs& a, hipStream_t stream) {
  if (a.bands <= kLdsBands) {
    hipLaunchKernelGGL(build_kernel<true>, dim3(1), dim3(kBuildThreads), 0, stream, a);
  } else {
    hipLaunchKernelGGL(build_kernel<false>, dim3(1), dim3(kBuildThreads), 0, stream, a);
  }
  return hipGetLastError();
}

}  // namespace hsi_amd
