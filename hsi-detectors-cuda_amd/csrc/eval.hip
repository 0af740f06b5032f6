// eval.hip - offline evaluation of a score map against a class mask (gfx950).
//
// The counterpart of the reference's scripts/eval_detection.py (roc_auc at
// :40-66, the operating point and per-class table at :100-140): ROC AUC as
// the Mann-Whitney statistic with ties counted half, precision/recall counts
// at a threshold, and the count, mean and minimum of the score per class.
//
// AUC = sum over positives p of (#negatives below p + #negatives tied with p
// / 2) / (P N), which is what the rank sum with tie-averaged ranks gives. It
// is computed exactly in integers: the negatives' keys are compacted and
// radix-sorted, every positive finds its two bounds by binary search, and
// 2 * below + tied is summed in 64-bit integers (order-free, so the result is
// deterministic). Keys are the score for higher-is-better maps and the
// negated score otherwise; -0 is folded into +0 and NaN ranks above every
// number (tied with other NaNs).
//
// Mask: label l <= max_label is valid (an annotated class), l == target is
// positive. Every per-block partial is reduced in block order.
#include <hipcub/hipcub.hpp>

#include "common.hpp"
#include "kernels.hpp"

namespace hsi_amd {
namespace {

constexpr int kEvalThreads = 256;

struct EvalPartial {
  unsigned pos, neg;
  unsigned tp, fp, fn, tn;
};

__device__ __forceinline__ float eval_key(float s, bool higher) {
  float k = higher ? s : -s;
  if (k != k) k = __builtin_inff();   // NaN: above every number
  return k + 0.0f;                    // -0 -> +0
}

// Per block: positive/negative counts, the confusion counts at the
// threshold, and per-class count/sum/min of the raw score.
__global__ __launch_bounds__(kEvalThreads) void eval_count_kernel(
    const float* __restrict__ score, const std::uint8_t* __restrict__ label, std::size_t n, int target,
    int max_label, int higher, float threshold, EvalPartial* __restrict__ part,
    unsigned* __restrict__ cls_count, double* __restrict__ cls_sum, float* __restrict__ cls_min) {
  __shared__ unsigned s_pos, s_neg, s_tp, s_fp, s_fn, s_tn;
  __shared__ unsigned c_count[kEvalMaxClasses];
  __shared__ float c_min[kEvalMaxClasses];
  __shared__ double c_sum[kEvalThreads];
  if (threadIdx.x == 0) s_pos = s_neg = s_tp = s_fp = s_fn = s_tn = 0u;
  if (threadIdx.x < kEvalMaxClasses) {
    c_count[threadIdx.x] = 0u;
    c_min[threadIdx.x] = __builtin_inff();
  }
  __syncthreads();
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(kEvalThreads) + threadIdx.x;
  int l = -1;
  float s = 0.f;
  if (i < n) {
    l = label[i];
    s = score[i];
    if (l > max_label) l = -1;
  }
  if (l >= 0) {
    const bool pos = l == target;
    const bool pred = higher ? s >= threshold : s <= threshold;
    atomicAdd(pos ? &s_pos : &s_neg, 1u);
    atomicAdd(pred ? (pos ? &s_tp : &s_fp) : (pos ? &s_fn : &s_tn), 1u);
    atomicAdd(&c_count[l], 1u);
    // float min over non-NaN scores (the reference's numpy min would give NaN)
    if (s == s) atomicMin(&c_min[l], s);
  }
  // per-class sums in a fixed order: one class at a time, tree over threads
  for (int c = 0; c <= max_label; ++c) {
    c_sum[threadIdx.x] = l == c ? static_cast<double>(s) : 0.0;
    __syncthreads();
    for (int span = kEvalThreads / 2; span > 0; span >>= 1) {
      if (threadIdx.x < static_cast<unsigned>(span)) c_sum[threadIdx.x] += c_sum[threadIdx.x + span];
      __syncthreads();
    }
    if (threadIdx.x == 0) cls_sum[static_cast<std::size_t>(blockIdx.x) * kEvalMaxClasses + c] = c_sum[0];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = EvalPartial{s_pos, s_neg, s_tp, s_fp, s_fn, s_tn};
  if (threadIdx.x <= static_cast<unsigned>(max_label)) {
    cls_count[static_cast<std::size_t>(blockIdx.x) * kEvalMaxClasses + threadIdx.x] = c_count[threadIdx.x];
    cls_min[static_cast<std::size_t>(blockIdx.x) * kEvalMaxClasses + threadIdx.x] = c_min[threadIdx.x];
  }
}

// One workgroup: exclusive scan of the per-block positive and negative
// counts (compaction offsets), totals, and the per-class reduction in block
// order. out: [0] P, [1] N, [2..5] tp fp fn tn; class c at 8 + 3c: count,
// sum, min (as doubles).
__global__ __launch_bounds__(1024) void eval_scan_kernel(const EvalPartial* __restrict__ part, int blocks,
                                                        int max_label, const unsigned* __restrict__ cls_count,
                                                        const double* __restrict__ cls_sum,
                                                        const float* __restrict__ cls_min,
                                                        unsigned* __restrict__ pos_off, unsigned* __restrict__ neg_off,
                                                        double* __restrict__ out) {
  __shared__ unsigned long long sp[1024], sn[1024];
  const int per = (blocks + 1023) / 1024;
  const int b0 = static_cast<int>(threadIdx.x) * per;
  unsigned long long p = 0, q = 0;
  for (int k = 0; k < per; ++k)
    if (b0 + k < blocks) {
      p += part[b0 + k].pos;
      q += part[b0 + k].neg;
    }
  sp[threadIdx.x] = p;
  sn[threadIdx.x] = q;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const unsigned long long a = threadIdx.x >= static_cast<unsigned>(off) ? sp[threadIdx.x - off] : 0ull;
    const unsigned long long b = threadIdx.x >= static_cast<unsigned>(off) ? sn[threadIdx.x - off] : 0ull;
    __syncthreads();
    sp[threadIdx.x] += a;
    sn[threadIdx.x] += b;
    __syncthreads();
  }
  unsigned long long rp = sp[threadIdx.x] - p, rn = sn[threadIdx.x] - q;
  for (int k = 0; k < per; ++k)
    if (b0 + k < blocks) {
      pos_off[b0 + k] = static_cast<unsigned>(rp);
      neg_off[b0 + k] = static_cast<unsigned>(rn);
      rp += part[b0 + k].pos;
      rn += part[b0 + k].neg;
    }
  if (threadIdx.x == 0) {
    unsigned long long tp = 0, fp = 0, fn = 0, tn = 0;
    for (int b = 0; b < blocks; ++b) {
      tp += part[b].tp;
      fp += part[b].fp;
      fn += part[b].fn;
      tn += part[b].tn;
    }
    out[0] = static_cast<double>(sp[1023]);
    out[1] = static_cast<double>(sn[1023]);
    out[2] = static_cast<double>(tp);
    out[3] = static_cast<double>(fp);
    out[4] = static_cast<double>(fn);
    out[5] = static_cast<double>(tn);
  }
  if (threadIdx.x <= static_cast<unsigned>(max_label)) {
    const int c = static_cast<int>(threadIdx.x);
    unsigned long long cnt = 0;
    double sum = 0.0;
    float mn = __builtin_inff();
    for (int b = 0; b < blocks; ++b) {
      const std::size_t e = static_cast<std::size_t>(b) * kEvalMaxClasses + c;
      cnt += cls_count[e];
      sum += cls_sum[e];
      mn = fminf(mn, cls_min[e]);
    }
    out[8 + 3 * c] = static_cast<double>(cnt);
    out[8 + 3 * c + 1] = cnt ? sum / static_cast<double>(cnt) : 0.0;
    out[8 + 3 * c + 2] = static_cast<double>(mn);
  }
}

// Keys of the valid pixels, positives and negatives apart, in raster order
// (block offsets from the scan, ranks within the block from an LDS scan).
__global__ __launch_bounds__(kEvalThreads) void eval_split_kernel(
    const float* __restrict__ score, const std::uint8_t* __restrict__ label, std::size_t n, int target,
    int max_label, int higher, const unsigned* __restrict__ pos_off, const unsigned* __restrict__ neg_off,
    float* __restrict__ pos_keys, float* __restrict__ neg_keys) {
  __shared__ unsigned sp[kEvalThreads], sn[kEvalThreads];
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(kEvalThreads) + threadIdx.x;
  int l = -1;
  float s = 0.f;
  if (i < n) {
    l = label[i];
    s = score[i];
    if (l > max_label) l = -1;
  }
  const unsigned fp = l >= 0 && l == target, fnn = l >= 0 && l != target;
  sp[threadIdx.x] = fp;
  sn[threadIdx.x] = fnn;
  __syncthreads();
  for (int off = 1; off < kEvalThreads; off <<= 1) {
    const unsigned a = threadIdx.x >= static_cast<unsigned>(off) ? sp[threadIdx.x - off] : 0u;
    const unsigned b = threadIdx.x >= static_cast<unsigned>(off) ? sn[threadIdx.x - off] : 0u;
    __syncthreads();
    sp[threadIdx.x] += a;
    sn[threadIdx.x] += b;
    __syncthreads();
  }
  const float k = eval_key(s, higher != 0);
  if (fp) pos_keys[pos_off[blockIdx.x] + sp[threadIdx.x] - 1] = k;
  if (fnn) neg_keys[neg_off[blockIdx.x] + sn[threadIdx.x] - 1] = k;
}

// 2 * (#negatives below) + (#negatives tied) per positive, summed exactly.
__global__ __launch_bounds__(kEvalThreads) void eval_auc_kernel(const float* __restrict__ pos_keys,
                                                               std::size_t npos,
                                                               const float* __restrict__ neg_sorted,
                                                               std::size_t nneg,
                                                               unsigned long long* __restrict__ acc) {
  __shared__ unsigned long long red[kEvalThreads];
  const std::size_t i = blockIdx.x * static_cast<std::size_t>(kEvalThreads) + threadIdx.x;
  unsigned long long v = 0;
  if (i < npos) {
    const float k = pos_keys[i];
    std::size_t lo = 0, hi = nneg;   // first index with key >= k
    while (lo < hi) {
      const std::size_t mid = (lo + hi) / 2;
      if (neg_sorted[mid] < k) lo = mid + 1;
      else hi = mid;
    }
    const std::size_t below = lo;
    hi = nneg;                       // first index with key > k
    while (lo < hi) {
      const std::size_t mid = (lo + hi) / 2;
      if (!(k < neg_sorted[mid])) lo = mid + 1;
      else hi = mid;
    }
    v = 2ull * below + (lo - below);
  }
  red[threadIdx.x] = v;
  __syncthreads();
  for (int span = kEvalThreads / 2; span > 0; span >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(span)) red[threadIdx.x] += red[threadIdx.x + span];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0]) atomicAdd(acc, red[0]);
}

__global__ void eval_zero_kernel(unsigned long long* acc) { *acc = 0ull; }

}  // namespace

constexpr int kEvalOut = 8 + 3 * kEvalMaxClasses;

std::size_t sort_temp_bytes(std::size_t n) {
  std::size_t temp = 0;
  if (hipcub::DeviceRadixSort::SortKeys(nullptr, temp, static_cast<const float*>(nullptr), static_cast<float*>(nullptr),
                                        static_cast<int>(n > 0 ? n : 1)) != hipSuccess)
    return 0;
  return temp;
}

std::size_t eval_scratch_bytes(std::size_t n) {
  const std::size_t blocks = div_up(n, kEvalThreads);
  const std::size_t parts[] = {blocks * sizeof(EvalPartial), blocks * kEvalMaxClasses * sizeof(unsigned),
                               blocks * kEvalMaxClasses * sizeof(double), blocks * kEvalMaxClasses * sizeof(float),
                               blocks * sizeof(unsigned), blocks * sizeof(unsigned), n * sizeof(float),
                               n * sizeof(float), n * sizeof(float), kEvalOut * sizeof(double),
                               sizeof(unsigned long long), sort_temp_bytes(n)};
  std::size_t total = 0;
  for (std::size_t b : parts) total += round_up(b, 256);
  return total;
}

hipError_t launch_evaluate(const float* score, const std::uint8_t* label, std::size_t n, int target, int max_label,
                           int higher, float threshold, void* scratch, std::size_t scratch_bytes, double* out_host,
                           hipStream_t stream) {
  if (n == 0 || n > 0x7fffffffu || max_label < 0 || max_label >= kEvalMaxClasses) return hipErrorInvalidValue;
  if (scratch_bytes < eval_scratch_bytes(n)) return hipErrorInvalidValue;
  const std::size_t blocks = div_up(n, kEvalThreads);
  auto carve = [&](std::size_t bytes) {
    char* p = static_cast<char*>(scratch);
    scratch = p + round_up(bytes, 256);
    return static_cast<void*>(p);
  };
  auto* part = static_cast<EvalPartial*>(carve(blocks * sizeof(EvalPartial)));
  auto* cls_count = static_cast<unsigned*>(carve(blocks * kEvalMaxClasses * sizeof(unsigned)));
  auto* cls_sum = static_cast<double*>(carve(blocks * kEvalMaxClasses * sizeof(double)));
  auto* cls_min = static_cast<float*>(carve(blocks * kEvalMaxClasses * sizeof(float)));
  auto* pos_off = static_cast<unsigned*>(carve(blocks * sizeof(unsigned)));
  auto* neg_off = static_cast<unsigned*>(carve(blocks * sizeof(unsigned)));
  auto* pos_keys = static_cast<float*>(carve(n * sizeof(float)));
  auto* neg_keys = static_cast<float*>(carve(n * sizeof(float)));
  auto* neg_sorted = static_cast<float*>(carve(n * sizeof(float)));
  auto* out = static_cast<double*>(carve(kEvalOut * sizeof(double)));
  auto* acc = static_cast<unsigned long long*>(carve(sizeof(unsigned long long)));
  std::size_t temp_bytes = sort_temp_bytes(n);
  void* temp = carve(temp_bytes);

  hipLaunchKernelGGL(eval_count_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kEvalThreads), 0, stream, score,
                     label, n, target, max_label, higher, threshold, part, cls_count, cls_sum, cls_min);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(eval_scan_kernel, dim3(1), dim3(1024), 0, stream, part, static_cast<int>(blocks), max_label,
                     cls_count, cls_sum, cls_min, pos_off, neg_off, out);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(eval_split_kernel, dim3(static_cast<unsigned>(blocks)), dim3(kEvalThreads), 0, stream, score,
                     label, n, target, max_label, higher, pos_off, neg_off, pos_keys, neg_keys);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  // The split's counts decide the sort and search sizes: read them back.
  double head[2];
  if ((e = hipMemcpyAsync(head, out, sizeof(head), hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  const std::size_t npos = static_cast<std::size_t>(head[0]), nneg = static_cast<std::size_t>(head[1]);
  hipLaunchKernelGGL(eval_zero_kernel, dim3(1), dim3(1), 0, stream, acc);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if (npos > 0 && nneg > 0) {
    if ((e = hipcub::DeviceRadixSort::SortKeys(temp, temp_bytes, neg_keys, neg_sorted, static_cast<int>(nneg), 0,
                                               32, stream)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(eval_auc_kernel, dim3(static_cast<unsigned>(div_up(npos, kEvalThreads))), dim3(kEvalThreads),
                       0, stream, pos_keys, npos, neg_sorted, nneg, acc);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  const std::size_t nout = 8 + 3 * static_cast<std::size_t>(max_label + 1);
  if ((e = hipMemcpyAsync(out_host, out, nout * sizeof(double), hipMemcpyDeviceToHost, stream)) != hipSuccess)
    return e;
  unsigned long long num = 0;
  if ((e = hipMemcpyAsync(&num, acc, sizeof(num), hipMemcpyDeviceToHost, stream)) != hipSuccess) return e;
  if ((e = hipStreamSynchronize(stream)) != hipSuccess) return e;
  // slot 6: the AUC numerator (2 * below + tied), exact while below 2^53
  out_host[6] = static_cast<double>(num);
  return hipSuccess;
}

}  // namespace hsi_amd
