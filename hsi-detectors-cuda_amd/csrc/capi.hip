// capi.hip - the C ABI (include/hsi_mi355x.h): context, validation, dispatch.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <mutex>
#include <string>
#include <vector>

#include "common.hpp"
#include "kernels.hpp"

using namespace hsi_amd;

struct hsi_ctx {
  int device = 0;
  std::string error;

  // Target library.
  int num_targets = 0;
  int lib_bands = 0;
  int lib_stride = 1;
  float* d_targets = nullptr;  // [kMaxTargets][kMaxBands] capacity, [T][B] used
  float* d_norms = nullptr;    // [kMaxTargets]
  double* d_sam_lib = nullptr; // [kMaxBands][kMaxTargets] capacity
  double* d_sam_rnorm = nullptr; // [kMaxTargets] 1/norm, for ranking
  int lib_degenerate = 0;        // some target has norm <= 0

  // Moments and statistics, allocated once at the largest supported size.
  int moment_bands = 0;        // bands of the last accumulation, 0 if none
  int stats_bands = 0;         // bands of the current statistics, 0 if none
  int stats_targets = 0;       // targets the statistics were built with
  bool layouts_dirty = false;  // plain score layouts not yet derived from the statistics
  bool mixed_dirty = false;    // mixed-precision layouts not yet derived
  double last_loading = 0.0;
  double* d_count = nullptr;
  double* d_sum = nullptr;
  double* d_sum_outer = nullptr;
  float* d_mean = nullptr;
  float* d_whitener = nullptr;
  float* d_target_white = nullptr;
  float* d_target_norm2 = nullptr;
  float* d_cem_weight = nullptr;
  unsigned* d_status = nullptr;      // [2]: covariance half, autocorrelation half
  double* d_build_work = nullptr;
  double* d_cem_lib = nullptr;
  float* d_white_blocks = nullptr;
  float* d_ace_u = nullptr;
  _Float16* d_white_hl = nullptr;
  _Float16* d_ace_hl = nullptr;
  float* d_col_scale = nullptr;
  float* d_row_inv = nullptr;
  float* d_target_inv = nullptr;

  // Learned pre-screen: FP32 masters, Adam moments, gradients, bf16 layouts.
  int mlp_bands = 0;
  long long mlp_step = 0;
  bool mlp_dirty = false;
  float* d_mlp_w = nullptr;
  float* d_mlp_m = nullptr;
  float* d_mlp_v = nullptr;
  float* d_mlp_grad = nullptr;
  void* d_mlp_a1 = nullptr;
  void* d_mlp_a2 = nullptr;
  float* d_mlp_tail = nullptr;
  void* d_mlp_a1h = nullptr;   // FP16 copies for the fused statistics + screening pass
  void* d_mlp_a2h = nullptr;
  float* d_mlp_act = nullptr;
  std::size_t mlp_act_bytes = 0;
  // Evaluation scratch.
  void* d_eval = nullptr;
  std::size_t eval_bytes = 0;
  // Screening scratch: per-wave regions of records and spectra, counters.
  void* d_screen_rec = nullptr;
  std::size_t screen_rec_bytes = 0;
  void* d_screen_counts = nullptr;
  std::size_t screen_counts_bytes = 0;

  // Q16 code (hsi_encode_q16 / hsi_set_q16_code): per-band offset and scale.
  int q16_bands = 0;
  float* d_q16_offset = nullptr;
  float* d_q16_scale = nullptr;
  unsigned* d_q16_extrema = nullptr;

  // Demosaic band positions (uploaded when they change).
  int* d_positions = nullptr;
  std::vector<int> positions;

  // Scratch sized per frame.
  double* d_mom_scratch = nullptr;
  std::size_t mom_scratch_bytes = 0;
  unsigned* d_det_scratch = nullptr;
  std::size_t det_scratch_bytes = 0;
};

namespace {

hsi_status fail(hsi_ctx* ctx, hsi_status status, const std::string& message) {
  if (ctx) ctx->error = message;
  return status;
}

hsi_status device_fail(hsi_ctx* ctx, const char* what, hipError_t err) {
  return fail(ctx, HSI_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(err));
}

// The mixed-form moments read the cube through one buffer view.
const char* const kMixedViewLimit =
    "the mixed-form statistics read the cube through one buffer view: bands x plane_stride elements must stay "
    "under 4 GiB";

#define HSI_HIP(ctx, expr, what)                          \
  do {                                                    \
    const hipError_t hsi_err_ = (expr);                   \
    if (hsi_err_ != hipSuccess) return device_fail(ctx, what, hsi_err_); \
  } while (0)

template <typename T>
hipError_t alloc(T** ptr, std::size_t count) {
  return hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(T) + 256);
}

hipError_t grow(void** ptr, std::size_t* have, std::size_t need) {
  if (*have >= need && *ptr) return hipSuccess;
  if (*ptr) {
    hipError_t err = hipFree(*ptr);
    if (err != hipSuccess) return err;
    *ptr = nullptr;
    *have = 0;
  }
  hipError_t err = hipMalloc(ptr, need);
  if (err == hipSuccess) *have = need;
  return err;
}

// Q16 cubes are accepted only where allow_q16 (the mixed statistics and the
// mixed RX/ACE kernels).
hsi_status check_cube(hsi_ctx* ctx, const void* d_cube, int height, int width, int bands,
                      std::size_t stride, hsi_dtype dtype, bool allow_q16 = false) {
  if (!d_cube) return fail(ctx, HSI_ERR_INVALID_ARG, "null cube");
  if (height <= 0 || width <= 0 || bands <= 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "cube shape must be positive");
  if (bands > kMaxBands)
    return fail(ctx, HSI_ERR_CAPACITY, "cube has " + std::to_string(bands) +
                                           " bands; at most " + std::to_string(kMaxBands) + " are supported");
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  if (stride < pixels || stride % 8 != 0)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "plane stride must be at least height*width and a multiple of 8 (use hsi_plane_stride)");
  if (dtype == HSI_Q16 && !allow_q16)
    return fail(ctx, HSI_ERR_INVALID_ARG, "Q16 cubes are read by the MIXED statistics and MIXED RX/ACE only");
  if (dtype != HSI_F32 && dtype != HSI_F16 && dtype != HSI_Q16)
    return fail(ctx, HSI_ERR_INVALID_ARG, "unknown cube dtype");
  if (reinterpret_cast<std::uintptr_t>(d_cube) % 16 != 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "cube must be 16-byte aligned");
  return HSI_OK;
}

hsi_status set_device(hsi_ctx* ctx) {
  HSI_HIP(ctx, hipSetDevice(ctx->device), "select device");
  return HSI_OK;
}

// Device address of hsi_detect_no_alloc's static scratch, per device: looked
// up once, by the first context created on that device (with the device
// current), and kept for the life of the process.
constexpr int kMaxDevices = 64;
std::mutex g_no_alloc_mutex;
unsigned* g_no_alloc_scratch[kMaxDevices] = {};
bool g_no_alloc_looked_up[kMaxDevices] = {};

void resolve_no_alloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lock(g_no_alloc_mutex);
  if (g_no_alloc_looked_up[device]) return;
  g_no_alloc_looked_up[device] = true;
  unsigned* ptr = nullptr;
  if (detect_no_alloc_scratch(&ptr) == hipSuccess) g_no_alloc_scratch[device] = ptr;
}

unsigned* no_alloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_no_alloc_mutex);
  return g_no_alloc_scratch[device];
}

hsi_status mlp_ready(hsi_ctx* ctx, int bands);  // a network is installed for `bands`
}  // namespace

extern "C" {

const char* hsi_status_name(hsi_status status) {
  switch (status) {
    case HSI_OK: return "ok";
    case HSI_ERR_INVALID_ARG: return "invalid argument";
    case HSI_ERR_CAPACITY: return "capacity exceeded";
    case HSI_ERR_NOT_POSITIVE_DEFINITE: return "not positive definite";
    case HSI_ERR_DEVICE: return "device error";
    case HSI_ERR_NO_STATISTICS: return "no statistics";
    case HSI_ERR_NO_LIBRARY: return "no library";
  }
  return "unknown status";
}

const char* hsi_ctx_last_error(const hsi_ctx* ctx) { return ctx ? ctx->error.c_str() : ""; }

// Band planes of a large frame start an odd number of 4 KiB pages apart. With
// a power-of-two plane size (1024x1024, 2048x2048) every plane would start on
// the same HBM channel, and the loads a wave has in flight across planes
// would queue on it: a BSQ read of 2048x2048x113 runs at 5.9 TB/s with
// planes 16 MiB apart and at 7.0 TB/s with them 16 MiB + 4 KiB apart.
constexpr std::size_t kPageFloats = 1024;         // 4 KiB of FP32
constexpr std::size_t kPadFromPixels = 1u << 16;  // smaller planes: multiple of 8 only

size_t hsi_plane_stride(int height, int width) {
  if (height <= 0 || width <= 0) return 0;
  const std::size_t pixels = static_cast<std::size_t>(height) * static_cast<std::size_t>(width);
  if (pixels < kPadFromPixels) return round_up(pixels, 8);
  std::size_t s = round_up(pixels, kPageFloats);
  if ((s / kPageFloats) % 2 == 0) s += kPageFloats;
  return s;
}

int hsi_max_targets(void) { return kMaxTargets; }
int hsi_max_bands(void) { return kMaxBands; }

void hsi_ctx_destroy(hsi_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  void* ptrs[] = {ctx->d_targets, ctx->d_norms, ctx->d_sam_lib, ctx->d_sam_rnorm, ctx->d_count, ctx->d_sum,
                  ctx->d_sum_outer, ctx->d_mean, ctx->d_whitener, ctx->d_target_white,
                  ctx->d_target_norm2, ctx->d_cem_weight, ctx->d_status, ctx->d_build_work,
                  ctx->d_cem_lib, ctx->d_white_blocks, ctx->d_ace_u, ctx->d_white_hl,
                  ctx->d_ace_hl, ctx->d_col_scale, ctx->d_row_inv, ctx->d_target_inv,
                  ctx->d_mom_scratch, ctx->d_det_scratch, ctx->d_mlp_w, ctx->d_mlp_m, ctx->d_mlp_v,
                  ctx->d_mlp_grad, ctx->d_mlp_a1, ctx->d_mlp_a2, ctx->d_mlp_tail, ctx->d_mlp_act,
                  ctx->d_mlp_a1h, ctx->d_mlp_a2h,
                  ctx->d_positions, ctx->d_screen_rec, ctx->d_screen_counts,
                  ctx->d_eval, ctx->d_q16_offset, ctx->d_q16_scale, ctx->d_q16_extrema};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  delete ctx;
}

hsi_status hsi_ctx_create(int device, hsi_ctx** out) {
  if (!out) return HSI_ERR_INVALID_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count)
    return HSI_ERR_DEVICE;
  hsi_ctx* ctx = new hsi_ctx();
  ctx->device = device;
  const std::size_t B = kMaxBands, T = kMaxTargets;
  hipError_t err = hipSetDevice(device);
  if (err == hipSuccess) err = alloc(&ctx->d_targets, T * B);
  if (err == hipSuccess) err = alloc(&ctx->d_norms, T);
  if (err == hipSuccess) err = alloc(&ctx->d_sam_lib, B * T);
  if (err == hipSuccess) err = alloc(&ctx->d_sam_rnorm, T);
  if (err == hipSuccess) err = alloc(&ctx->d_count, 1);
  if (err == hipSuccess) err = alloc(&ctx->d_sum, B);
  if (err == hipSuccess) err = alloc(&ctx->d_sum_outer, B * B);
  if (err == hipSuccess) err = alloc(&ctx->d_mean, B);
  if (err == hipSuccess) err = alloc(&ctx->d_whitener, B * B);
  if (err == hipSuccess) err = alloc(&ctx->d_target_white, T * B);
  if (err == hipSuccess) err = alloc(&ctx->d_target_norm2, T);
  if (err == hipSuccess) err = alloc(&ctx->d_cem_weight, T * B);
  if (err == hipSuccess) err = alloc(&ctx->d_status, 2);
  if (err == hipSuccess) err = alloc(&ctx->d_build_work, build_work_doubles(kMaxBands, kMaxTargets));
  if (err == hipSuccess) err = alloc(&ctx->d_cem_lib, B * T);
  if (err == hipSuccess) err = alloc(&ctx->d_white_blocks, rxace_white_floats(kMaxBands));
  if (err == hipSuccess) err = alloc(&ctx->d_ace_u, rxace_target_floats(kMaxBands, kMaxTargets));
  if (err == hipSuccess) err = alloc(&ctx->d_white_hl, mixed_white_halves(kMaxBands));
  if (err == hipSuccess) err = alloc(&ctx->d_ace_hl, mixed_target_halves(kMaxBands, kMaxTargets));
  if (err == hipSuccess) err = alloc(&ctx->d_col_scale, B);
  if (err == hipSuccess) err = alloc(&ctx->d_row_inv, B);
  if (err == hipSuccess) err = alloc(&ctx->d_target_inv, T);
  if (err == hipSuccess) err = alloc(&ctx->d_q16_offset, B);
  if (err == hipSuccess) err = alloc(&ctx->d_q16_scale, B);
  if (err == hipSuccess) err = alloc(&ctx->d_q16_extrema, q16_extrema_words(kMaxBands));
  if (err == hipSuccess) err = hipMemset(ctx->d_sum_outer, 0, B * B * sizeof(double));
  if (err == hipSuccess) err = hipMemset(ctx->d_status, 0, 2 * sizeof(unsigned));
  if (err == hipSuccess) err = hipDeviceSynchronize();
  if (err != hipSuccess) {
    hsi_ctx_destroy(ctx);
    return HSI_ERR_DEVICE;
  }
  resolve_no_alloc_scratch(device);
  *out = ctx;
  return HSI_OK;
}

hsi_status hsi_ctx_reserve(hsi_ctx* ctx, size_t max_pixels, int max_bands) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (max_pixels == 0 || max_bands <= 0 || max_bands > kMaxBands)
    return fail(ctx, HSI_ERR_INVALID_ARG, "reserve: bad size");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  std::size_t mom = 0;
  for (int b = 1; b <= max_bands; ++b) {
    const std::size_t need = moments_scratch_doubles(max_pixels, b) * sizeof(double);
    if (need > mom) mom = need;
  }
  HSI_HIP(ctx, grow(reinterpret_cast<void**>(&ctx->d_mom_scratch), &ctx->mom_scratch_bytes, mom),
          "reserve moment scratch");
  HSI_HIP(ctx, grow(reinterpret_cast<void**>(&ctx->d_det_scratch), &ctx->det_scratch_bytes,
                    detect_scratch_words(max_pixels) * sizeof(unsigned)),
          "reserve detect scratch");
  return HSI_OK;
}

hsi_status hsi_set_library(hsi_ctx* ctx, const float* values, const float* norms, int num_targets,
                           int bands) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (!values || num_targets <= 0 || bands <= 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "library needs at least one target and one band");
  if (num_targets > kMaxTargets || bands > kMaxBands)
    return fail(ctx, HSI_ERR_CAPACITY,
                "library of " + std::to_string(num_targets) + " targets x " + std::to_string(bands) +
                    " bands exceeds the supported " + std::to_string(kMaxTargets) + " x " +
                    std::to_string(kMaxBands));
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  // Norms as SpectralLibrary::norms forms them: double sum, float root.
  std::vector<float> n(static_cast<std::size_t>(num_targets));
  for (int t = 0; t < num_targets; ++t) {
    if (norms) {
      n[t] = norms[t];
    } else {
      double acc = 0.0;
      for (int b = 0; b < bands; ++b) {
        const double v = values[static_cast<std::size_t>(t) * bands + b];
        acc += v * v;
      }
      n[t] = static_cast<float>(std::sqrt(acc));
    }
  }
  const int ls = sam_cem_lib_stride(num_targets);
  std::vector<double> lib(static_cast<std::size_t>(bands) * ls, 0.0);
  for (int t = 0; t < num_targets; ++t)
    for (int b = 0; b < bands; ++b)
      lib[static_cast<std::size_t>(b) * ls + t] = values[static_cast<std::size_t>(t) * bands + b];
  HSI_HIP(ctx, hipMemcpy(ctx->d_targets, values, sizeof(float) * num_targets * bands,
                         hipMemcpyHostToDevice), "upload library");
  HSI_HIP(ctx, hipMemcpy(ctx->d_norms, n.data(), sizeof(float) * num_targets, hipMemcpyHostToDevice),
          "upload norms");
  HSI_HIP(ctx, hipMemcpy(ctx->d_sam_lib, lib.data(), sizeof(double) * lib.size(), hipMemcpyHostToDevice),
          "upload library layout");
  std::vector<double> rn(static_cast<std::size_t>(kMaxTargets), 0.0);
  int degenerate = 0;
  for (int t = 0; t < num_targets; ++t) {
    if (n[t] > 0.0f) rn[t] = 1.0 / static_cast<double>(n[t]);
    else degenerate = 1;
  }
  HSI_HIP(ctx, hipMemcpy(ctx->d_sam_rnorm, rn.data(), sizeof(double) * rn.size(), hipMemcpyHostToDevice),
          "upload reciprocal norms");
  ctx->lib_degenerate = degenerate;
  ctx->num_targets = num_targets;
  ctx->lib_bands = bands;
  ctx->lib_stride = ls;
  ctx->stats_bands = 0;  // statistics embed the library; rebuild them
  return HSI_OK;
}

hsi_status hsi_stats_accumulate(hsi_ctx* ctx, const void* d_cube, hsi_dtype dtype, int height, int width,
                                int bands, size_t plane_stride, hsi_precision precision,
                                const float* d_shift_or_null, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, dtype, true);
  if (st != HSI_OK) return st;
  if (precision != HSI_PLAIN && precision != HSI_MIXED)
    return fail(ctx, HSI_ERR_INVALID_ARG, "unknown precision");
  if (precision == HSI_PLAIN && dtype != HSI_F32)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "the plain (FP64) moments read F32 cubes; use HSI_MIXED for F16 and Q16 cubes");
  if (dtype == HSI_Q16 && ctx->q16_bands != bands)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "no Q16 code for " + std::to_string(bands) + " bands: encode with hsi_encode_q16 or install one "
                "with hsi_set_q16_code");
  if (precision == HSI_MIXED && !mixed_moments_fit(bands, plane_stride, dtype == HSI_F32 ? 4 : 2, false))
    return fail(ctx, HSI_ERR_INVALID_ARG, kMixedViewLimit);
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  HSI_HIP(ctx, grow(reinterpret_cast<void**>(&ctx->d_mom_scratch), &ctx->mom_scratch_bytes,
                    moments_scratch_doubles(pixels, bands) * sizeof(double)),
          "allocate moment scratch");
  MomentArgs m;
  m.cube = d_cube;
  m.dtype = dtype;
  m.pixels = pixels;
  m.bands = bands;
  m.stride = plane_stride;
  m.precision = precision;
  m.shift = precision == HSI_MIXED ? d_shift_or_null : nullptr;
  m.q_offset = ctx->d_q16_offset;
  m.q_scale = ctx->d_q16_scale;
  m.count = ctx->d_count;
  m.sum = ctx->d_sum;
  m.sum_outer = ctx->d_sum_outer;
  HSI_HIP(ctx, launch_moments(m, ctx->d_mom_scratch, static_cast<hipStream_t>(stream)), "accumulate moments");
  ctx->moment_bands = bands;
  return HSI_OK;
}

hsi_status hsi_stats_accumulate_screen(hsi_ctx* ctx, const float* d_cube, int height, int width, int bands,
                                       size_t plane_stride, const float* d_shift_or_null, float cut,
                                       hsi_detection* d_candidates, uint32_t* d_count, int capacity,
                                       float* d_spectra, size_t spectra_stride, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, HSI_F32);
  if (st != HSI_OK) return st;
  if (bands > kFusedScreenMaxBands)
    return fail(ctx, HSI_ERR_CAPACITY, "statistics with screening support at most " +
                                           std::to_string(kFusedScreenMaxBands) +
                                           " bands (use hsi_stats_accumulate and hsi_mlp_screen)");
  if ((st = mlp_ready(ctx, bands)) != HSI_OK) return st;
  if (!d_candidates || !d_count || capacity < 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "screening needs a candidate list, its count and a capacity >= 0");
  const std::size_t pb = static_cast<std::size_t>((bands + 31) / 32 * 32);
  if (d_spectra && spectra_stride < pb)
    return fail(ctx, HSI_ERR_INVALID_ARG, "candidate spectra need a pitch >= 32 * ceil(bands / 32)");
  if (!mixed_moments_fit(bands, plane_stride, 4, true)) return fail(ctx, HSI_ERR_INVALID_ARG, kMixedViewLimit);
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (ctx->mlp_dirty) {
    HSI_HIP(ctx, launch_mlp_layouts(ctx->d_mlp_w, bands, ctx->d_mlp_a1, ctx->d_mlp_a2, ctx->d_mlp_tail,
                                    ctx->d_mlp_a1h, ctx->d_mlp_a2h, s),
            "pre-screen layouts");
    ctx->mlp_dirty = false;
  }
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  HSI_HIP(ctx, grow(reinterpret_cast<void**>(&ctx->d_mom_scratch), &ctx->mom_scratch_bytes,
                    moments_scratch_doubles(pixels, bands) * sizeof(double)),
          "allocate moment scratch");
  MomentArgs m;
  m.cube = d_cube;
  m.dtype = HSI_F32;
  m.pixels = pixels;
  m.bands = bands;
  m.stride = plane_stride;
  m.precision = HSI_MIXED;
  m.shift = d_shift_or_null;
  m.count = ctx->d_count;
  m.sum = ctx->d_sum;
  m.sum_outer = ctx->d_sum_outer;
  FusedScreenArgs fs;
  fs.cut = cut;
  fs.width = width;
  fs.w1h = ctx->d_mlp_a1h;
  fs.w2h = ctx->d_mlp_a2h;
  fs.tail = ctx->d_mlp_tail;
  fs.cand = d_candidates;
  fs.count = d_count;
  fs.cap = static_cast<unsigned>(capacity);
  fs.spectra = d_spectra;
  fs.spectra_stride = spectra_stride;
  const std::size_t cap = static_cast<std::size_t>(capacity);
  HSI_HIP(ctx, grow(&ctx->d_screen_rec, &ctx->screen_rec_bytes, (cap ? cap : 1) * sizeof(hsi_detection)),
          "screening scratch");
  HSI_HIP(ctx, grow(&ctx->d_screen_counts, &ctx->screen_counts_bytes, 2 * kScreenMaxWaves * sizeof(unsigned)),
          "screening scratch");
  fs.scratch_rec = static_cast<hsi_detection*>(ctx->d_screen_rec);
  fs.counts = static_cast<unsigned*>(ctx->d_screen_counts);
  HSI_HIP(ctx, launch_moments_screen(m, fs, ctx->d_mom_scratch, s), "accumulate moments with screening");
  ctx->moment_bands = bands;
  return HSI_OK;
}

hsi_status hsi_encode_q16(hsi_ctx* ctx, const float* d_src, int height, int width, int bands, size_t plane_stride,
                          uint16_t* d_dst, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_src, height, width, bands, plane_stride, HSI_F32);
  if (st != HSI_OK) return st;
  if (!d_dst || reinterpret_cast<std::uintptr_t>(d_dst) % 16 != 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "Q16 destination must be a 16-byte aligned device buffer");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, launch_encode_q16(d_src, static_cast<std::size_t>(height) * width, bands, plane_stride,
                                 ctx->d_q16_extrema, d_dst, ctx->d_q16_offset, ctx->d_q16_scale,
                                 static_cast<hipStream_t>(stream)),
          "encode Q16");
  ctx->q16_bands = bands;
  return HSI_OK;
}

hsi_status hsi_set_q16_code(hsi_ctx* ctx, int bands, const float* offset, const float* scale) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (bands <= 0 || bands > kMaxBands || !offset || !scale)
    return fail(ctx, HSI_ERR_INVALID_ARG, "Q16 code needs 1.." + std::to_string(kMaxBands) + " bands of offsets and scales");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, hipMemcpy(ctx->d_q16_offset, offset, bands * sizeof(float), hipMemcpyHostToDevice), "upload Q16 code");
  HSI_HIP(ctx, hipMemcpy(ctx->d_q16_scale, scale, bands * sizeof(float), hipMemcpyHostToDevice), "upload Q16 code");
  ctx->q16_bands = bands;
  return HSI_OK;
}

hsi_status hsi_get_q16_code(hsi_ctx* ctx, int bands, float* offset, float* scale) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (ctx->q16_bands <= 0 || bands != ctx->q16_bands)
    return fail(ctx, HSI_ERR_NO_STATISTICS, "no Q16 code for this band count");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, hipDeviceSynchronize(), "wait for device");
  if (offset)
    HSI_HIP(ctx, hipMemcpy(offset, ctx->d_q16_offset, bands * sizeof(float), hipMemcpyDeviceToHost), "read Q16 code");
  if (scale)
    HSI_HIP(ctx, hipMemcpy(scale, ctx->d_q16_scale, bands * sizeof(float), hipMemcpyDeviceToHost), "read Q16 code");
  return HSI_OK;
}

// The score kernels read the statistics in derived layouts (band-major
// doubles for CEM, MFMA-ordered blocks for RX/ACE, scaled FP16 hi/lo parts for
// the mixed form). They are derived on the first score that needs them after
// each build or upload, on that score's stream, so a frame that scores only
// SAM and CEM never pays for the RX/ACE layouts.
static hsi_status ensure_layouts(hsi_ctx* ctx, hipStream_t stream) {
  if (!ctx->layouts_dirty) return HSI_OK;
  const int targets = ctx->stats_targets;
  LayoutArgs la;
  la.bands = ctx->stats_bands;
  la.num_targets = targets;
  la.lib_stride = sam_cem_lib_stride(targets > 0 ? targets : 1);
  la.whitener = ctx->d_whitener;
  la.target_white = ctx->d_target_white;
  la.cem_weight = ctx->d_cem_weight;
  la.mean = ctx->d_mean;
  la.cem_lib = targets > 0 ? ctx->d_cem_lib : nullptr;
  la.white_blocks = ctx->d_white_blocks;
  la.ace_u = ctx->d_ace_u;
  HSI_HIP(ctx, launch_layouts(la, stream), "derive statistics layouts");
  ctx->layouts_dirty = false;
  return HSI_OK;
}

static hsi_status ensure_mixed_layouts(hsi_ctx* ctx, hipStream_t stream) {
  if (!ctx->mixed_dirty) return HSI_OK;
  MixedLayoutArgs ma;
  ma.bands = ctx->stats_bands;
  ma.num_targets = ctx->stats_targets;
  ma.whitener = ctx->d_whitener;
  ma.target_white = ctx->d_target_white;
  ma.white_hl = ctx->d_white_hl;
  ma.ace_hl = ctx->d_ace_hl;
  ma.col_scale = ctx->d_col_scale;
  ma.row_inv = ctx->d_row_inv;
  ma.target_inv = ctx->d_target_inv;
  HSI_HIP(ctx, launch_mixed_layouts(ma, stream), "derive mixed-precision layouts");
  ctx->mixed_dirty = false;
  return HSI_OK;
}

}  // extern "C"

// The build of a context's accumulated moments: checks, then its arguments
// (device pointers only, so a batch can keep them in a table).
static void fill_build_args(hsi_ctx* ctx, int bands, double diagonal_loading, BuildArgs* out) {
  BuildArgs& a = *out;
  a.bands = bands;
  a.num_targets = ctx->num_targets;
  a.loading = diagonal_loading;
  a.count = ctx->d_count;
  a.sum = ctx->d_sum;
  a.sum_outer = ctx->d_sum_outer;
  a.targets = ctx->d_targets;
  a.work = ctx->d_build_work;
  a.mean = ctx->d_mean;
  a.whitener = ctx->d_whitener;
  a.target_white = ctx->d_target_white;
  a.target_norm2 = ctx->d_target_norm2;
  a.cem_weight = ctx->d_cem_weight;
  a.status = ctx->d_status;
}

static hsi_status build_args_of(hsi_ctx* ctx, double diagonal_loading, BuildArgs* out) {
  if (ctx->moment_bands <= 0) return fail(ctx, HSI_ERR_NO_STATISTICS, "no scene statistics accumulated yet");
  const int bands = ctx->moment_bands;
  if (ctx->num_targets > 0 && ctx->lib_bands != bands)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "library has " + std::to_string(ctx->lib_bands) + " bands but the statistics cover " +
                    std::to_string(bands));
  if (!(diagonal_loading >= 0.0) || !std::isfinite(diagonal_loading))
    return fail(ctx, HSI_ERR_INVALID_ARG, "diagonal loading must be finite and non-negative");
  fill_build_args(ctx, bands, diagonal_loading, out);
  return HSI_OK;
}

// What a launched build leaves the context with.
static void mark_built(hsi_ctx* ctx, double diagonal_loading) {
  ctx->layouts_dirty = true;
  ctx->mixed_dirty = true;
  ctx->stats_bands = ctx->moment_bands;
  ctx->stats_targets = ctx->num_targets;
  ctx->last_loading = diagonal_loading;
}

extern "C" {

hsi_status hsi_stats_build(hsi_ctx* ctx, double diagonal_loading, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  BuildArgs a;
  const hsi_status st = build_args_of(ctx, diagonal_loading, &a);
  if (st != HSI_OK) return st;
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, launch_build(a, static_cast<hipStream_t>(stream)), "build statistics");
  mark_built(ctx, diagonal_loading);
  return HSI_OK;
}

hsi_status hsi_stats_check(hsi_ctx* ctx, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (ctx->stats_bands <= 0) return fail(ctx, HSI_ERR_NO_STATISTICS, "no statistics built");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, hipStreamSynchronize(static_cast<hipStream_t>(stream)), "wait for statistics");
  unsigned words[2] = {0u, 0u};
  HSI_HIP(ctx, hipMemcpy(words, ctx->d_status, sizeof(words), hipMemcpyDeviceToHost),
          "read statistics status");
  const unsigned status = words[0] | words[1];
  if (status == 0) return HSI_OK;
  char loading[64];
  std::snprintf(loading, sizeof loading, "%f", ctx->last_loading);
  // The reference's messages (detector.cpp:72-77, 97-100), then the
  // device's own reason when its stricter pivot rule fired.
  const unsigned word = (words[0] & HSI_STATUS_COV_NOT_PD) ? words[0] : words[1];
  const std::string tail = (word & HSI_STATUS_BELOW_TOLERANCE)
                               ? " (a Cholesky pivot was positive but below 4*256*eps of its diagonal: "
                                 "numerically singular)"
                               : "";
  if (status & HSI_STATUS_COV_NOT_PD)
    return fail(ctx, HSI_ERR_NOT_POSITIVE_DEFINITE,
                std::string("scene covariance is not positive definite at loading ") + loading +
                    "; hyperspectral bands are strongly correlated, so try a larger --loading" + tail);
  return fail(ctx, HSI_ERR_NOT_POSITIVE_DEFINITE,
              std::string("scene autocorrelation is not positive definite at loading ") + loading +
                  "; try a larger --loading" + tail);
}

hsi_status hsi_stats_status_copy(hsi_ctx* ctx, uint32_t* dst, hsi_stream stream) {
  if (!ctx || !dst) return fail(ctx, HSI_ERR_INVALID_ARG, "status copy needs a destination");
  if (ctx->stats_bands <= 0) return fail(ctx, HSI_ERR_NO_STATISTICS, "no statistics built");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, hipMemcpyAsync(dst, ctx->d_status, 2 * sizeof(unsigned), hipMemcpyDefault,
                              static_cast<hipStream_t>(stream)),
          "copy statistics status");
  return HSI_OK;
}

hsi_status hsi_stats_read(hsi_ctx* ctx, float* mean, float* whitener, float* target_white,
                          float* target_norm2, float* cem_weight, double* count, double* sum,
                          double* sum_outer) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, hipDeviceSynchronize(), "wait for device");
  const std::size_t mb = static_cast<std::size_t>(ctx->moment_bands);
  if (count || sum || sum_outer) {
    if (mb == 0) return fail(ctx, HSI_ERR_NO_STATISTICS, "no moments accumulated");
    if (count) HSI_HIP(ctx, hipMemcpy(count, ctx->d_count, sizeof(double), hipMemcpyDeviceToHost), "read count");
    if (sum) HSI_HIP(ctx, hipMemcpy(sum, ctx->d_sum, mb * sizeof(double), hipMemcpyDeviceToHost), "read sum");
    if (sum_outer)
      HSI_HIP(ctx, hipMemcpy(sum_outer, ctx->d_sum_outer, mb * mb * sizeof(double), hipMemcpyDeviceToHost),
              "read sum_outer");
  }
  const std::size_t b = static_cast<std::size_t>(ctx->stats_bands);
  const std::size_t t = static_cast<std::size_t>(ctx->stats_targets);
  if (mean || whitener || target_white || target_norm2 || cem_weight) {
    if (b == 0) return fail(ctx, HSI_ERR_NO_STATISTICS, "no statistics built");
    if (mean) HSI_HIP(ctx, hipMemcpy(mean, ctx->d_mean, b * sizeof(float), hipMemcpyDeviceToHost), "read mean");
    if (whitener)
      HSI_HIP(ctx, hipMemcpy(whitener, ctx->d_whitener, b * b * sizeof(float), hipMemcpyDeviceToHost),
              "read whitener");
    if (t > 0) {
      if (target_white)
        HSI_HIP(ctx, hipMemcpy(target_white, ctx->d_target_white, t * b * sizeof(float), hipMemcpyDeviceToHost),
                "read target_white");
      if (target_norm2)
        HSI_HIP(ctx, hipMemcpy(target_norm2, ctx->d_target_norm2, t * sizeof(float), hipMemcpyDeviceToHost),
                "read target_norm2");
      if (cem_weight)
        HSI_HIP(ctx, hipMemcpy(cem_weight, ctx->d_cem_weight, t * b * sizeof(float), hipMemcpyDeviceToHost),
                "read cem_weight");
    }
  }
  return HSI_OK;
}

hsi_status hsi_stats_upload(hsi_ctx* ctx, int bands, const float* mean, const float* whitener,
                            const float* target_white, const float* target_norm2,
                            const float* cem_weight) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (bands <= 0 || bands > kMaxBands || !mean || !whitener)
    return fail(ctx, HSI_ERR_INVALID_ARG, "upload needs bands, mean and whitener");
  int targets = 0;
  if (ctx->num_targets > 0 && ctx->lib_bands == bands && target_white && target_norm2 && cem_weight)
    targets = ctx->num_targets;
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const std::size_t b = static_cast<std::size_t>(bands);
  const std::size_t t = static_cast<std::size_t>(targets);
  HSI_HIP(ctx, hipMemcpy(ctx->d_mean, mean, b * sizeof(float), hipMemcpyHostToDevice), "upload mean");
  HSI_HIP(ctx, hipMemcpy(ctx->d_whitener, whitener, b * b * sizeof(float), hipMemcpyHostToDevice),
          "upload whitener");
  if (t > 0) {
    HSI_HIP(ctx, hipMemcpy(ctx->d_target_white, target_white, t * b * sizeof(float), hipMemcpyHostToDevice),
            "upload target_white");
    HSI_HIP(ctx, hipMemcpy(ctx->d_target_norm2, target_norm2, t * sizeof(float), hipMemcpyHostToDevice),
            "upload target_norm2");
    HSI_HIP(ctx, hipMemcpy(ctx->d_cem_weight, cem_weight, t * b * sizeof(float), hipMemcpyHostToDevice),
            "upload cem_weight");
  }
  HSI_HIP(ctx, hipMemset(ctx->d_status, 0, 2 * sizeof(unsigned)), "reset status");
  HSI_HIP(ctx, hipDeviceSynchronize(), "upload statistics");
  ctx->layouts_dirty = true;
  ctx->mixed_dirty = true;
  ctx->stats_bands = bands;
  ctx->stats_targets = targets;
  return HSI_OK;
}

hsi_status hsi_score_sam_cem(hsi_ctx* ctx, const void* d_cube, hsi_dtype dtype, int height, int width,
                             int bands, size_t plane_stride, float* d_sam_score, int32_t* d_sam_target,
                             float* d_cem_score, int32_t* d_cem_target, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, dtype);
  if (st != HSI_OK) return st;
  if (!d_sam_score && !d_cem_score) return HSI_OK;
  if (ctx->num_targets <= 0) return fail(ctx, HSI_ERR_NO_LIBRARY, "empty spectral library");
  if (ctx->lib_bands != bands)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "library has " + std::to_string(ctx->lib_bands) + " bands but the cube has " +
                    std::to_string(bands));
  for (const void* m : {static_cast<const void*>(d_sam_score), static_cast<const void*>(d_sam_target),
                        static_cast<const void*>(d_cem_score), static_cast<const void*>(d_cem_target)})
    if (reinterpret_cast<std::uintptr_t>(m) % 16 != 0)
      return fail(ctx, HSI_ERR_INVALID_ARG, "SAM/CEM score and target maps must be 16-byte aligned");
  if (d_cem_score && (ctx->stats_bands != bands || ctx->stats_targets != ctx->num_targets))
    return fail(ctx, HSI_ERR_NO_STATISTICS, "CEM needs statistics built for this library and band count");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  if (d_cem_score) {
    const hsi_status lst = ensure_layouts(ctx, static_cast<hipStream_t>(stream));
    if (lst != HSI_OK) return lst;
  }
  SamCemArgs a;
  a.cube = d_cube;
  a.dtype = dtype;
  a.pixels = static_cast<std::size_t>(height) * width;
  a.bands = bands;
  a.stride = plane_stride;
  a.num_targets = ctx->num_targets;
  a.lib_stride = ctx->lib_stride;
  a.sam_lib = ctx->d_sam_lib;
  a.sam_norms = ctx->d_norms;
  a.sam_rnorms = ctx->d_sam_rnorm;
  a.sam_exact_all = ctx->lib_degenerate;
  a.cem_lib = ctx->d_cem_lib;
  a.sam_score = d_sam_score;
  a.sam_target = d_sam_target;
  a.cem_score = d_cem_score;
  a.cem_target = d_cem_target;
  HSI_HIP(ctx, launch_sam_cem(a, static_cast<hipStream_t>(stream)), "score SAM/CEM");
  return HSI_OK;
}

hsi_status hsi_score(hsi_ctx* ctx, hsi_detector detector, hsi_precision precision, const void* d_cube,
                     hsi_dtype dtype, int height, int width, int bands, size_t plane_stride,
                     float* d_score, int32_t* d_target, float* d_aux, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (!d_score) return fail(ctx, HSI_ERR_INVALID_ARG, "null score map");
  if (precision != HSI_PLAIN && precision != HSI_MIXED && precision != HSI_FP32)
    return fail(ctx, HSI_ERR_INVALID_ARG, "unknown precision");
  switch (detector) {
    case HSI_DET_SAM:
    case HSI_DET_CEM:
      if (precision != HSI_PLAIN)
        return fail(ctx, HSI_ERR_INVALID_ARG, "the mixed and FP32 forms apply to RX and ACE only");
      return detector == HSI_DET_SAM
                 ? hsi_score_sam_cem(ctx, d_cube, dtype, height, width, bands, plane_stride, d_score,
                                     d_target, nullptr, nullptr, stream)
                 : hsi_score_sam_cem(ctx, d_cube, dtype, height, width, bands, plane_stride, nullptr,
                                     nullptr, d_score, d_target, stream);
    case HSI_DET_RX:
    case HSI_DET_ACE:
      break;
    default:
      return fail(ctx, HSI_ERR_INVALID_ARG, "unknown detector");
  }
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, dtype, precision == HSI_MIXED);
  if (st != HSI_OK) return st;
  if (dtype == HSI_Q16 && ctx->q16_bands != bands)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "no Q16 code for " + std::to_string(bands) + " bands: encode with hsi_encode_q16 or install one "
                "with hsi_set_q16_code");
  if (bands > kMaxRxBands)
    return fail(ctx, HSI_ERR_CAPACITY, "RX/ACE support at most " + std::to_string(kMaxRxBands) +
                                           " bands (the whitener tile must fit in LDS)");
  if (ctx->stats_bands != bands)
    return fail(ctx, HSI_ERR_NO_STATISTICS, "RX/ACE need statistics built for this band count");
  if (detector == HSI_DET_ACE && ctx->stats_targets <= 0)
    return fail(ctx, HSI_ERR_NO_LIBRARY, "ACE needs a target library");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  {
    const hipStream_t ls = static_cast<hipStream_t>(stream);
    const hsi_status lst = precision == HSI_MIXED ? ensure_mixed_layouts(ctx, ls) : ensure_layouts(ctx, ls);
    if (lst != HSI_OK) return lst;
  }
  RxAceArgs a;
  a.cube = d_cube;
  a.dtype = dtype;
  a.pixels = static_cast<std::size_t>(height) * width;
  a.bands = bands;
  a.stride = plane_stride;
  a.num_targets = detector == HSI_DET_ACE ? ctx->stats_targets : 0;
  a.mean = ctx->d_mean;
  a.white_blocks = ctx->d_white_blocks;
  a.ace_u = ctx->d_ace_u;
  a.target_norm2 = ctx->d_target_norm2;
  a.white_hl = ctx->d_white_hl;
  a.ace_hl = ctx->d_ace_hl;
  a.col_scale = ctx->d_col_scale;
  a.row_inv = ctx->d_row_inv;
  a.target_inv = ctx->d_target_inv;
  a.q_offset = ctx->d_q16_offset;
  a.q_scale = ctx->d_q16_scale;
  if (detector == HSI_DET_RX) {
    a.rx = d_score;
  } else {
    a.rx = d_aux;
    a.ace_score = d_score;
    a.ace_target = d_target;
  }
  hipStream_t s = static_cast<hipStream_t>(stream);
  const hipError_t err = precision == HSI_PLAIN  ? launch_rxace_plain(a, s)
                         : precision == HSI_FP32 ? launch_rxace_fp32(a, s)
                                                 : launch_rxace_mixed(a, s);
  HSI_HIP(ctx, err, "score RX/ACE");
  return HSI_OK;
}

hsi_status hsi_detect(hsi_ctx* ctx, const float* d_score, const int32_t* d_target, int height, int width,
                      hsi_detection_params params, hsi_detection* d_out, uint32_t* d_count,
                      hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (!d_score || !d_out || !d_count || height <= 0 || width <= 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "detect: null buffer or bad shape");
  if (params.max_detections < 0 || params.nms_radius < 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "detect: negative capacity or radius");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  HSI_HIP(ctx, grow(reinterpret_cast<void**>(&ctx->d_det_scratch), &ctx->det_scratch_bytes,
                    detect_scratch_words(pixels) * sizeof(unsigned)),
          "allocate detect scratch");
  HSI_HIP(ctx, launch_detect(d_score, d_target, height, width, params, d_out, d_count,
                             ctx->d_det_scratch, static_cast<hipStream_t>(stream)),
          "detect");
  return HSI_OK;
}

hsi_status hsi_detect_no_alloc(hsi_ctx* ctx, const float* d_score, const int32_t* d_target, int height,
                               int width, hsi_detection_params params, hsi_detection* d_out, uint32_t* d_count,
                               hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (!d_score || !d_out || !d_count || height <= 0 || width <= 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "detect: null buffer or bad shape");
  if (params.max_detections < 0 || params.nms_radius < 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "detect: negative capacity or radius");
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  if (pixels > kDetectNoAllocMaxPixels)
    return fail(ctx, HSI_ERR_CAPACITY,
                "detect without allocation: " + std::to_string(pixels) + " pixels exceed the static scratch's " +
                    std::to_string(kDetectNoAllocMaxPixels) + " (use hsi_detect)");
  unsigned* scratch = no_alloc_scratch(ctx->device);
  if (!scratch)
    return fail(ctx, HSI_ERR_DEVICE, "detect without allocation: the static scratch has no address on this device");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, launch_detect(d_score, d_target, height, width, params, d_out, d_count, scratch,
                             static_cast<hipStream_t>(stream)),
          "detect");
  return HSI_OK;
}

hsi_status hsi_narrow_to_half(const float* d_src, void* d_dst, size_t count, hsi_stream stream) {
  if (!d_src || !d_dst) return HSI_ERR_INVALID_ARG;
  // The kernel moves eight elements per lane with 16-byte accesses.
  if (reinterpret_cast<std::uintptr_t>(d_src) % 16 != 0 || reinterpret_cast<std::uintptr_t>(d_dst) % 16 != 0)
    return HSI_ERR_INVALID_ARG;
  if (count == 0) return HSI_OK;
  if (launch_narrow_to_half(d_src, static_cast<_Float16*>(d_dst), count, static_cast<hipStream_t>(stream)) !=
      hipSuccess)
    return HSI_ERR_DEVICE;
  return HSI_OK;
}

hsi_status hsi_transpose_bip_to_bsq(const float* d_src, float* d_dst, int height, int width, int bands,
                                    size_t plane_stride, hsi_stream stream) {
  if (!d_src || !d_dst || height <= 0 || width <= 0 || bands <= 0) return HSI_ERR_INVALID_ARG;
  if (plane_stride < static_cast<std::size_t>(height) * width) return HSI_ERR_INVALID_ARG;
  if (launch_transpose_bip_to_bsq(d_src, d_dst, height, width, bands, plane_stride,
                                  static_cast<hipStream_t>(stream)) != hipSuccess)
    return HSI_ERR_DEVICE;
  return HSI_OK;
}

hsi_status hsi_evaluate(hsi_ctx* ctx, const float* d_score, const uint8_t* d_labels, size_t n, int target_class,
                        int max_label, int higher_is_better, float threshold, hsi_eval_result* result,
                        double* class_stats, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (!d_score || !d_labels || !result || n == 0 || n > 0x7fffffffu)
    return fail(ctx, HSI_ERR_INVALID_ARG, "evaluate: need a score map, labels, a result and 1..2^31-1 pixels");
  if (max_label < 0 || max_label >= kEvalMaxClasses || target_class < 0 || target_class > max_label)
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "evaluate: need 0 <= target_class <= max_label < " + std::to_string(kEvalMaxClasses));
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, grow(&ctx->d_eval, &ctx->eval_bytes, eval_scratch_bytes(n)), "evaluation scratch");
  double out[8 + 3 * kEvalMaxClasses];
  HSI_HIP(ctx, launch_evaluate(d_score, d_labels, n, target_class, max_label, higher_is_better != 0, threshold,
                               ctx->d_eval, ctx->eval_bytes, out, static_cast<hipStream_t>(stream)),
          "evaluate");
  result->positives = static_cast<uint64_t>(out[0]);
  result->negatives = static_cast<uint64_t>(out[1]);
  result->tp = static_cast<uint64_t>(out[2]);
  result->fp = static_cast<uint64_t>(out[3]);
  result->fn = static_cast<uint64_t>(out[4]);
  result->tn = static_cast<uint64_t>(out[5]);
  const double pn = out[0] * out[1];
  result->auc = pn > 0.0 ? out[6] / (2.0 * pn) : std::numeric_limits<double>::quiet_NaN();
  if (class_stats)
    for (int c = 0; c <= max_label; ++c)
      for (int k = 0; k < 3; ++k) class_stats[3 * c + k] = out[8 + 3 * c + k];
  return HSI_OK;
}

hsi_status hsi_demosaic(hsi_ctx* ctx, const void* d_raw, hsi_raw_type raw_type, int raw_height, int raw_width,
                        int period, int bands, const int32_t* band_order, float* d_cube, size_t plane_stride,
                        hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (!d_raw || !d_cube || period <= 0 || raw_height < period || raw_width < period)
    return fail(ctx, HSI_ERR_INVALID_ARG, "demosaic: image is smaller than one mosaic tile");
  if (bands <= 0 || bands > period * period || bands > kMaxBands)
    return fail(ctx, HSI_ERR_INVALID_ARG, "demosaic: bands must be in 1.." + std::to_string(period * period));
  if (raw_type != HSI_RAW_U8 && raw_type != HSI_RAW_U16 && raw_type != HSI_RAW_F32)
    return fail(ctx, HSI_ERR_INVALID_ARG, "demosaic: unknown raw sample type");
  const std::size_t pixels = static_cast<std::size_t>(raw_height / period) * (raw_width / period);
  if (plane_stride < pixels) return fail(ctx, HSI_ERR_INVALID_ARG, "demosaic: plane stride below the pixel count");
  std::vector<int> pos(static_cast<std::size_t>(bands));
  for (int b = 0; b < bands; ++b) {
    pos[static_cast<std::size_t>(b)] = band_order ? band_order[b] : b;
    if (pos[static_cast<std::size_t>(b)] < 0 || pos[static_cast<std::size_t>(b)] >= period * period)
      return fail(ctx, HSI_ERR_INVALID_ARG, "demosaic: band_order entries must be in 0..period^2-1");
  }
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  if (!ctx->d_positions) HSI_HIP(ctx, alloc(&ctx->d_positions, kMaxBands), "allocate band positions");
  if (pos != ctx->positions) {
    HSI_HIP(ctx, hipMemcpy(ctx->d_positions, pos.data(), pos.size() * sizeof(int), hipMemcpyHostToDevice),
            "upload band positions");
    ctx->positions = pos;
  }
  HSI_HIP(ctx, launch_demosaic(d_raw, static_cast<int>(raw_type), raw_height, raw_width, period, bands,
                               ctx->d_positions, d_cube, plane_stride, static_cast<hipStream_t>(stream)),
          "demosaic");
  return HSI_OK;
}

}  // extern "C"

// ---- learned pre-screen ------------------------------------------------------

namespace {

std::uint64_t splitmix(std::uint64_t& state) {
  std::uint64_t z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

hsi_status mlp_alloc(hsi_ctx* ctx) {
  if (ctx->d_mlp_w) return HSI_OK;
  const std::size_t P = mlp_param_count(kMaxBands);
  HSI_HIP(ctx, alloc(&ctx->d_mlp_w, P), "allocate pre-screen weights");
  HSI_HIP(ctx, alloc(&ctx->d_mlp_m, P), "allocate pre-screen moments");
  HSI_HIP(ctx, alloc(&ctx->d_mlp_v, P), "allocate pre-screen moments");
  HSI_HIP(ctx, alloc(&ctx->d_mlp_grad, P), "allocate pre-screen gradients");
  HSI_HIP(ctx, hipMalloc(&ctx->d_mlp_a1, mlp_a1_halves(kMaxBands) * 2 + 256), "allocate pre-screen layouts");
  HSI_HIP(ctx, hipMalloc(&ctx->d_mlp_a2, mlp_a2_halves() * 2 + 256), "allocate pre-screen layouts");
  HSI_HIP(ctx, alloc(&ctx->d_mlp_tail, mlp_tail_floats()), "allocate pre-screen layouts");
  HSI_HIP(ctx, hipMalloc(&ctx->d_mlp_a1h, mlp_a1_halves(kMaxBands) * 2 + 256), "allocate pre-screen layouts");
  HSI_HIP(ctx, hipMalloc(&ctx->d_mlp_a2h, mlp_a2_halves() * 2 + 256), "allocate pre-screen layouts");
  return HSI_OK;
}

hsi_status mlp_install(hsi_ctx* ctx, int bands, const std::vector<float>& w) {
  hsi_status st = mlp_alloc(ctx);
  if (st != HSI_OK) return st;
  const std::size_t P = w.size();
  HSI_HIP(ctx, hipMemcpy(ctx->d_mlp_w, w.data(), P * sizeof(float), hipMemcpyHostToDevice), "upload weights");
  HSI_HIP(ctx, hipMemset(ctx->d_mlp_m, 0, P * sizeof(float)), "reset moments");
  HSI_HIP(ctx, hipMemset(ctx->d_mlp_v, 0, P * sizeof(float)), "reset moments");
  HSI_HIP(ctx, hipDeviceSynchronize(), "install weights");
  ctx->mlp_bands = bands;
  ctx->mlp_step = 0;
  ctx->mlp_dirty = true;
  return HSI_OK;
}

hsi_status mlp_ready(hsi_ctx* ctx, int bands) {
  if (ctx->mlp_bands <= 0) return fail(ctx, HSI_ERR_INVALID_ARG, "no pre-screen network: call hsi_mlp_init or hsi_mlp_set_weights");
  if (ctx->mlp_bands != bands)
    return fail(ctx, HSI_ERR_INVALID_ARG, "pre-screen network takes " + std::to_string(ctx->mlp_bands) +
                                              " bands, input has " + std::to_string(bands));
  return HSI_OK;
}

hsi_status mlp_check_samples(hsi_ctx* ctx, const float* d_x, std::size_t x_stride, int n, const float* d_labels) {
  if (!d_x || !d_labels || n <= 0 || x_stride < static_cast<std::size_t>(n))
    return fail(ctx, HSI_ERR_INVALID_ARG, "training batch needs samples, labels, n > 0 and x_stride >= n");
  return HSI_OK;
}

}  // namespace

extern "C" {

size_t hsi_mlp_param_count(int bands) {
  return bands > 0 && bands <= kMaxBands ? mlp_param_count(bands) : 0;
}

hsi_status hsi_mlp_init(hsi_ctx* ctx, int bands, uint64_t seed) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (bands <= 0 || bands > kMaxBands) return fail(ctx, HSI_ERR_INVALID_ARG, "pre-screen bands out of range");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  std::vector<float> w(mlp_param_count(bands), 0.f);
  std::uint64_t state = seed;
  auto uniform = [&](float limit) {
    const double u = static_cast<double>(splitmix(state) >> 11) * (1.0 / 9007199254740992.0);
    return static_cast<float>((2.0 * u - 1.0) * limit);
  };
  // He-uniform for the ReLU layers, Glorot-uniform for the output; zero biases.
  const std::size_t o_b1 = 64u * bands, o_w2 = o_b1 + 64, o_b2 = o_w2 + 32 * 64, o_w3 = o_b2 + 32;
  for (std::size_t i = 0; i < o_b1; ++i) w[i] = uniform(std::sqrt(6.0f / bands));
  for (std::size_t i = o_w2; i < o_b2; ++i) w[i] = uniform(std::sqrt(6.0f / 64.0f));
  for (std::size_t i = o_w3; i < o_w3 + 32; ++i) w[i] = uniform(std::sqrt(6.0f / 33.0f));
  return mlp_install(ctx, bands, w);
}

hsi_status hsi_mlp_set_weights(hsi_ctx* ctx, int bands, const float* weights) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (bands <= 0 || bands > kMaxBands || !weights)
    return fail(ctx, HSI_ERR_INVALID_ARG, "pre-screen weights need bands in range and a buffer");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  std::vector<float> w(weights, weights + mlp_param_count(bands));
  return mlp_install(ctx, bands, w);
}

hsi_status hsi_mlp_get_weights(hsi_ctx* ctx, float* weights) {
  if (!ctx || !weights) return fail(ctx, HSI_ERR_INVALID_ARG, "null buffer");
  if (ctx->mlp_bands <= 0) return fail(ctx, HSI_ERR_INVALID_ARG, "no pre-screen network");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, hipDeviceSynchronize(), "wait for device");
  HSI_HIP(ctx, hipMemcpy(weights, ctx->d_mlp_w, mlp_param_count(ctx->mlp_bands) * sizeof(float),
                         hipMemcpyDeviceToHost),
          "read weights");
  return HSI_OK;
}

static hsi_status mlp_grad_common(hsi_ctx* ctx, const float* d_x, size_t x_stride, int n, const float* d_labels,
                                  float pos_weight, float* d_grad, float* d_loss, hipStream_t s) {
  HSI_HIP(ctx, grow(reinterpret_cast<void**>(&ctx->d_mlp_act), &ctx->mlp_act_bytes,
                    mlp_act_floats(n) * sizeof(float)),
          "allocate training scratch");
  HSI_HIP(ctx, launch_mlp_gradients(d_x, x_stride, n, ctx->mlp_bands, d_labels, pos_weight, ctx->d_mlp_w,
                                    ctx->d_mlp_act, d_grad, d_loss, s),
          "pre-screen gradients");
  return HSI_OK;
}

hsi_status hsi_mlp_gradients(hsi_ctx* ctx, const float* d_x, size_t x_stride, int n, const float* d_labels,
                             float pos_weight, float* d_grad, float* d_loss, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = mlp_ready(ctx, ctx->mlp_bands);
  if (st != HSI_OK) return st;
  if ((st = mlp_check_samples(ctx, d_x, x_stride, n, d_labels)) != HSI_OK) return st;
  if (!d_grad) return fail(ctx, HSI_ERR_INVALID_ARG, "null gradient buffer");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  return mlp_grad_common(ctx, d_x, x_stride, n, d_labels, pos_weight, d_grad, d_loss,
                         static_cast<hipStream_t>(stream));
}

hsi_status hsi_mlp_train_step(hsi_ctx* ctx, const float* d_x, size_t x_stride, int n, const float* d_labels,
                              const hsi_mlp_train_params* params, float* d_loss, hsi_stream stream) {
  if (!ctx || !params) return HSI_ERR_INVALID_ARG;
  hsi_status st = mlp_ready(ctx, ctx->mlp_bands);
  if (st != HSI_OK) return st;
  if ((st = mlp_check_samples(ctx, d_x, x_stride, n, d_labels)) != HSI_OK) return st;
  if (!(params->learning_rate > 0.f) || !(params->beta1 >= 0.f && params->beta1 < 1.f) ||
      !(params->beta2 >= 0.f && params->beta2 < 1.f) || !(params->epsilon > 0.f))
    return fail(ctx, HSI_ERR_INVALID_ARG, "Adam needs lr > 0, betas in [0, 1), epsilon > 0");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if ((st = mlp_grad_common(ctx, d_x, x_stride, n, d_labels, params->pos_weight, ctx->d_mlp_grad, d_loss, s)) != HSI_OK)
    return st;
  ++ctx->mlp_step;
  const double t = static_cast<double>(ctx->mlp_step);
  const float c1 = static_cast<float>(1.0 / (1.0 - std::pow(static_cast<double>(params->beta1), t)));
  const float c2 = static_cast<float>(1.0 / (1.0 - std::pow(static_cast<double>(params->beta2), t)));
  HSI_HIP(ctx, launch_adam(ctx->d_mlp_w, ctx->d_mlp_m, ctx->d_mlp_v, ctx->d_mlp_grad,
                           mlp_param_count(ctx->mlp_bands), params->learning_rate, params->beta1, params->beta2,
                           params->epsilon, c1, c2, s),
          "Adam update");
  ctx->mlp_dirty = true;
  return HSI_OK;
}

hsi_status hsi_mlp_infer(hsi_ctx* ctx, const void* d_cube, hsi_dtype dtype, int height, int width, int bands,
                         size_t plane_stride, int center_on_statistics_mean, float* d_logit, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, dtype);
  if (st != HSI_OK) return st;
  if ((st = mlp_ready(ctx, bands)) != HSI_OK) return st;
  if (!d_logit) return fail(ctx, HSI_ERR_INVALID_ARG, "null logit map");
  if (center_on_statistics_mean && ctx->stats_bands != bands)
    return fail(ctx, HSI_ERR_NO_STATISTICS, "centring on the scene mean needs statistics for this band count");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (ctx->mlp_dirty) {
    HSI_HIP(ctx, launch_mlp_layouts(ctx->d_mlp_w, bands, ctx->d_mlp_a1, ctx->d_mlp_a2, ctx->d_mlp_tail,
                                      ctx->d_mlp_a1h, ctx->d_mlp_a2h, s),
            "pre-screen layouts");
    ctx->mlp_dirty = false;
  }
  MlpInferArgs a;
  a.cube = d_cube;
  a.dtype = dtype;
  a.pixels = static_cast<std::size_t>(height) * width;
  a.bands = bands;
  a.stride = plane_stride;
  a.center = center_on_statistics_mean ? ctx->d_mean : nullptr;
  a.a1 = ctx->d_mlp_a1;
  a.a2 = ctx->d_mlp_a2;
  a.tail = ctx->d_mlp_tail;
  a.logit = d_logit;
  HSI_HIP(ctx, launch_mlp_infer(a, s), "pre-screen inference");
  return HSI_OK;
}

hsi_status hsi_gather_pixels(hsi_ctx* ctx, const float* d_cube, int height, int width, int bands,
                             size_t plane_stride, const int32_t* d_index, int n, int center_on_statistics_mean,
                             float* d_out, size_t out_stride, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, HSI_F32);
  if (st != HSI_OK) return st;
  if (!d_index || !d_out || n < 0 || out_stride < static_cast<std::size_t>(n))
    return fail(ctx, HSI_ERR_INVALID_ARG, "gather needs an index list, an output and out_stride >= n");
  if (center_on_statistics_mean && ctx->stats_bands != bands)
    return fail(ctx, HSI_ERR_NO_STATISTICS, "centring on the scene mean needs statistics for this band count");
  if (n == 0) return HSI_OK;
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  HSI_HIP(ctx, launch_gather(d_cube, plane_stride, static_cast<std::size_t>(height) * width, bands, d_index, n,
                             center_on_statistics_mean ? ctx->d_mean : nullptr,
                             d_out, out_stride, static_cast<hipStream_t>(stream)),
          "gather pixels");
  return HSI_OK;
}

}  // extern "C"

namespace {
hsi_status score_candidates(hsi_ctx* ctx, hsi_detector detector, hsi_precision precision, const float* d_cube,
                            int height, int width, int bands, size_t plane_stride,
                            const hsi_detection* d_candidates, const uint32_t* d_count, int capacity,
                            const float* d_spectra, size_t spectra_stride, float* d_score, int32_t* d_target,
                            float* d_aux, hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  if (detector != HSI_DET_RX && detector != HSI_DET_ACE)
    return fail(ctx, HSI_ERR_INVALID_ARG, "candidate rescoring applies to RX and ACE");
  if (precision != HSI_PLAIN && precision != HSI_MIXED)
    return fail(ctx, HSI_ERR_INVALID_ARG, "unknown precision");
  if (precision == HSI_MIXED && !d_spectra)
    return fail(ctx, HSI_ERR_INVALID_ARG, "the mixed form rescores candidate spectra rows only");
  if (precision == HSI_MIXED && (spectra_stride % 4 != 0 || reinterpret_cast<std::uintptr_t>(d_spectra) % 16 != 0))
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "mixed-form candidate spectra need 16-byte aligned rows (pitch a multiple of 4)");
  if (precision == HSI_MIXED && capacity > 0 &&
      static_cast<std::size_t>(capacity) * spectra_stride * sizeof(float) >= (static_cast<std::size_t>(1) << 31))
    return fail(ctx, HSI_ERR_CAPACITY, "mixed-form candidate spectra must span less than 2 GiB");
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, HSI_F32);
  if (st != HSI_OK) return st;
  if (!d_candidates || !d_count || capacity < 0 || !d_score)
    return fail(ctx, HSI_ERR_INVALID_ARG, "candidate rescoring needs a list, its count and a score map");
  if (bands > kMaxRxBands)
    return fail(ctx, HSI_ERR_CAPACITY, "RX/ACE support at most " + std::to_string(kMaxRxBands) + " bands");
  if (ctx->stats_bands != bands)
    return fail(ctx, HSI_ERR_NO_STATISTICS, "RX/ACE need statistics built for this band count");
  if (detector == HSI_DET_ACE && ctx->stats_targets <= 0)
    return fail(ctx, HSI_ERR_NO_LIBRARY, "ACE needs a target library");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if ((st = precision == HSI_PLAIN ? ensure_layouts(ctx, s) : ensure_mixed_layouts(ctx, s)) != HSI_OK) return st;
  const std::size_t pixels = static_cast<std::size_t>(height) * width;
  HSI_HIP(ctx, launch_fill_scores(d_score, detector == HSI_DET_ACE ? d_target : nullptr,
                                  detector == HSI_DET_ACE ? d_aux : nullptr, pixels, s),
          "clear score maps");
  if (capacity == 0) return HSI_OK;
  RxAceArgs a;
  a.cube = d_cube;
  a.dtype = HSI_F32;
  a.pixels = pixels;
  a.bands = bands;
  a.stride = plane_stride;
  a.num_targets = detector == HSI_DET_ACE ? ctx->stats_targets : 0;
  a.mean = ctx->d_mean;
  a.white_blocks = ctx->d_white_blocks;
  a.ace_u = ctx->d_ace_u;
  a.target_norm2 = ctx->d_target_norm2;
  a.white_hl = ctx->d_white_hl;
  a.ace_hl = ctx->d_ace_hl;
  a.col_scale = ctx->d_col_scale;
  a.row_inv = ctx->d_row_inv;
  a.target_inv = ctx->d_target_inv;
  if (detector == HSI_DET_RX) {
    a.rx = d_score;
  } else {
    a.rx = d_aux;
    a.ace_score = d_score;
    a.ace_target = d_target;
  }
  const unsigned cap = static_cast<unsigned>(capacity);
  const hipError_t err =
      precision == HSI_PLAIN
          ? launch_rxace_candidates(a, d_candidates, d_count, cap, width, d_spectra, spectra_stride, s)
          : launch_rxace_mixed_candidates(a, d_candidates, d_count, cap, width, d_spectra, spectra_stride, s);
  HSI_HIP(ctx, err, "rescore candidates");
  return HSI_OK;
}
}  // namespace

extern "C" {

hsi_status hsi_score_candidates(hsi_ctx* ctx, hsi_detector detector, const float* d_cube, int height, int width,
                                int bands, size_t plane_stride, const hsi_detection* d_candidates,
                                const uint32_t* d_count, int capacity, float* d_score, int32_t* d_target,
                                float* d_aux, hsi_stream stream) {
  return score_candidates(ctx, detector, HSI_PLAIN, d_cube, height, width, bands, plane_stride, d_candidates,
                          d_count, capacity, nullptr, 0, d_score, d_target, d_aux, stream);
}

hsi_status hsi_score_candidate_spectra(hsi_ctx* ctx, hsi_detector detector, hsi_precision precision,
                                       const float* d_cube, int height, int width, int bands,
                                       size_t plane_stride, const hsi_detection* d_candidates,
                                       const uint32_t* d_count, int capacity, const float* d_spectra,
                                       size_t spectra_stride, float* d_score, int32_t* d_target, float* d_aux,
                                       hsi_stream stream) {
  if (!d_spectra || spectra_stride < static_cast<size_t>(bands < 0 ? 0 : bands))
    return ctx ? fail(ctx, HSI_ERR_INVALID_ARG, "candidate spectra need a buffer with pitch >= bands")
               : HSI_ERR_INVALID_ARG;
  return score_candidates(ctx, detector, precision, d_cube, height, width, bands, plane_stride, d_candidates,
                          d_count, capacity, d_spectra, spectra_stride, d_score, d_target, d_aux, stream);
}

hsi_status hsi_mlp_screen(hsi_ctx* ctx, const float* d_cube, int height, int width, int bands,
                          size_t plane_stride, float cut, hsi_detection* d_candidates, uint32_t* d_count,
                          int capacity, float* d_spectra, size_t spectra_stride, float* d_logit,
                          hsi_stream stream) {
  if (!ctx) return HSI_ERR_INVALID_ARG;
  hsi_status st = check_cube(ctx, d_cube, height, width, bands, plane_stride, HSI_F32);
  if (st != HSI_OK) return st;
  if ((st = mlp_ready(ctx, bands)) != HSI_OK) return st;
  if (!d_candidates || !d_count || capacity < 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "screening needs a candidate list, its count and a capacity >= 0");
  const std::size_t pb = static_cast<std::size_t>((bands + 31) / 32 * 32);
  if (d_spectra && (spectra_stride < pb || spectra_stride % 4 != 0))
    return fail(ctx, HSI_ERR_INVALID_ARG,
                "candidate spectra need a pitch >= 32 * ceil(bands / 32), a multiple of 4");
  // The pass writes each row with 16-byte stores.
  if (d_spectra && reinterpret_cast<std::uintptr_t>(d_spectra) % 16 != 0)
    return fail(ctx, HSI_ERR_INVALID_ARG, "candidate spectra must be 16-byte aligned");
  if (ctx->stats_bands != bands)
    return fail(ctx, HSI_ERR_NO_STATISTICS, "screening centres on the scene mean: build statistics first");
  if (set_device(ctx) != HSI_OK) return HSI_ERR_DEVICE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (ctx->mlp_dirty) {
    HSI_HIP(ctx, launch_mlp_layouts(ctx->d_mlp_w, bands, ctx->d_mlp_a1, ctx->d_mlp_a2, ctx->d_mlp_tail,
                                      ctx->d_mlp_a1h, ctx->d_mlp_a2h, s),
            "pre-screen layouts");
    ctx->mlp_dirty = false;
  }
  MlpInferArgs a;
  a.cube = d_cube;
  a.dtype = HSI_F32;
  a.pixels = static_cast<std::size_t>(height) * width;
  a.bands = bands;
  a.stride = plane_stride;
  a.center = ctx->d_mean;
  a.a1 = ctx->d_mlp_a1;
  a.a2 = ctx->d_mlp_a2;
  a.tail = ctx->d_mlp_tail;
  a.logit = d_logit;
  a.cut = cut;
  a.cand = d_candidates;
  a.cand_count = d_count;
  a.cand_cap = static_cast<unsigned>(capacity);
  a.spectra = d_spectra;
  a.spectra_stride = spectra_stride;
  a.width = width;
  const std::size_t cap = static_cast<std::size_t>(capacity);
  HSI_HIP(ctx, grow(&ctx->d_screen_rec, &ctx->screen_rec_bytes, (cap ? cap : 1) * sizeof(hsi_detection)),
          "screening scratch");
  HSI_HIP(ctx, grow(&ctx->d_screen_counts, &ctx->screen_counts_bytes, 2 * kScreenMaxWaves * sizeof(unsigned)),
          "screening scratch");
  a.scratch_rec = static_cast<hsi_detection*>(ctx->d_screen_rec);
  a.scratch_counts = static_cast<unsigned*>(ctx->d_screen_counts);
  HSI_HIP(ctx, launch_mlp_infer(a, s), "pre-screen screening");
  return HSI_OK;
}

}  // extern "C"

// ---- batched scenes -------------------------------------------------------------
//
// Scenes of different sizes, band counts and library sizes scored together:
// one hsi_ctx per scene holds its library and statistics, and the per-pixel
// scoring runs as one persistent launch per kernel shape class - SAM/CEM
// grouped by library stride (1..32 targets -> 1, 2, 4, 8, 16, 32), RX/ACE by
// 16-band row blocks and target tiles - over device tables of the scenes.
// Launch shapes (pixels per thread, grid size, tiles per work item) come from
// an autotuner that times the candidates on the batch itself.

struct hsi_batch {
  int device = 0;
  std::string error;
  std::vector<hsi_ctx*> ctx;
  std::vector<hsi_scene> scenes;
  struct SamClass {
    int lib_stride = 1;
    std::vector<SamSceneDev> table;
    unsigned blocks[2] = {0, 0};  // total 256*PPT-pixel blocks for PPT = 4, 2
    SamSceneDev* d_table[2] = {nullptr, nullptr};
    int* d_block_scene[2] = {nullptr, nullptr};
    int ppt = 4;
    int workgroups = 0;           // 0: one workgroup per block
  };
  struct RxClass {
    int nb = 0, tt = 0;
    std::vector<RxSceneDev> table;
    RxSceneDev* d_table = nullptr;
    std::vector<RxWorkItem> items[3];  // tiles per item: 16, 64, 256
    RxWorkItem* d_items[3] = {nullptr, nullptr, nullptr};
    int chunk = 1;                     // index into items
    int workgroups = 0;                // 0: one workgroup per item
  };
  std::vector<SamClass> sam;
  std::vector<RxClass> rx;
  bool tuned = false;
  // Batched statistics: the scenes that need them and their build arguments
  // in two device tables, one per build form (matrices in LDS / in global
  // memory), each built by one launch.
  std::vector<int> stats_scenes;
  BuildArgs* d_build[2] = {nullptr, nullptr};
  int build_count[2] = {0, 0};
  // Statistics + scoring captured once as a graph (hsi_batch_run), on a
  // stream of the batch's own, ordered against the caller's by events.
  hipStream_t graph_stream = nullptr;
  hipEvent_t graph_in = nullptr, graph_out = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double graph_loading = -1.0;
};

namespace {

hsi_status bfail(hsi_batch* b, hsi_status st, const std::string& msg) {
  if (b) b->error = msg;
  return st;
}

const int kRxChunks[3] = {16, 64, 256};

void batch_release_graph(hsi_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec), b->exec = nullptr;
  if (b->graph) (void)hipGraphDestroy(b->graph), b->graph = nullptr;
  b->graph_loading = -1.0;
}

void batch_release_tables(hsi_batch* b) {
  batch_release_graph(b);
  for (auto*& p : b->d_build)
    if (p) (void)hipFree(p), p = nullptr;
  b->build_count[0] = b->build_count[1] = 0;
  b->stats_scenes.clear();
  for (auto& c : b->sam) {
    for (auto*& p : c.d_table)
      if (p) (void)hipFree(p), p = nullptr;
    for (auto*& p : c.d_block_scene)
      if (p) (void)hipFree(p), p = nullptr;
  }
  for (auto& c : b->rx) {
    if (c.d_table) (void)hipFree(c.d_table), c.d_table = nullptr;
    for (auto*& p : c.d_items)
      if (p) (void)hipFree(p), p = nullptr;
  }
  b->sam.clear();
  b->rx.clear();
}

template <typename T>
hipError_t upload_vec(T** dst, const std::vector<T>& v) {
  if (v.empty()) return hipSuccess;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(dst), v.size() * sizeof(T));
  if (e != hipSuccess) return e;
  return hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

// wgs > 0: persistent grid; 0: one workgroup per block or item; -1: one
// launch of the single-scene kernel per scene (the tuner keeps whichever
// shape is fastest for the class).
constexpr int kPerScene = -1;

hipError_t run_sam_class(const hsi_batch::SamClass& c, int ppt, int wgs, hipStream_t s) {
  if (wgs == kPerScene) {
    for (const SamSceneDev& d : c.table) {
      SamCemArgs a;
      a.cube = d.cube;
      a.dtype = HSI_F32;
      a.pixels = d.pixels;
      a.bands = d.bands;
      a.stride = d.stride;
      a.num_targets = d.num_targets;
      a.lib_stride = c.lib_stride;
      a.sam_lib = d.sam_lib;
      a.sam_norms = d.sam_norms;
      a.sam_rnorms = d.sam_rnorms;
      a.sam_exact_all = d.exact_all;
      a.cem_lib = d.cem_lib;
      a.sam_score = d.sam_score;
      a.sam_target = d.sam_target;
      a.cem_score = d.cem_score;
      a.cem_target = d.cem_target;
      const hipError_t e = launch_sam_cem(a, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  const int k = ppt == 4 ? 0 : 1;
  return launch_sam_cem_batch(c.lib_stride, ppt, c.d_table[k], c.d_block_scene[k], c.blocks[k], wgs, s);
}

hipError_t run_rx_class(const hsi_batch::RxClass& c, int chunk, int wgs, hipStream_t s) {
  if (wgs == kPerScene) {
    for (const RxSceneDev& d : c.table) {
      RxAceArgs a;
      a.cube = d.cube;
      a.dtype = HSI_F32;
      a.pixels = d.pixels;
      a.bands = d.bands;
      a.stride = d.stride;
      a.num_targets = d.num_targets;
      a.mean = d.mean;
      a.white_blocks = d.white_blocks;
      a.ace_u = d.ace_u;
      a.target_norm2 = d.target_norm2;
      a.rx = d.rx;
      a.ace_score = d.ace_score;
      a.ace_target = d.ace_target;
      const hipError_t e = launch_rxace_plain(a, s);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
  return launch_rxace_batch(c.nb, c.tt, c.d_table, c.d_items[chunk], static_cast<int>(c.items[chunk].size()), wgs, s);
}

}  // namespace

extern "C" {

hsi_status hsi_batch_create(int device, hsi_batch** out) {
  if (!out) return HSI_ERR_INVALID_ARG;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) return HSI_ERR_DEVICE;
  *out = new hsi_batch();
  (*out)->device = device;
  return HSI_OK;
}

void hsi_batch_destroy(hsi_batch* b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->graph_stream) (void)hipStreamSynchronize(b->graph_stream);
  batch_release_tables(b);
  if (b->graph_in) (void)hipEventDestroy(b->graph_in);
  if (b->graph_out) (void)hipEventDestroy(b->graph_out);
  if (b->graph_stream) (void)hipStreamDestroy(b->graph_stream);
  for (hsi_ctx* c : b->ctx) hsi_ctx_destroy(c);
  delete b;
}

const char* hsi_batch_last_error(const hsi_batch* b) { return b ? b->error.c_str() : ""; }

hsi_ctx* hsi_batch_context(hsi_batch* b, int scene) {
  if (!b || scene < 0 || scene >= static_cast<int>(b->ctx.size())) return nullptr;
  return b->ctx[static_cast<std::size_t>(scene)];
}

hsi_status hsi_batch_set_scenes(hsi_batch* b, const hsi_scene* scenes, int count) {
  if (!b) return HSI_ERR_INVALID_ARG;
  if (!scenes || count <= 0) return bfail(b, HSI_ERR_INVALID_ARG, "a batch needs at least one scene");
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "select device");
  batch_release_tables(b);
  while (static_cast<int>(b->ctx.size()) < count) {
    hsi_ctx* c = nullptr;
    const hsi_status st = hsi_ctx_create(b->device, &c);
    if (st != HSI_OK) return bfail(b, st, "cannot create a scene context");
    b->ctx.push_back(c);
  }
  while (static_cast<int>(b->ctx.size()) > count) {
    hsi_ctx_destroy(b->ctx.back());
    b->ctx.pop_back();
  }
  b->scenes.assign(scenes, scenes + count);
  b->tuned = false;
  for (int i = 0; i < count; ++i) {
    const hsi_scene& sc = scenes[i];
    hsi_ctx* c = b->ctx[static_cast<std::size_t>(i)];
    hsi_status st = check_cube(c, sc.d_cube, sc.height, sc.width, sc.bands, sc.plane_stride, HSI_F32);
    if (st != HSI_OK) return bfail(b, st, "scene " + std::to_string(i) + ": " + c->error);
    if (sc.bands > kMaxRxBands && (sc.d_rx_score || sc.d_ace_score))
      return bfail(b, HSI_ERR_CAPACITY, "scene " + std::to_string(i) + ": RX/ACE support at most " +
                                            std::to_string(kMaxRxBands) + " bands");
    const bool wants_lib = sc.d_sam_score || sc.d_cem_score || sc.d_ace_score;
    if (wants_lib && (!sc.library || sc.num_targets <= 0))
      return bfail(b, HSI_ERR_NO_LIBRARY, "scene " + std::to_string(i) + ": SAM/CEM/ACE need a library");
    for (const void* m : {static_cast<const void*>(sc.d_sam_score), static_cast<const void*>(sc.d_sam_target),
                          static_cast<const void*>(sc.d_cem_score), static_cast<const void*>(sc.d_cem_target)})
      if (reinterpret_cast<std::uintptr_t>(m) % 16 != 0)
        return bfail(b, HSI_ERR_INVALID_ARG, "scene " + std::to_string(i) + ": SAM/CEM maps must be 16-byte aligned");
    if (sc.library && sc.num_targets > 0) {
      st = hsi_set_library(c, sc.library, nullptr, sc.num_targets, sc.bands);
      if (st != HSI_OK) return bfail(b, st, "scene " + std::to_string(i) + ": " + c->error);
    }
    st = hsi_ctx_reserve(c, static_cast<std::size_t>(sc.height) * sc.width, sc.bands);
    if (st != HSI_OK) return bfail(b, st, "scene " + std::to_string(i) + ": " + c->error);
  }
  // SAM/CEM classes by library stride.
  for (int i = 0; i < count; ++i) {
    const hsi_scene& sc = scenes[i];
    if (!sc.d_sam_score && !sc.d_cem_score) continue;
    hsi_ctx* c = b->ctx[static_cast<std::size_t>(i)];
    hsi_batch::SamClass* cls = nullptr;
    for (auto& k : b->sam)
      if (k.lib_stride == c->lib_stride) cls = &k;
    if (!cls) {
      b->sam.emplace_back();
      cls = &b->sam.back();
      cls->lib_stride = c->lib_stride;
      cls->ppt = c->lib_stride <= 8 ? 4 : 2;
    }
    SamSceneDev d{};
    d.cube = sc.d_cube;
    d.pixels = static_cast<std::size_t>(sc.height) * sc.width;
    d.bands = sc.bands;
    d.stride = sc.plane_stride;
    d.num_targets = c->num_targets;
    d.exact_all = c->lib_degenerate;
    d.sam_lib = c->d_sam_lib;
    d.sam_norms = c->d_norms;
    d.sam_rnorms = c->d_sam_rnorm;
    d.cem_lib = c->d_cem_lib;
    d.sam_score = sc.d_sam_score;
    d.sam_target = sc.d_sam_target;
    d.cem_score = sc.d_cem_score;
    d.cem_target = sc.d_cem_target;
    cls->table.push_back(d);
  }
  for (auto& cls : b->sam) {
    const int ppts[2] = {4, 2};
    for (int k = 0; k < 2; ++k) {
      std::vector<SamSceneDev> t = cls.table;
      std::vector<int> owner;
      unsigned total = 0;
      for (std::size_t j = 0; j < t.size(); ++j) {
        t[j].block_begin = total;
        const unsigned nblk = static_cast<unsigned>(div_up(t[j].pixels, static_cast<std::size_t>(256 * ppts[k])));
        owner.insert(owner.end(), nblk, static_cast<int>(j));
        total += nblk;
      }
      cls.blocks[k] = total;
      if (upload_vec(&cls.d_table[k], t) != hipSuccess || upload_vec(&cls.d_block_scene[k], owner) != hipSuccess)
        return bfail(b, HSI_ERR_DEVICE, "upload scene table");
    }
  }
  // RX/ACE classes by row blocks and target tiles.
  for (int i = 0; i < count; ++i) {
    const hsi_scene& sc = scenes[i];
    if (!sc.d_rx_score && !sc.d_ace_score) continue;
    hsi_ctx* c = b->ctx[static_cast<std::size_t>(i)];
    const int nb = rx_blocks(sc.bands);
    const int T = sc.d_ace_score ? c->num_targets : 0;
    const int tt = T == 0 ? 0 : (T <= 16 ? 1 : 2);
    hsi_batch::RxClass* cls = nullptr;
    for (auto& k : b->rx)
      if (k.nb == nb && k.tt == tt) cls = &k;
    if (!cls) {
      b->rx.emplace_back();
      cls = &b->rx.back();
      cls->nb = nb;
      cls->tt = tt;
    }
    RxSceneDev d{};
    d.cube = sc.d_cube;
    d.pixels = static_cast<std::size_t>(sc.height) * sc.width;
    d.bands = sc.bands;
    d.stride = sc.plane_stride;
    d.num_targets = T;
    d.mean = c->d_mean;
    d.white_blocks = c->d_white_blocks;
    d.ace_u = c->d_ace_u;
    d.target_norm2 = c->d_target_norm2;
    d.rx = sc.d_rx_score;
    d.ace_score = sc.d_ace_score;
    d.ace_target = sc.d_ace_target;
    const int scene_index = static_cast<int>(cls->table.size());
    cls->table.push_back(d);
    const std::size_t tiles = div_up(d.pixels, 16);
    for (int k = 0; k < 3; ++k)
      for (std::size_t t0 = 0; t0 < tiles; t0 += static_cast<std::size_t>(kRxChunks[k]))
        cls->items[k].push_back(RxWorkItem{scene_index, t0, std::min(tiles, t0 + static_cast<std::size_t>(kRxChunks[k]))});
  }
  for (auto& cls : b->rx) {
    if (upload_vec(&cls.d_table, cls.table) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "upload scene table");
    for (int k = 0; k < 3; ++k)
      if (upload_vec(&cls.d_items[k], cls.items[k]) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "upload work items");
  }
  // Build tables: every scene whose detectors need statistics.
  std::vector<BuildArgs> forms[2];
  for (int i = 0; i < count; ++i) {
    const hsi_scene& sc = scenes[i];
    if (!sc.d_rx_score && !sc.d_ace_score && !sc.d_cem_score) continue;
    BuildArgs a;
    fill_build_args(b->ctx[static_cast<std::size_t>(i)], sc.bands, 0.0, &a);
    forms[build_in_lds(sc.bands) ? 0 : 1].push_back(a);
    b->stats_scenes.push_back(i);
  }
  for (int k = 0; k < 2; ++k) {
    b->build_count[k] = static_cast<int>(forms[k].size());
    if (upload_vec(&b->d_build[k], forms[k]) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "upload build table");
  }
  return HSI_OK;
}

hsi_status hsi_batch_statistics(hsi_batch* b, double loading, hsi_stream stream) {
  if (!b) return HSI_ERR_INVALID_ARG;
  if (!(loading >= 0.0) || !std::isfinite(loading))
    return bfail(b, HSI_ERR_INVALID_ARG, "diagonal loading must be finite and non-negative");
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "select device");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // Moments scene by scene (each launch spans the GPU), then every build in
  // one launch per build form - the builds are one or two workgroups each.
  for (const int i : b->stats_scenes) {
    const hsi_scene& sc = b->scenes[static_cast<std::size_t>(i)];
    hsi_ctx* c = b->ctx[static_cast<std::size_t>(i)];
    const hsi_status st = hsi_stats_accumulate(c, sc.d_cube, HSI_F32, sc.height, sc.width, sc.bands,
                                               sc.plane_stride, HSI_PLAIN, nullptr, stream);
    if (st != HSI_OK) return bfail(b, st, "scene " + std::to_string(i) + ": " + c->error);
  }
  for (int k = 0; k < 2; ++k) {
    const hipError_t e = launch_build_batch(b->d_build[k], b->build_count[k], k == 0, loading, s);
    if (e != hipSuccess) return bfail(b, HSI_ERR_DEVICE, std::string("batched build: ") + hipGetErrorString(e));
  }
  for (const int i : b->stats_scenes) {
    hsi_ctx* c = b->ctx[static_cast<std::size_t>(i)];
    mark_built(c, loading);
    const hsi_status st = ensure_layouts(c, s);
    if (st != HSI_OK) return bfail(b, st, "scene " + std::to_string(i) + ": " + c->error);
  }
  return HSI_OK;
}

hsi_status hsi_batch_check(hsi_batch* b, int* failed_scene) {
  if (!b) return HSI_ERR_INVALID_ARG;
  if (failed_scene) *failed_scene = -1;
  for (std::size_t i = 0; i < b->scenes.size(); ++i) {
    hsi_ctx* c = b->ctx[i];
    if (c->stats_bands <= 0) continue;
    const hsi_status st = hsi_stats_check(c, nullptr);
    if (st != HSI_OK) {
      if (failed_scene) *failed_scene = static_cast<int>(i);
      return bfail(b, st, "scene " + std::to_string(i) + ": " + c->error);
    }
  }
  return HSI_OK;
}

hsi_status hsi_batch_score(hsi_batch* b, hsi_stream stream) {
  if (!b) return HSI_ERR_INVALID_ARG;
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "select device");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  for (std::size_t i = 0; i < b->scenes.size(); ++i) {
    const hsi_scene& sc = b->scenes[i];
    if ((sc.d_cem_score || sc.d_rx_score || sc.d_ace_score) && b->ctx[i]->stats_bands != sc.bands)
      return bfail(b, HSI_ERR_NO_STATISTICS, "scene " + std::to_string(i) + ": run hsi_batch_statistics first");
  }
  for (const auto& c : b->sam) {
    const hipError_t e = run_sam_class(c, c.ppt, c.workgroups, s);
    if (e != hipSuccess) return bfail(b, HSI_ERR_DEVICE, std::string("batched SAM/CEM: ") + hipGetErrorString(e));
  }
  for (const auto& c : b->rx) {
    const hipError_t e = run_rx_class(c, c.chunk, c.workgroups, s);
    if (e != hipSuccess) return bfail(b, HSI_ERR_DEVICE, std::string("batched RX/ACE: ") + hipGetErrorString(e));
  }
  return HSI_OK;
}

hsi_status hsi_batch_run(hsi_batch* b, double loading, int use_graph, hsi_stream stream) {
  if (!b) return HSI_ERR_INVALID_ARG;
  if (!use_graph) {
    const hsi_status st = hsi_batch_statistics(b, loading, stream);
    return st == HSI_OK ? hsi_batch_score(b, stream) : st;
  }
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "select device");
  if (!b->graph_stream) {
    if (hipStreamCreateWithFlags(&b->graph_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&b->graph_in, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->graph_out, hipEventDisableTiming) != hipSuccess)
      return bfail(b, HSI_ERR_DEVICE, "create the batch's graph stream");
  }
  if (!b->exec || b->graph_loading != loading) {
    batch_release_graph(b);
    if (hipStreamBeginCapture(b->graph_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return bfail(b, HSI_ERR_DEVICE, "begin graph capture");
    hsi_status st = hsi_batch_statistics(b, loading, b->graph_stream);
    if (st == HSI_OK) st = hsi_batch_score(b, b->graph_stream);
    hipGraph_t g = nullptr;
    const hipError_t end = hipStreamEndCapture(b->graph_stream, &g);
    if (st != HSI_OK || end != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return st != HSI_OK ? st : bfail(b, HSI_ERR_DEVICE, std::string("end graph capture: ") + hipGetErrorString(end));
    }
    b->graph = g;
    if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
      batch_release_graph(b);
      return bfail(b, HSI_ERR_DEVICE, "instantiate the batch graph");
    }
    b->graph_loading = loading;
  }
  const hipStream_t caller = static_cast<hipStream_t>(stream);
  if (hipEventRecord(b->graph_in, caller) != hipSuccess || hipStreamWaitEvent(b->graph_stream, b->graph_in, 0) != hipSuccess ||
      hipGraphLaunch(b->exec, b->graph_stream) != hipSuccess || hipEventRecord(b->graph_out, b->graph_stream) != hipSuccess ||
      hipStreamWaitEvent(caller, b->graph_out, 0) != hipSuccess)
    return bfail(b, HSI_ERR_DEVICE, "launch the batch graph");
  return HSI_OK;
}

hsi_status hsi_batch_autotune(hsi_batch* b, hsi_stream stream, char* report, size_t report_len) {
  if (!b) return HSI_ERR_INVALID_ARG;
  batch_release_graph(b);  // launch shapes change
  if (hipSetDevice(b->device) != hipSuccess) return bfail(b, HSI_ERR_DEVICE, "select device");
  const hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return bfail(b, HSI_ERR_DEVICE, "create events");
  std::string text;
  auto time_ms = [&](auto&& launch) -> float {
    if (launch() != hipSuccess) return -1.f;  // warm-up; an invalid shape is skipped
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < 3; ++r) (void)launch();
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    return ms / 3.f;
  };
  for (auto& c : b->sam) {
    float best = -1.f;
    for (int ppt : {4, 2}) {
      if (c.lib_stride > 8 && ppt == 4) continue;
      for (int wgs : {0, 2048, 8192, kPerScene}) {
        if (wgs == kPerScene && ppt != (c.lib_stride <= 8 ? 4 : 2)) continue;  // the single kernel's own shape
        const float ms = time_ms([&] { return run_sam_class(c, ppt, wgs, s); });
        if (ms >= 0.f && (best < 0.f || ms < best)) {
          best = ms;
          c.ppt = ppt;
          c.workgroups = wgs;
        }
      }
    }
    text += "sam/cem targets<=" + std::to_string(c.lib_stride) + " scenes=" + std::to_string(c.table.size()) +
            ": ppt " + std::to_string(c.ppt) + ", workgroups " +
            (c.workgroups > 0 ? std::to_string(c.workgroups) + " persistent"
                              : std::string(c.workgroups == kPerScene ? "per-scene launches" : "one per block")) + ", " +
            std::to_string(best * 1e3f) + " us\n";
  }
  for (auto& c : b->rx) {
    float best = -1.f;
    for (int chunk = 0; chunk < 3; ++chunk)
      for (int wgs : {0, 256, 512, 1024, kPerScene}) {
        if (wgs == kPerScene && chunk != 0) continue;
        const float ms = time_ms([&] { return run_rx_class(c, chunk, wgs, s); });
        if (ms >= 0.f && (best < 0.f || ms < best)) {
          best = ms;
          c.chunk = chunk;
          c.workgroups = wgs;
        }
      }
    text += "rx/ace bands<=" + std::to_string(16 * c.nb) + " target tiles=" + std::to_string(c.tt) +
            " scenes=" + std::to_string(c.table.size()) + ": tiles/item " + std::to_string(kRxChunks[c.chunk]) +
            ", workgroups " + (c.workgroups > 0 ? std::to_string(c.workgroups) + " persistent"
                                               : std::string(c.workgroups == kPerScene ? "per-scene launches" : "one per item")) +
            ", " + std::to_string(best * 1e3f) + " us\n";
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->tuned = true;
  if (report && report_len > 0) {
    std::strncpy(report, text.c_str(), report_len - 1);
    report[report_len - 1] = '\0';
  }
  return HSI_OK;
}

}  // extern "C"
