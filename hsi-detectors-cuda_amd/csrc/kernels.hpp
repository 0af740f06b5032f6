// kernels.hpp - launchers shared between the kernel files and the C ABI.
#pragma once

#include "common.hpp"

namespace hsi_amd {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef _Float16 h4 __attribute__((ext_vector_type(4)));

// FP16 hi/lo split of two FP32 values d0, d1: hi = both rounded to FP16 by
// one packed conversion, lo = d - hi (exact in FP32, hi read back from its
// packed halves) rounded by another - the same values as splitting each
// value on its own, with two conversions fewer per pair. Returns {hi pair,
// lo pair} as packed words.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef _Float16 h2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint2 split_pair(float d0, float d1) {
  const h2v h = __builtin_convertvector((f2v{d0, d1}), h2v);
  const f2v hf = __builtin_convertvector(h, f2v);
  const h2v l = __builtin_convertvector((f2v{d0 - hf.x, d1 - hf.y}), h2v);
  return make_uint2(__builtin_bit_cast(unsigned, h), __builtin_bit_cast(unsigned, l));
}

// Eight values split into FP16 hi and lo parts (split_pair four times).
__device__ __forceinline__ void split8(const float (&v)[8], h8& hi, h8& lo) {
  unsigned hw[4], lw[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const uint2 s = split_pair(v[2 * q], v[2 * q + 1]);
    hw[q] = s.x;
    lw[q] = s.y;
  }
  typedef unsigned u4v __attribute__((ext_vector_type(4)));
  hi = __builtin_bit_cast(h8, (u4v{hw[0], hw[1], hw[2], hw[3]}));
  lo = __builtin_bit_cast(h8, (u4v{lw[0], lw[1], lw[2], lw[3]}));
}

// ---- SAM / CEM ---------------------------------------------------------------
struct SamCemArgs {
  const void* cube = nullptr;
  hsi_dtype dtype = HSI_F32;
  std::size_t pixels = 0;
  int bands = 0;
  std::size_t stride = 0;
  int num_targets = 0;
  int lib_stride = 1;              // sam_cem_lib_stride(num_targets)
  const double* sam_lib = nullptr;  // [bands][lib_stride] targets, zero padded
  const float* sam_norms = nullptr; // [num_targets]
  const double* sam_rnorms = nullptr; // [lib_stride] 1/norm (0 padded), for ranking only
  int sam_exact_all = 0;            // 1 if some target has norm <= 0: exact path throughout
  const double* cem_lib = nullptr;  // [bands][lib_stride] CEM filters, zero padded
  float* sam_score = nullptr;
  int* sam_target = nullptr;
  float* cem_score = nullptr;
  int* cem_target = nullptr;
};
int sam_cem_lib_stride(int num_targets);
hipError_t launch_sam_cem(const SamCemArgs& args, hipStream_t stream);

// ---- statistics --------------------------------------------------------------
// Bands padded for the 16x16 matrix-core blocks, with one extra "ones" band
// at index `bands` so the same accumulation yields sum(x) and the count.
__host__ __device__ inline int moment_blocks(int bands) { return (bands + 1 + 15) / 16; }
inline int moment_pairs(int bands) {
  const int nb = moment_blocks(bands);
  return nb * (nb + 1) / 2;
}
// Scratch doubles for the per-workgroup partials.
std::size_t moments_scratch_doubles(std::size_t pixels, int bands);
struct MomentArgs {
  const void* cube = nullptr;
  hsi_dtype dtype = HSI_F32;       // PLAIN: F32 only; MIXED: F32, F16, Q16
  std::size_t pixels = 0;
  int bands = 0;
  std::size_t stride = 0;
  hsi_precision precision = HSI_PLAIN;
  const float* shift = nullptr;    // MIXED: caller's reference [bands], or null (sample mean)
  const float* q_offset = nullptr; // Q16 code [bands]
  const float* q_scale = nullptr;
  double* count = nullptr;
  double* sum = nullptr;
  double* sum_outer = nullptr;
};
hipError_t launch_moments(const MomentArgs& args, double* scratch, hipStream_t stream);
// Whether the mixed-form moments (screen: with the fused screening pass) can
// read a cube of `bands` planes of `stride` elements of `elem_bytes` bytes
// through one buffer view (under 4 GiB, with every staging offset below 2^32).
bool mixed_moments_fit(int bands, std::size_t stride, std::size_t elem_bytes, bool screen);

// The learned pre-screen fused into the mixed-form moments (F32 cube, bands
// <= kFusedScreenMaxBands): from the same read of each pixel, the network's
// logit of (pixel - shift) - the FP16 hi part the moments stage - and, at
// logit >= cut, a record {x, y, row, logit} with the pixel's spectrum
// (shift + hi + lo in FP32) in spectra row `row` (nullable; pitch >= 32 *
// ceil(bands / 32), elements from `bands` to that boundary written 0). Each
// wave of the pass owns cap / waves rows; the records are compacted into
// cand[0 .. *count) afterwards (the rows stay where the pass wrote them), and
// a wave that found more than its share raises *count past `cap`.
// counts / scratch_rec: 2 * kScreenMaxWaves counters and `cap` records of
// scratch, as hsi_mlp_screen uses.
constexpr int kFusedScreenMaxBands = 127;
struct FusedScreenArgs {
  float cut = 0.f;
  int width = 0;
  const void* w1h = nullptr;      // FP16 layer-1 fragments (launch_mlp_layouts a1h)
  const void* w2h = nullptr;      // FP16 layer-2 fragments (a2h)
  const float* tail = nullptr;    // b1[64] | b2[32] | w3[32] | b3
  hsi_detection* cand = nullptr;
  unsigned* count = nullptr;
  unsigned cap = 0;
  float* spectra = nullptr;
  std::size_t spectra_stride = 0;
  unsigned* counts = nullptr;
  hsi_detection* scratch_rec = nullptr;
};
hipError_t launch_moments_screen(const MomentArgs& args, const FusedScreenArgs& fs, double* scratch,
                                 hipStream_t stream);

struct BuildArgs {
  int bands = 0;
  int num_targets = 0;
  double loading = 0.0;
  const double* count = nullptr;
  const double* sum = nullptr;
  const double* sum_outer = nullptr;
  const float* targets = nullptr;  // [T][B]
  double* work = nullptr;          // build_work_doubles(bands, T)
  float* mean = nullptr;
  float* whitener = nullptr;
  float* target_white = nullptr;
  float* target_norm2 = nullptr;
  float* cem_weight = nullptr;
  unsigned* status = nullptr;
};
std::size_t build_work_doubles(int bands, int num_targets);
hipError_t launch_build(const BuildArgs& args, hipStream_t stream);
// Whether a build of `bands` keeps its matrices in LDS (one kernel form) or
// in its global work area (the other); a batched build launches each form
// once, for all scenes of that form in a device table, with `loading`
// replacing each entry's.
bool build_in_lds(int bands);
hipError_t launch_build_batch(const BuildArgs* d_table, int count, bool in_lds, double loading, hipStream_t stream);

// Derived layouts the scoring kernels read, refreshed after every build or
// upload: CEM filters as doubles band-major for the SAM/CEM kernel; the
// whitener as doubles in 16x16 lower blocks, and the whitened targets, for
// the RX/ACE kernels.
struct LayoutArgs {
  int bands = 0;
  int num_targets = 0;
  int lib_stride = 1;
  const float* whitener = nullptr;
  const float* target_white = nullptr;
  const float* cem_weight = nullptr;
  const float* mean = nullptr;
  double* cem_lib = nullptr;        // [bands][lib_stride]
  float* white_blocks = nullptr;    // rxace_white_floats(bands)
  float* ace_u = nullptr;           // rxace_target_floats(bands, T)
};
hipError_t launch_layouts(const LayoutArgs& args, hipStream_t stream);

// Layouts of the mixed-precision RX/ACE kernel: the whitener and whitened
// targets rescaled by powers of two and split into FP16 hi + lo parts.
struct MixedLayoutArgs {
  int bands = 0;
  int num_targets = 0;
  const float* whitener = nullptr;
  const float* target_white = nullptr;
  _Float16* white_hl = nullptr;   // mixed_white_halves(bands)
  _Float16* ace_hl = nullptr;     // mixed_target_halves(bands, T)
  float* col_scale = nullptr;     // [kMaxBands] y' = y * col_scale
  float* row_inv = nullptr;       // [kMaxBands] z = z'' * row_inv
  float* target_inv = nullptr;    // [kMaxTargets] z.u = (z.u)'' * target_inv
};
std::size_t mixed_white_halves(int bands);
std::size_t mixed_target_halves(int bands, int num_targets);
hipError_t launch_mixed_layouts(const MixedLayoutArgs& args, hipStream_t stream);

// ---- RX / ACE ----------------------------------------------------------------
__host__ __device__ inline int rx_blocks(int bands) { return (bands + 15) / 16; }
std::size_t rxace_white_floats(int bands);
std::size_t rxace_target_floats(int bands, int num_targets);

struct RxAceArgs {
  const void* cube = nullptr;
  hsi_dtype dtype = HSI_F32;
  std::size_t pixels = 0;
  int bands = 0;
  std::size_t stride = 0;
  int num_targets = 0;  // 0 for RX only
  const float* mean = nullptr;
  const float* white_blocks = nullptr;
  const float* ace_u = nullptr;
  const float* target_norm2 = nullptr;
  const _Float16* white_hl = nullptr;
  const _Float16* ace_hl = nullptr;
  const float* col_scale = nullptr;
  const float* row_inv = nullptr;
  const float* target_inv = nullptr;
  const float* q_offset = nullptr;  // Q16 code [bands] (dtype HSI_Q16)
  const float* q_scale = nullptr;
  float* rx = nullptr;
  float* ace_score = nullptr;
  int* ace_target = nullptr;
};
hipError_t launch_rxace_plain(const RxAceArgs& args, hipStream_t stream);
// The plain kernel over a candidate list (raster-order detections of a
// pre-screen): scores only pixels (y * width + x) of the first
// min(*count, capacity) records, written at their own positions. With
// spectra non-null, record i's bands are read from row cand[i].target of
// spectra (pitch spectra_stride >= bands; a row outside [0, capacity) names
// no pixel) instead of the cube.
hipError_t launch_rxace_candidates(const RxAceArgs& args, const hsi_detection* cand, const unsigned* count,
                                   unsigned capacity, int width, const float* spectra,
                                   std::size_t spectra_stride, hipStream_t stream);
hipError_t launch_fill_scores(float* score, int* target, float* aux, std::size_t pixels, hipStream_t stream);

// ---- offline evaluation ------------------------------------------------------------
constexpr int kEvalMaxClasses = 32;
std::size_t eval_scratch_bytes(std::size_t n);
// Synchronous on `stream`. out_host: [0] P, [1] N, [2..5] tp fp fn tn,
// [6] AUC numerator (2 * below + tied, summed over positives), class c at
// 8 + 3c: count, mean, min.
hipError_t launch_evaluate(const float* score, const std::uint8_t* label, std::size_t n, int target, int max_label,
                           int higher, float threshold, void* scratch, std::size_t scratch_bytes, double* out_host,
                           hipStream_t stream);

// Batched plain RX/ACE: scenes of one row-block count NB (= rx_blocks(bands))
// and target-tile count TT (0: RX only, 1: <= 16 targets, 2: <= 32), FP32
// cubes, statistics layouts already derived.
struct RxSceneDev {
  const void* cube;
  std::size_t pixels;
  int bands;
  std::size_t stride;
  int num_targets;
  const float* mean;
  const float* white_blocks;
  const float* ace_u;
  const float* target_norm2;
  float* rx;
  float* ace_score;
  int* ace_target;
};
struct RxWorkItem {
  int scene;
  std::size_t tile_begin, tile_end;  // 16-pixel tiles of that scene
};
// Batched SAM/CEM over scenes sharing a library stride (next power of two of
// the target count); FP32 cubes. block_begin numbers the scenes' 256*PPT
// pixel blocks in order; a NULL score map skips that detector for the scene.
struct SamSceneDev {
  const void* cube;
  std::size_t pixels;
  int bands;
  std::size_t stride;
  int num_targets;
  int exact_all;
  const double* sam_lib;
  const float* sam_norms;
  const double* sam_rnorms;
  const double* cem_lib;
  float* sam_score;
  int* sam_target;
  float* cem_score;
  int* cem_target;
  unsigned block_begin;
};
// block_scene[b] = scene of pixel block b. SAM and CEM run fused up to 16
// targets, at 32 targets as two launches. workgroups > 0: a persistent grid
// of that size; <= 0: one workgroup per block.
hipError_t launch_sam_cem_batch(int lib_stride, int ppt, const SamSceneDev* scenes, const int* block_scene,
                                unsigned total_blocks, int workgroups, hipStream_t stream);
// workgroups > 0: a persistent grid of that size; <= 0: one workgroup per item.
hipError_t launch_rxace_batch(int nb, int tt, const RxSceneDev* scenes, const RxWorkItem* work, int num_items,
                              int workgroups, hipStream_t stream);
hipError_t launch_rxace_mixed(const RxAceArgs& args, hipStream_t stream);
// The FP32 form of the plain kernel: whitening and ACE products on
// v_mfma_f32_16x16x4f32 with FP32 accumulation (same layouts as plain).
hipError_t launch_rxace_fp32(const RxAceArgs& args, hipStream_t stream);
// The precision form over candidate spectra rows, with launch_rxace_candidates'
// list semantics (FP32 rows; spectra_stride >= bands; the rows' bytes below
// 2 GiB); an overflowed list scores the whole FP32 cube.
hipError_t launch_rxace_mixed_candidates(const RxAceArgs& args, const hsi_detection* cand, const unsigned* count,
                                         unsigned capacity, int width, const float* spectra,
                                         std::size_t spectra_stride, hipStream_t stream);

// ---- learned pre-screen (mlp.hip) --------------------------------------------
struct MlpInferArgs {
  const void* cube = nullptr;
  hsi_dtype dtype = HSI_F32;
  std::size_t pixels = 0;
  int bands = 0;
  std::size_t stride = 0;
  const float* center = nullptr;  // nullable: no centring
  const void* a1 = nullptr;       // bf16 layer-1 fragments
  const void* a2 = nullptr;       // bf16 layer-2 fragments
  const float* tail = nullptr;    // b1 | b2 | w3 | b3
  float* logit = nullptr;         // nullable when screening
  // Screening (cand non-null, F32 cube): pixels with logit >= cut are
  // listed in cand[0 .. *cand_count) (x, y, 0, logit), in a fixed order
  // (region of the wave that scanned them, then raster order); with spectra
  // non-null each one's bands go to a row of spectra (pixel-major, pitch
  // spectra_stride >= 32 * ceil(bands / 32), a multiple of 4, cand_cap rows),
  // the row a wave's region slot, named in the record's target. A count above cand_cap
  // means the list overflowed. The scratch holds the per-wave regions:
  // cand_cap records,
  // 2 * kScreenMaxWaves counters.
  float cut = 0.f;
  hsi_detection* cand = nullptr;
  unsigned* cand_count = nullptr;
  unsigned cand_cap = 0;
  float* spectra = nullptr;
  std::size_t spectra_stride = 0;
  int width = 0;
  hsi_detection* scratch_rec = nullptr;
  unsigned* scratch_counts = nullptr;
};
constexpr int kScreenMaxWaves = 4096;
// Per-wave candidate regions -> one list: counts[0 .. waves) of regions of rc
// records (rec_in) are scanned into prefix[] and copied to rec_out[0 ..
// total); *total is raised past cap when a region overflowed.
hipError_t launch_screen_compact(const unsigned* counts, int waves, unsigned rc, unsigned cap, unsigned* prefix,
                                 unsigned* total, const hsi_detection* rec_in, hsi_detection* rec_out,
                                 hipStream_t stream);
std::size_t mlp_param_count(int bands);
std::size_t mlp_a1_halves(int bands);
std::size_t mlp_a2_halves();
std::size_t mlp_tail_floats();
std::size_t mlp_act_floats(int n);
// a1h / a2h (nullable): FP16 copies of the a1 / a2 fragments (same halves).
hipError_t launch_mlp_layouts(const float* w, int bands, void* a1, void* a2, float* tail, void* a1h, void* a2h,
                              hipStream_t stream);
hipError_t launch_mlp_infer(const MlpInferArgs& args, hipStream_t stream);
hipError_t launch_mlp_gradients(const float* x, std::size_t x_stride, int n, int bands, const float* labels,
                                float pos_weight, const float* w, float* act, float* grad, float* loss,
                                hipStream_t stream);
hipError_t launch_adam(float* w, float* m, float* v, const float* g, std::size_t count, float lr, float beta1,
                       float beta2, float eps, float c1, float c2, hipStream_t stream);
hipError_t launch_gather(const float* cube, std::size_t stride, std::size_t pixels, int bands, const int* index,
                         int n, const float* center, float* out, std::size_t out_stride, hipStream_t stream);

// ---- ingest (ingest.hip) -------------------------------------------------------
// Q16 encoding of an F32 cube (16-byte aligned planes, stride % 8 == 0):
// per-band extrema into `extrema` (q16_extrema_words(bands) words of
// scratch), then the codes into dst and the code itself into q_offset /
// q_scale [bands].
std::size_t q16_extrema_words(int bands);
hipError_t launch_encode_q16(const float* src, std::size_t pixels, int bands, std::size_t stride,
                             unsigned* extrema, std::uint16_t* dst, float* q_offset, float* q_scale,
                             hipStream_t stream);
// raw_type: 0 = u8, 1 = u16, 2 = f32; d_position[bands] = mosaic positions.
hipError_t launch_demosaic(const void* raw, int raw_type, int raw_height, int raw_width, int period, int bands,
                           const int* d_position, float* cube, std::size_t stride, hipStream_t stream);

// ---- detect ------------------------------------------------------------------
std::size_t detect_scratch_words(std::size_t pixels);
// The static scratch of hsi_detect_no_alloc covers frames up to this size
// (4096 x 4096 pixels: 2.5 MiB of device memory per GPU, held until the
// process ends). Address on the current device, looked up once per device.
constexpr std::size_t kDetectNoAllocMaxPixels = std::size_t(1) << 24;
hipError_t detect_no_alloc_scratch(unsigned** out);
hipError_t launch_detect(const float* score, const int* target, int height, int width,
                         hsi_detection_params params, hsi_detection* out, unsigned* count,
                         unsigned* scratch, hipStream_t stream);

// ---- ingest ------------------------------------------------------------------
hipError_t launch_narrow_to_half(const float* src, _Float16* dst, std::size_t count,
                                 hipStream_t stream);
hipError_t launch_transpose_bip_to_bsq(const float* src, float* dst, int height, int width,
                                       int bands, std::size_t stride, hipStream_t stream);

}  // namespace hsi_amd
