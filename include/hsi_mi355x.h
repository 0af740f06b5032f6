/*
 * hsi_mi355x.h - C ABI of the MI355X (gfx950) four-detector frame-scoring
 * library: SAM, CEM, RX and ACE, their scene statistics, and
 * threshold-and-compact.
 *
 * This is the drop-in boundary for the reference's device layer. It replaces
 * the C++ launchers in include/hsi/sam_cuda.hpp, stats_cuda.hpp and
 * detect_cuda.hpp, whose signatures carry cudaError_t/cudaStream_t, with
 * plain pointers, sizes and status codes. Each declaration names the
 * reference interface it stands in for.
 *
 * Conventions (kept from the reference, include/hsi/sam_cuda.hpp:59-96):
 *  - Device buffers (cubes, score maps, detection lists, counters) belong to
 *    the caller. The context owns only what it derives: the target library
 *    in device layouts, the scene statistics, and reusable scratch.
 *  - Launch functions are asynchronous on the caller's stream (`hsi_stream`
 *    is a hipStream_t; NULL is the default stream). Functions documented as
 *    synchronous block the host.
 *  - Cubes are band-sequential (BSQ) with a padded plane stride: element
 *    (band b, pixel p) is at cube[b * plane_stride + p], pixel p = y*W + x.
 *  - Errors come back as hsi_status; hsi_ctx_last_error() holds the message.
 *
 * A context is single-threaded: calls on one context must not race.
 */
#ifndef HSI_MI355X_H
#define HSI_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. The reference launchers return cudaErrorInvalidValue for
 * bad shapes, null pointers and an oversize library (sam.cu:347-353,
 * 371-376; detect.cu:91-97); build_statistics returns false with a message
 * naming --loading (detector.cpp:72-77, 97-100). */
typedef enum hsi_status {
  HSI_OK = 0,
  HSI_ERR_INVALID_ARG = 1,
  HSI_ERR_CAPACITY = 2,                /* library larger than supported */
  HSI_ERR_NOT_POSITIVE_DEFINITE = 3,   /* statistics failed; raise loading */
  HSI_ERR_DEVICE = 4,                  /* HIP runtime error */
  HSI_ERR_NO_STATISTICS = 5,           /* RX/ACE/CEM before statistics exist */
  HSI_ERR_NO_LIBRARY = 6               /* SAM/ACE/CEM before a library is set */
} hsi_status;

/* Detectors, in the order of hsi::Detector (include/hsi/detector.hpp:23-42). */
typedef enum hsi_detector {
  HSI_DET_SAM = 0,
  HSI_DET_RX = 1,
  HSI_DET_ACE = 2,
  HSI_DET_CEM = 3
} hsi_detector;

/* Element type of a device cube. F16 replaces the reference's SamVariant::Half
 * cube (sam_cuda.hpp:31, sam.cu:210-274). Q16 is a 16-bit per-band affine
 * code (no reference counterpart): element q of band b stands for
 * fmaf(scale[b], q, offset[b]) in FP32, with offset[b] the band's minimum and
 * scale[b] its range / 65535, so the code spends all 16 bits on the band's
 * own range where FP16 keeps 11 significant bits of each value. The code's
 * offsets and scales belong to the context (hsi_encode_q16,
 * hsi_set_q16_code); Q16 cubes are read by the MIXED statistics and the MIXED
 * RX/ACE kernels. */
typedef enum hsi_dtype { HSI_F32 = 0, HSI_F16 = 1, HSI_Q16 = 2 } hsi_dtype;

/* Arithmetic form. PLAIN accumulates in FP64 (the oracle's precision) on an
 * FP32 or FP16 cube. MIXED runs the RX/ACE whitening and the covariance on
 * FP16 matrix cores with FP32 accumulation and FP64 cross-tile reduction.
 * FP32 (RX/ACE scoring only; the statistics stay PLAIN) runs the RX/ACE
 * whitening and ACE's products on FP32 matrix cores (v_mfma_f32_16x16x4f32)
 * with FP32 accumulation: the plain method at single precision. */
typedef enum hsi_precision { HSI_PLAIN = 0, HSI_MIXED = 1, HSI_FP32 = 2 } hsi_precision;

/* 16-byte detection record, layout-identical to hsi::Detection
 * (include/hsi/types.hpp:101-106). */
typedef struct hsi_detection {
  int32_t x;
  int32_t y;
  int32_t target;
  float score;
} hsi_detection;

/* hsi::DetectionParams (include/hsi/types.hpp:122-140) with the polarity as
 * an int: 0 = lower is better (SAM), 1 = higher is better (RX, ACE, CEM). */
typedef struct hsi_detection_params {
  float threshold;
  int32_t nms_radius;
  int32_t max_detections;
  int32_t higher_is_better;
} hsi_detection_params;

typedef struct hsi_ctx hsi_ctx;
typedef void* hsi_stream;

/* ---- context ------------------------------------------------------------ */

/* Create a context on HIP device `device`. Replaces the process-global
 * constant-memory library of the reference (sam.cu:22-23): every context
 * holds its own library and statistics. */
hsi_status hsi_ctx_create(int device, hsi_ctx** out);
void hsi_ctx_destroy(hsi_ctx* ctx);
/* Message for the last failing call on this context ("" if none). */
const char* hsi_ctx_last_error(const hsi_ctx* ctx);
const char* hsi_status_name(hsi_status status);

/* Pre-size the context's scratch for frames up to max_pixels x max_bands so
 * that later launches never allocate (needed before stream capture).
 * Synchronous. */
hsi_status hsi_ctx_reserve(hsi_ctx* ctx, size_t max_pixels, int max_bands);

/* ---- geometry ----------------------------------------------------------- */

/* Plane stride in elements of a device BSQ cube. Below 65,536 pixels: the
 * pixel count rounded up to a multiple of 8, so every band plane starts
 * 16-byte aligned for FP32 and FP16 alike. From 65,536 pixels on: rounded up
 * to an odd multiple of 1024 elements, so consecutive planes start on
 * different HBM channels (a power-of-two plane size would put every plane on
 * the same one). Any stride >= height*width that is a multiple of 8 is
 * accepted by the scoring entry points; this one is the fastest to read.
 * Replaces bsq_plane_stride (sam_cuda.hpp:41-52, sam.cu:338-343). */
size_t hsi_plane_stride(int height, int width);

/* ---- target library ----------------------------------------------------- */

/* Largest library a context accepts (targets), and the largest band count. */
int hsi_max_targets(void);
int hsi_max_bands(void);

/* Install a target library: `values` is target-major [num_targets x bands]
 * host memory and `norms` the per-target L2 norms (float, as
 * SpectralLibrary::norms computes them, types.cpp:68-76), or NULL to have
 * them computed the same way. Synchronous. Replaces upload_targets
 * (sam_cuda.hpp:59-70, sam.cu:347-369), without its global state and with
 * no reciprocal-norm shortcut. */
hsi_status hsi_set_library(hsi_ctx* ctx, const float* values, const float* norms,
                           int num_targets, int bands);

/* ---- scene statistics --------------------------------------------------- */

/* Accumulate the scene's raw first and second moments over a device cube
 * (count, sum[B], lower triangle of sum x x^T) in FP64, replacing any
 * previous accumulation. PLAIN forms each product exactly in FP64 and reads
 * F32 cubes. MIXED reads F32, F16 or Q16 cubes: it accumulates the moments of
 * d = x - r, r = d_shift_or_null[B] (device floats, e.g. the previous frame's
 * mean: the reference's lagged mode, linalg.hpp:78-83) or, when NULL, the
 * mean of a sparse sample of the frame; products are formed on FP16 matrix
 * cores from a hi/lo split of d with FP32 tile accumulators, tiles are
 * reduced in FP64, and the raw moments are rebuilt from the shifted ones in
 * FP64 as moments_from_shifted does (linalg.cpp:81-103); like the mixed
 * scoring it reads the cube through one buffer view, so bands x plane_stride
 * elements must stay under 4 GiB (HSI_ERR_INVALID_ARG otherwise). PLAIN ignores the
 * shift (its products are exact). Replaces launch_reset_moments +
 * launch_accumulate_mean + launch_accumulate_moments (stats_cuda.hpp:45-72,
 * stats.cu:174-214; their `reference` argument is d_shift_or_null) and
 * SpectralMoments::add over a frame (detectors_cpu.cpp:77-93). */
hsi_status hsi_stats_accumulate(hsi_ctx* ctx, const void* d_cube, hsi_dtype dtype,
                                int height, int width, int bands, size_t plane_stride,
                                hsi_precision precision, const float* d_shift_or_null,
                                hsi_stream stream);

/* hsi_stats_accumulate (F32 cube, HSI_MIXED, same shift rule) with the
 * learned pre-screen (hsi_mlp_screen's network, installed for `bands`) run
 * on the same read of every pixel: no counterpart in the reference; it
 * replaces a statistics pass plus a screening pass over the frame with one.
 * The network's input is the pixel minus the shift - the caller's
 * d_shift_or_null (the previous frame's mean, say) or the sampled mean - in
 * FP16, since the frame's own mean is known only after the pass. Every pixel
 * whose logit is >= cut is listed in d_candidates[0 .. *d_count) as
 * {x, y, row, logit}, in a fixed order that is not raster order; with
 * d_spectra non-null (pitch spectra_stride >= 32 * ceil(bands / 32)) row
 * `row` receives the pixel's bands as the pass staged them, shift + (hi +
 * lo) in FP32 (within 2^-21 |x - shift| + 2^-23 |x| + 2^-24 of each value),
 * and 0
 * from `bands` to the 32-band boundary; rows are distinct but not in list
 * order. The pass splits the capacity evenly over its screening waves (at
 * most 1024, each screening its own 32 of every 128 pixels of a workgroup's
 * range); if any of them flags more than its share, or the total exceeds
 * `capacity`,
 * *d_count is set above `capacity` (the list overflowed, and
 * hsi_score_candidate_spectra then scores the whole frame). At most 127
 * bands, in a cube under 4 GiB as for the mixed statistics. Asynchronous: build the statistics, then rescore the rows. */
hsi_status hsi_stats_accumulate_screen(hsi_ctx* ctx, const float* d_cube, int height, int width,
                                       int bands, size_t plane_stride,
                                       const float* d_shift_or_null, float cut,
                                       hsi_detection* d_candidates, uint32_t* d_count,
                                       int capacity, float* d_spectra, size_t spectra_stride,
                                       hsi_stream stream);

/* Derive DetectorStatistics on the device in FP64 from the accumulated
 * moments and the installed library: covariance and autocorrelation with
 * diagonal loading, their Cholesky factors, the whitener L^-1, whitened
 * centred targets and their squared norms, and the CEM filters
 * R^-1 d / (d^T R^-1 d); narrowed to float as the reference does. A failed
 * factorisation is recorded in a device status word (read it with
 * hsi_stats_check). Replaces build_statistics (detector.cpp:47-152),
 * SpectralMoments::mean/covariance/correlation (linalg.cpp:45-79) and
 * cholesky/triangular_inverse_lower (linalg.cpp:105-162). Asynchronous. */
hsi_status hsi_stats_build(hsi_ctx* ctx, double diagonal_loading, hsi_stream stream);

/* Synchronous: wait for `stream`, then report HSI_OK or
 * HSI_ERR_NOT_POSITIVE_DEFINITE with the reference's message (which names
 * --loading) in hsi_ctx_last_error; when the failing pivot was positive but
 * below the device tolerance (HSI_STATUS_BELOW_TOLERANCE) the message says
 * so after the reference's text. */
hsi_status hsi_stats_check(hsi_ctx* ctx, hsi_stream stream);

/* Bits of the statistics status words. A Cholesky pivot fails when it is not
 * positive - the reference's rule (linalg.cpp:115-117) - or when it is
 * positive but at most 4*256*eps of its diagonal entry before factorisation:
 * the device rejects such rounding-level pivots, which a rank-deficient
 * matrix leaves in a blocked factorisation where the reference's sequential
 * one lands on exactly zero. BELOW_TOLERANCE is set beside COV/CORR_NOT_PD
 * in that second case; the reference would have built statistics there. */
#define HSI_STATUS_COV_NOT_PD 1u
#define HSI_STATUS_CORR_NOT_PD 2u
#define HSI_STATUS_BELOW_TOLERANCE 4u

/* Asynchronous: copy the two status words of the last build (covariance
 * path, autocorrelation path; 0 = factorised, else HSI_STATUS_* bits) to
 * `dst` (device or pinned host memory, 8 bytes) on `stream`. The
 * non-blocking form of hsi_stats_check, for pipelines and captured graphs.
 * No reference counterpart: build_statistics is synchronous there
 * (detector.cpp:47-152). */
hsi_status hsi_stats_status_copy(hsi_ctx* ctx, uint32_t* dst, hsi_stream stream);

/* Synchronous copy of the statistics to host memory; any pointer may be
 * NULL. Sizes: mean[B], whitener[B*B] (row-major, upper triangle zero),
 * target_white[T*B], target_norm2[T], cem_weight[T*B], sum[B],
 * sum_outer[B*B] (lower triangle). Replaces read_moments
 * (stats_cuda.hpp:167-173) and the DetectorStatistics fields
 * (detector.hpp:62-85). */
hsi_status hsi_stats_read(hsi_ctx* ctx, float* mean, float* whitener,
                          float* target_white, float* target_norm2,
                          float* cem_weight, double* count, double* sum,
                          double* sum_outer);

/* Synchronous: install statistics computed elsewhere (host memory, same
 * layouts as hsi_stats_read), e.g. a previous frame's, or the CPU scorer's
 * for a parity check. target_white/target_norm2/cem_weight may be NULL when
 * the library is empty. */
hsi_status hsi_stats_upload(hsi_ctx* ctx, int bands, const float* mean,
                            const float* whitener, const float* target_white,
                            const float* target_norm2, const float* cem_weight);

/* ---- scoring -------------------------------------------------------------- */

/* Score every pixel of a device cube with one detector.
 *   SAM: d_score = smallest spectral angle over the library, d_target = its
 *        index (first wins a tie). Replaces launch_sam_best
 *        (sam_cuda.hpp:72-96) and matches sam_best_cpu (sam_cpu.cpp:48-77).
 *   RX:  d_score = ||L^-1 (x - mean)||^2 (rx_cpu, detectors_cpu.cpp:95-113);
 *        d_target ignored.
 *   ACE: d_score = best (z.u)^2/(uu zz) clamped to [0,1], d_target its index
 *        or -1, d_aux (nullable) = the RX statistic from the same whitening
 *        (ace_best_cpu, detectors_cpu.cpp:115-156).
 *   CEM: d_score = best w_t . x, d_target its index (cem_best_cpu,
 *        detectors_cpu.cpp:158-187).
 * d_target may be NULL. PLAIN accumulates in FP64 and reads an F32 or F16
 * cube; MIXED (RX/ACE only) whitens on FP16 matrix cores. SAM and CEM maps
 * must be 16-byte aligned (any device allocation is), else INVALID_ARG. */
hsi_status hsi_score(hsi_ctx* ctx, hsi_detector detector, hsi_precision precision,
                     const void* d_cube, hsi_dtype dtype, int height, int width,
                     int bands, size_t plane_stride, float* d_score,
                     int32_t* d_target, float* d_aux, hsi_stream stream);

/* SAM and CEM from one read of the cube: the two bandwidth-bound detectors
 * share their pixel loads. Outputs as hsi_score for each; any of the four
 * output pointers may be NULL (a NULL score map skips that detector). */
hsi_status hsi_score_sam_cem(hsi_ctx* ctx, const void* d_cube, hsi_dtype dtype,
                             int height, int width, int bands, size_t plane_stride,
                             float* d_sam_score, int32_t* d_sam_target,
                             float* d_cem_score, int32_t* d_cem_target,
                             hsi_stream stream);

/* ---- threshold and compact ----------------------------------------------- */

/* Threshold a score map, suppress non-maxima within nms_radius (a neighbour
 * strictly better, or equal with a smaller linear index, suppresses), and
 * compact survivors into d_out in raster order, as detect_cpu does
 * (sam_cpu.cpp:86-138) - including polarity, which the reference kernel
 * ignores (detect.cu:43-87). *d_count receives the number found, which may
 * exceed max_detections; only the first max_detections are written. The
 * count is overwritten, so it need not be pre-zeroed. d_target may be NULL
 * (target 0). Replaces launch_detect (detect_cuda.hpp:16-19). */
hsi_status hsi_detect(hsi_ctx* ctx, const float* d_score, const int32_t* d_target,
                      int height, int width, hsi_detection_params params,
                      hsi_detection* d_out, uint32_t* d_count, hsi_stream stream);

/* hsi_detect without any allocation. No counterpart in the reference: for
 * callers that cannot allocate (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same arguments, results and
 * errors as hsi_detect, but its scratch is one static device array per GPU
 * (2.5 MiB, held until the process ends) instead of the context's, so it
 * calls no hipMalloc/hipFree. Frames of at most HSI_DETECT_NO_ALLOC_MAX_PIXELS
 * pixels (4096 x 4096); larger ones return HSI_ERR_CAPACITY. Every call on
 * one GPU shares that array: calls that overlap on one GPU (other streams or
 * threads) race and must be serialised by the caller. */
#define HSI_DETECT_NO_ALLOC_MAX_PIXELS 16777216u
hsi_status hsi_detect_no_alloc(hsi_ctx* ctx, const float* d_score, const int32_t* d_target,
                               int height, int width, hsi_detection_params params,
                               hsi_detection* d_out, uint32_t* d_count, hsi_stream stream);

/* ---- learned pre-screen --------------------------------------------------------
 *
 * No counterpart in the reference. A B -> 64 -> 32 -> 1 ReLU network (FP32
 * master weights, bf16 matrix-core inference with FP32 accumulation) scores
 * each pixel's chance of being an RX/ACE detection; pixels above a
 * conservative cut are rescored exactly (hsi_score_candidates) and the rest
 * are skipped. Parameters are packed W1[64][B] | b1[64] | W2[32][64] |
 * b2[32] | w3[32] | b3. One network per context. */

typedef struct hsi_mlp_train_params {
  float learning_rate; /* Adam step size */
  float beta1;         /* Adam first-moment decay, e.g. 0.9 */
  float beta2;         /* Adam second-moment decay, e.g. 0.999 */
  float epsilon;       /* Adam denominator floor, e.g. 1e-8 */
  float pos_weight;    /* loss weight of positive samples (recall over precision) */
} hsi_mlp_train_params;

/* Number of packed parameters for `bands` inputs (0 if out of range). */
size_t hsi_mlp_param_count(int bands);

/* Synchronous: a fresh network for `bands` inputs, seeded (He-uniform hidden
 * layers, zero biases), with Adam state cleared. */
hsi_status hsi_mlp_init(hsi_ctx* ctx, int bands, uint64_t seed);

/* Synchronous: install packed host weights; clears the Adam state. */
hsi_status hsi_mlp_set_weights(hsi_ctx* ctx, int bands, const float* weights);

/* Synchronous: copy the packed weights to host memory. */
hsi_status hsi_mlp_get_weights(hsi_ctx* ctx, float* weights);

/* Mean weighted binary cross-entropy gradient of a batch, without updating:
 * d_x is band-major, x[b * x_stride + s] for s < n (inputs already centred),
 * d_labels[n] in {0, 1}; d_grad receives hsi_mlp_param_count(B) floats and
 * d_loss (nullable) the mean loss. Asynchronous. */
hsi_status hsi_mlp_gradients(hsi_ctx* ctx, const float* d_x, size_t x_stride, int n,
                             const float* d_labels, float pos_weight, float* d_grad,
                             float* d_loss, hsi_stream stream);

/* One Adam step on a batch (same inputs as hsi_mlp_gradients). The forward
 * and backward passes, the gradient reductions and the update all run on
 * the device in a fixed order. Asynchronous. */
hsi_status hsi_mlp_train_step(hsi_ctx* ctx, const float* d_x, size_t x_stride, int n,
                              const float* d_labels, const hsi_mlp_train_params* params,
                              float* d_loss, hsi_stream stream);

/* Logit of every pixel of a device cube (F32 or F16), the input centred on
 * the statistics' mean when center_on_statistics_mean is non-zero (the
 * statistics must be built for this band count). Asynchronous. */
hsi_status hsi_mlp_infer(hsi_ctx* ctx, const void* d_cube, hsi_dtype dtype, int height,
                         int width, int bands, size_t plane_stride,
                         int center_on_statistics_mean, float* d_logit, hsi_stream stream);

/* Copy pixels d_index[0..n) of an F32 cube into a band-major sample matrix
 * out[b * out_stride + s], optionally centred on the statistics' mean: the
 * training-set builder. An index outside the frame yields a zero sample.
 * Asynchronous. */
hsi_status hsi_gather_pixels(hsi_ctx* ctx, const float* d_cube, int height, int width,
                             int bands, size_t plane_stride, const int32_t* d_index, int n,
                             int center_on_statistics_mean, float* d_out, size_t out_stride,
                             hsi_stream stream);

/* Exact (PLAIN) RX or ACE for the pixels of a candidate list only: the
 * first *d_count records of d_candidates (x, y as written by hsi_detect over
 * a logit map). Every other pixel of the score map gets 0 (ACE: target -1,
 * aux 0). If *d_count exceeds `capacity` the list overflowed and the whole
 * frame is scored instead (decided on the device, no host round trip).
 * Candidates' scores equal hsi_score's for the same pixels; records outside
 * the frame are skipped. F32 cube. Asynchronous. */
hsi_status hsi_score_candidates(hsi_ctx* ctx, hsi_detector detector, const float* d_cube,
                                int height, int width, int bands, size_t plane_stride,
                                const hsi_detection* d_candidates, const uint32_t* d_count,
                                int capacity, float* d_score, int32_t* d_target,
                                float* d_aux, hsi_stream stream);

/* Screening in one pass over the cube (F32, centred on the statistics'
 * mean, which must be built for this band count): every pixel whose logit
 * is >= cut is listed in d_candidates[0 .. *d_count) as {x, y, row, logit},
 * in a fixed order that is not raster order. With d_spectra non-null its
 * bands are written to row `row` of d_spectra ([capacity][spectra_stride],
 * spectra_stride >= 32 * ceil(bands / 32), a multiple of 4, 16-byte
 * aligned rows; bands past the last read 0) - straight from the pass, with
 * no copy: the rows are distinct but not in list order (each wave of the
 * pass owns a block of them); with d_spectra NULL, row is 0. The pass splits the capacity
 * evenly over its waves; if any wave finds more candidates than its share,
 * or the total exceeds `capacity`, *d_count is set above `capacity` (the
 * list overflowed; hsi_score_candidate_spectra then scores the whole frame).
 * d_logit (nullable) receives the full logit map. The library keeps scratch
 * of the list's size. Asynchronous. */
hsi_status hsi_mlp_screen(hsi_ctx* ctx, const float* d_cube, int height, int width, int bands,
                          size_t plane_stride, float cut, hsi_detection* d_candidates,
                          uint32_t* d_count, int capacity, float* d_spectra,
                          size_t spectra_stride, float* d_logit, hsi_stream stream);

/* hsi_score_candidates over the rows that hsi_mlp_screen wrote: record i
 * is scored from row d_candidates[i].target of d_spectra and written at its
 * (x, y). Only a list that hsi_mlp_screen wrote (with these spectra) is
 * valid input: the target field of any other list (hsi_detect's target
 * index) names no spectra row. Rows: capacity rows of pitch spectra_stride
 * >= bands; a record whose row is outside [0, capacity) is skipped; a row's
 * elements past `bands` are ignored. The whole-frame fallback on overflow
 * reads d_cube. HSI_PLAIN: the same results as hsi_score_candidates.
 * HSI_MIXED: the precision form of hsi_score (FP16 hi/lo matrix cores) on
 * the rows, and on the cube for the fallback; the rows must be 16-byte
 * aligned (spectra_stride a multiple of 4) and span less than 2 GiB
 * (capacity * spectra_stride * 4 bytes). Asynchronous. */
hsi_status hsi_score_candidate_spectra(hsi_ctx* ctx, hsi_detector detector, hsi_precision precision,
                                       const float* d_cube, int height, int width, int bands,
                                       size_t plane_stride, const hsi_detection* d_candidates,
                                       const uint32_t* d_count, int capacity, const float* d_spectra,
                                       size_t spectra_stride, float* d_score, int32_t* d_target,
                                       float* d_aux, hsi_stream stream);

/* ---- offline evaluation ------------------------------------------------------
 *
 * Replaces scripts/eval_detection.py (roc_auc :40-66, operating point and
 * class table :100-140) for a device score map and a per-pixel class mask
 * (uint8). Pixels with label <= max_label are annotated; label ==
 * target_class is positive, every other annotated label negative.
 *   auc: Mann-Whitney AUC with ties counted half (the rank sum with
 *        tie-averaged ranks), exact; the score is ranked as is when
 *        higher_is_better, negated otherwise (SAM angles); NaN when one side
 *        is empty. NaN scores rank above every number.
 *   tp, fp, fn, tn: predicted = score >= threshold (higher_is_better) or
 *        score <= threshold, against the positive label.
 * class_stats (nullable, host, 3 * (max_label + 1) doubles): per label its
 * pixel count, mean score and minimum score (NaN scores skipped in the
 * minimum). Synchronous; the library keeps scratch of the map's size. */
typedef struct hsi_eval_result {
  double auc;
  uint64_t positives;
  uint64_t negatives;
  uint64_t tp;
  uint64_t fp;
  uint64_t fn;
  uint64_t tn;
} hsi_eval_result;

hsi_status hsi_evaluate(hsi_ctx* ctx, const float* d_score, const uint8_t* d_labels, size_t n,
                        int target_class, int max_label, int higher_is_better, float threshold,
                        hsi_eval_result* result, double* class_stats, hsi_stream stream);

/* ---- batched scenes ----------------------------------------------------------
 *
 * No counterpart in the reference (its pipeline scores one shape at a time).
 * Scenes of uneven size, band count (<= 256; RX/ACE <= 224) and library
 * size (1-32 targets) are scored together: each scene keeps its own
 * context (library, statistics), and SAM/CEM and RX/ACE each run as one
 * persistent launch per kernel shape class over device tables of the
 * scenes. hsi_batch_autotune times the candidate launch shapes (pixels per
 * thread, grid size, tiles per work item) on the batch and keeps the
 * fastest. Cubes are F32 with a hsi_plane_stride stride; SAM/CEM maps are
 * 16-byte aligned. A NULL score map skips that detector for the scene; RX
 * and ACE of a scene share one pass when both are requested. */
typedef struct hsi_batch hsi_batch;

typedef struct hsi_scene {
  const float* d_cube;
  int height, width, bands;
  size_t plane_stride;
  const float* library;  /* host [num_targets x bands], NULL if none */
  int num_targets;
  float* d_sam_score;
  int32_t* d_sam_target;
  float* d_cem_score;
  int32_t* d_cem_target;
  float* d_rx_score;
  float* d_ace_score;
  int32_t* d_ace_target;
} hsi_scene;

hsi_status hsi_batch_create(int device, hsi_batch** out);
void hsi_batch_destroy(hsi_batch* batch);
const char* hsi_batch_last_error(const hsi_batch* batch);

/* Synchronous: install the scenes (libraries uploaded, device tables built).
 * The cubes and maps named here must stay valid while the batch is used. */
hsi_status hsi_batch_set_scenes(hsi_batch* batch, const hsi_scene* scenes, int count);

/* Statistics of every scene that needs them (RX, ACE, CEM): moments scene
 * by scene, then every scene's FP64 build in one launch per build form.
 * Asynchronous. */
hsi_status hsi_batch_statistics(hsi_batch* batch, double diagonal_loading, hsi_stream stream);

/* Synchronous: HSI_OK, or the first scene whose factorisation failed. */
hsi_status hsi_batch_check(hsi_batch* batch, int* failed_scene);

/* All requested detectors of all scenes: one persistent launch per shape
 * class. Asynchronous. */
hsi_status hsi_batch_score(hsi_batch* batch, hsi_stream stream);

/* hsi_batch_statistics then hsi_batch_score. With use_graph, the first call
 * (and the first after hsi_batch_set_scenes, hsi_batch_autotune or a change
 * of loading) captures both as one hipGraph on a stream of the batch's own,
 * and every call replays it there, ordered after the caller's earlier work
 * on `stream` and before its later work (events). The cubes' contents may
 * change between calls; their addresses may not. Asynchronous. */
hsi_status hsi_batch_run(hsi_batch* batch, double diagonal_loading, int use_graph, hsi_stream stream);

/* Synchronous: time the candidate launch shapes of every class on the
 * current scenes (statistics must be built) and keep the fastest; a text
 * report of the choices is written to `report` (may be NULL). */
hsi_status hsi_batch_autotune(hsi_batch* batch, hsi_stream stream, char* report, size_t report_len);

/* The context of scene `scene` (for hsi_detect, hsi_stats_read, ...), or
 * NULL. Owned by the batch. */
hsi_ctx* hsi_batch_context(hsi_batch* batch, int scene);

/* ---- ingest helpers ---------------------------------------------------------- */

/* Encode an F32 device cube as Q16 on the device: one pass finds each band's
 * minimum and maximum over the frame's pixels, a second writes
 * q = round((x - offset[b]) / scale[b]) with offset[b] = minimum and
 * scale[b] = (maximum - minimum) / 65535 (0 for a constant band, coded 0),
 * and the context keeps offset/scale as its Q16 code for `bands` bands.
 * Decoding is within scale[b] / 2 of x (plus FP32 rounding). Elements must
 * be finite. d_dst has the same plane stride as d_src (uint16 elements);
 * plane padding is coded 0. Asynchronous. The storage form of the precision
 * path: no reference counterpart (the reference narrows to FP16,
 * sam.cu:311-316). */
hsi_status hsi_encode_q16(hsi_ctx* ctx, const float* d_src, int height, int width,
                          int bands, size_t plane_stride, uint16_t* d_dst, hsi_stream stream);

/* Synchronous: install a Q16 code (host offset[bands], scale[bands]) on the
 * context, e.g. the one another context encoded with; and read the current
 * code back (HSI_ERR_NO_STATISTICS if there is none). */
hsi_status hsi_set_q16_code(hsi_ctx* ctx, int bands, const float* offset, const float* scale);
hsi_status hsi_get_q16_code(hsi_ctx* ctx, int bands, float* offset, float* scale);

/* FP32 -> FP16 narrowing of `count` elements (round to nearest even). Both
 * buffers must be 16-byte aligned (INVALID_ARG otherwise); any count.
 * Replaces launch_narrow_to_half (sam_cuda.hpp:107-109, sam.cu:311-316). */
hsi_status hsi_narrow_to_half(const float* d_src, void* d_dst, size_t count,
                              hsi_stream stream);

/* BIP -> padded BSQ transpose on the device. Replaces
 * launch_transpose_bip_to_bsq (sam_cuda.hpp:98-105, sam.cu:281-309). */
hsi_status hsi_transpose_bip_to_bsq(const float* d_src, float* d_dst, int height,
                                    int width, int bands, size_t plane_stride,
                                    hsi_stream stream);

/* Sample type of a raw snapshot-mosaic frame. */
typedef enum hsi_raw_type { HSI_RAW_U8 = 0, HSI_RAW_U16 = 1, HSI_RAW_F32 = 2 } hsi_raw_type;

/* Demosaic a snapshot-mosaic frame (raw_height x raw_width samples, a
 * period x period filter tile) on the device into a BSQ cube of
 * (raw_height / period) x (raw_width / period) pixels and `bands` planes:
 * band b takes the samples at mosaic position band_order[b] (row
 * position / period, column position % period of each tile), or position b
 * when band_order is NULL; integer samples are scaled by 1/255 or 1/65535 in
 * float. Replaces the host demosaic of the hot-folder loader
 * (hot_loader.cpp:144-182). band_order is host memory, copied; synchronous
 * with respect to that copy, then asynchronous. */
hsi_status hsi_demosaic(hsi_ctx* ctx, const void* d_raw, hsi_raw_type raw_type, int raw_height,
                        int raw_width, int period, int bands, const int32_t* band_order,
                        float* d_cube, size_t plane_stride, hsi_stream stream);

/* ---- host helpers ------------------------------------------------------------ */

/* RX threshold for a per-pixel false-alarm rate under the chi-squared
 * approximation with `bands` degrees of freedom; NaN for invalid arguments
 * (the reference throws). Replaces rx_threshold_for_false_alarm_rate
 * (detectors.hpp:56-62, detectors_cpu.cpp:189-206). */
double hsi_rx_threshold(int bands, double false_alarm_rate);

#ifdef __cplusplus
}
#endif

#endif /* HSI_MI355X_H */
