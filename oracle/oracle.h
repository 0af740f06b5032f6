/*
 * oracle.h - CPU restatement of the reference's frame-scoring arithmetic.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product library links, loads or
 * calls this code; it exists so tests, the smoke check and the benchmark's
 * CPU-baseline leg have an independent checker for the HIP kernels.
 *
 * Every function restates one reference function with the same operation
 * order, the same accumulation precision (double) and the same narrowing
 * points (double -> float), so that on the same inputs its outputs are
 * bit-identical to the compiled reference. That claim is pinned by the golden
 * fixtures under tests/golden/ (written by the compiled reference) and by the
 * property tests in tests/test_oracle_*.py.
 *
 * Cubes are band-sequential (BSQ): element (band b, pixel p) lives at
 * cube[b * plane_stride + p]. The reference always uses plane_stride ==
 * pixels; a larger stride is accepted so that a padded device layout can be
 * read without repacking. Scoring functions take a pixel range [p0, p1) so a
 * caller may split a frame across threads; outputs are indexed by absolute
 * pixel number.
 */
#ifndef HSI_ORACLE_H
#define HSI_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One detected pixel: 16 bytes, same layout as hsi::Detection
 * (include/hsi/types.hpp:101-106). */
typedef struct {
  int32_t x;
  int32_t y;
  int32_t target;
  float score;
} orc_detection;

/* ---- library (src/types.cpp:53-80) ------------------------------------- */
void orc_library_norms(const float* targets, int num_targets, int bands,
                       float* out_norms);

/* ---- SAM (src/sam_cpu.cpp:13-77) --------------------------------------- */
double orc_sam_angle(double dot, double pixel_sq, double target_norm);
void orc_sam_all(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                 int bands, const float* targets, const float* norms,
                 int num_targets, size_t out_stride, float* out_score);
void orc_sam_best(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                  int bands, const float* targets, const float* norms,
                  int num_targets, float* out_score, int32_t* out_target);

/* ---- detect (src/sam_cpu.cpp:86-138) ----------------------------------- */
/* Returns the number of detections found; writes at most `capacity`. */
int64_t orc_detect(const float* score, const int32_t* target, int height,
                   int width, float threshold, int nms_radius,
                   int higher_is_better, orc_detection* out, int64_t capacity);

/* ---- mean spectrum over a mask (src/sam_cpu.cpp:140-164) --------------- */
int orc_mean_spectrum(const float* cube, size_t pixels, int bands,
                      const uint8_t* mask, float* out_values);

/* ---- moments (src/linalg.cpp:10-103, src/detectors_cpu.cpp:77-93) ------ */
/* Adds pixels [p0, p1) into count/sum/sum_outer (lower triangle of a
 * row-major bands*bands matrix), exactly as SpectralMoments::add. */
void orc_moments_accumulate(const float* cube, size_t plane_stride, size_t p0,
                            size_t p1, int bands, double* count, double* sum,
                            double* sum_outer);
void orc_moments_from_shifted(const double* s1, const double* s2_lower,
                              double count, const double* reference, int bands,
                              double* out_sum, double* out_sum_outer);
void orc_moments_mean(double count, const double* sum, int bands, double* out);
void orc_moments_correlation(double count, const double* sum_outer, int bands,
                             double loading, double* out);
void orc_moments_covariance(double count, const double* sum,
                            const double* sum_outer, int bands, double loading,
                            double* out);

/* ---- dense linear algebra (src/linalg.cpp:105-180) --------------------- */
int orc_cholesky(double* a, int n);
void orc_cholesky_solve(const double* l, int n, const double* b, double* x);
void orc_cholesky_inverse(const double* l, int n, double* inv);
void orc_triangular_inverse_lower(const double* l, int n, double* inv);
double orc_quadratic_form(const double* m, const double* y, int n);

/* ---- statistics (src/detector.cpp:47-152) ------------------------------ */
/* Return codes for orc_build_statistics. */
enum {
  ORC_STATS_OK = 0,
  ORC_STATS_EMPTY = 1,
  ORC_STATS_BAND_MISMATCH = 2,
  ORC_STATS_COV_NOT_PD = 3,
  ORC_STATS_CORR_NOT_PD = 4
};
int orc_build_statistics(double count, const double* sum,
                         const double* sum_outer, int bands,
                         const float* targets, int num_targets,
                         int library_bands, double loading, float* mean,
                         float* whitener, float* target_white,
                         float* target_norm2, float* cem_weight, char* error,
                         int error_len);

/* ---- second-order detectors (src/detectors_cpu.cpp:95-187) ------------- */
void orc_rx(const float* cube, size_t plane_stride, size_t p0, size_t p1,
            int bands, const float* mean, const float* whitener,
            float* out_score);
void orc_ace_best(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                  int bands, const float* mean, const float* whitener,
                  const float* target_white, const float* target_norm2,
                  int num_targets, float* out_score, int32_t* out_target,
                  float* out_rx);
void orc_cem_best(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                  int bands, const float* cem_weight, int num_targets,
                  float* out_score, int32_t* out_target);

/* ---- chi-squared RX threshold (src/detectors_cpu.cpp:40-73, 189-206) --- */
double orc_gamma_p(double a, double x);
double orc_rx_threshold(int bands, double rate);

#ifdef __cplusplus
}
#endif

#endif /* HSI_ORACLE_H */
