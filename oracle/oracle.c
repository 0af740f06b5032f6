/*
 * oracle.c - CPU restatement of the reference scorer (test infrastructure).
 *
 * Only tests/, the smoke check and the benchmark's CPU-baseline leg use this
 * file. Each function names the reference function it restates; the
 * arithmetic (double accumulation in index order, float narrowing at the same
 * points, libm sqrt/acos/lgamma/exp/log) is kept identical so outputs are
 * bit-identical to the compiled reference on the same inputs.
 *
 * Build with -ffp-contract=off: the reference is built for plain x86-64,
 * where no fused multiply-add is formed, and a contracted a*b+c would round
 * differently wherever the product is inexact.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* float(pi/2), promoted: include/hsi/sam.hpp:25 kMaxSamAngle. */
static const float kMaxSamAngleF = 1.57079632679489661923f;

/* ------------------------------------------------------------------------ */
/* src/types.cpp:68-76 SpectralLibrary::norms: double sum of squares, sqrt,
 * narrowed to float. */
void orc_library_norms(const float* targets, int num_targets, int bands,
                       float* out_norms) {
  for (int t = 0; t < num_targets; ++t) {
    double acc = 0.0;
    const float* v = targets + (size_t)t * (size_t)bands;
    for (int b = 0; b < bands; ++b) acc += (double)v[b] * v[b];
    out_norms[t] = (float)sqrt(acc);
  }
}

/* src/sam_cpu.cpp:13-18 angle_of. */
double orc_sam_angle(double dot, double pixel_sq, double target_norm) {
  if (pixel_sq <= 0.0 || target_norm <= 0.0) return (double)kMaxSamAngleF;
  double cosine = dot / (sqrt(pixel_sq) * target_norm);
  if (cosine < -1.0) cosine = -1.0;
  if (cosine > 1.0) cosine = 1.0;
  return acos(cosine);
}

/* Band-strided gather of one pixel's squared norm: sam_cpu.cpp:31-35. */
static double pixel_sq_of(const float* cube, size_t stride, size_t p,
                          int bands) {
  double s = 0.0;
  for (int b = 0; b < bands; ++b) {
    const double v = cube[(size_t)b * stride + p];
    s += v * v;
  }
  return s;
}

static double dot_of(const float* cube, size_t stride, size_t p, int bands,
                     const float* target) {
  double dot = 0.0;
  for (int b = 0; b < bands; ++b) {
    dot += (double)cube[(size_t)b * stride + p] * target[b];
  }
  return dot;
}

/* src/sam_cpu.cpp:22-46 sam_cpu: every target, target-major output. */
void orc_sam_all(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                 int bands, const float* targets, const float* norms,
                 int num_targets, size_t out_stride, float* out_score) {
  for (size_t p = p0; p < p1; ++p) {
    const double pixel_sq = pixel_sq_of(cube, plane_stride, p, bands);
    for (int t = 0; t < num_targets; ++t) {
      const double dot = dot_of(cube, plane_stride, p, bands,
                                targets + (size_t)t * (size_t)bands);
      out_score[(size_t)t * out_stride + p] =
          (float)orc_sam_angle(dot, pixel_sq, (double)norms[t]);
    }
  }
}

/* src/sam_cpu.cpp:48-77 sam_best_cpu: smallest angle, first target wins a
 * tie (strict <), -1 if no target produced a finite angle. */
void orc_sam_best(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                  int bands, const float* targets, const float* norms,
                  int num_targets, float* out_score, int32_t* out_target) {
  for (size_t p = p0; p < p1; ++p) {
    const double pixel_sq = pixel_sq_of(cube, plane_stride, p, bands);
    double best = INFINITY;
    int32_t best_target = -1;
    for (int t = 0; t < num_targets; ++t) {
      const double dot = dot_of(cube, plane_stride, p, bands,
                                targets + (size_t)t * (size_t)bands);
      const double a = orc_sam_angle(dot, pixel_sq, (double)norms[t]);
      if (a < best) {
        best = a;
        best_target = t;
      }
    }
    out_score[p] = (float)best;
    out_target[p] = best_target;
  }
}

/* ------------------------------------------------------------------------ */
/* src/sam_cpu.cpp:86-93 better / over_threshold. */
static int better(float a, float b, int higher) { return higher ? a > b : a < b; }
static int over_threshold(float s, float thr, int higher) {
  return higher ? s >= thr : s <= thr;
}

/* src/sam_cpu.cpp:96-112 local_best: suppressed by a strictly better
 * neighbour, or an equal one with a smaller linear index. */
static int local_best(const float* scores, int x, int y, int height, int width,
                      int radius, float self, int higher) {
  const size_t self_index = (size_t)y * (size_t)width + (size_t)x;
  for (int dy = -radius; dy <= radius; ++dy) {
    const int ny = y + dy;
    if (ny < 0 || ny >= height) continue;
    for (int dx = -radius; dx <= radius; ++dx) {
      if (dx == 0 && dy == 0) continue;
      const int nx = x + dx;
      if (nx < 0 || nx >= width) continue;
      const size_t n = (size_t)ny * (size_t)width + (size_t)nx;
      if (better(scores[n], self, higher)) return 0;
      if (scores[n] == self && n < self_index) return 0;
    }
  }
  return 1;
}

/* src/sam_cpu.cpp:116-138 detect_cpu: raster order. */
int64_t orc_detect(const float* score, const int32_t* target, int height,
                   int width, float threshold, int nms_radius,
                   int higher_is_better, orc_detection* out, int64_t capacity) {
  int64_t found = 0;
  for (int y = 0; y < height; ++y) {
    for (int x = 0; x < width; ++x) {
      const size_t p = (size_t)y * (size_t)width + (size_t)x;
      if (!over_threshold(score[p], threshold, higher_is_better)) continue;
      if (nms_radius > 0 && !local_best(score, x, y, height, width, nms_radius,
                                        score[p], higher_is_better)) {
        continue;
      }
      if (found < capacity && out) {
        out[found].x = x;
        out[found].y = y;
        out[found].target = target ? target[p] : 0;
        out[found].score = score[p];
      }
      ++found;
    }
  }
  return found;
}

/* src/sam_cpu.cpp:140-164 mean_spectrum. Returns 0, or -1 when the mask
 * selects nothing (the reference throws). */
int orc_mean_spectrum(const float* cube, size_t pixels, int bands,
                      const uint8_t* mask, float* out_values) {
  size_t count = 0;
  double* acc = (double*)calloc((size_t)bands, sizeof(double));
  if (!acc) return -1;
  for (size_t p = 0; p < pixels; ++p) {
    if (!mask[p]) continue;
    ++count;
    for (int b = 0; b < bands; ++b) acc[b] += cube[(size_t)b * pixels + p];
  }
  if (count == 0) {
    free(acc);
    return -1;
  }
  for (int b = 0; b < bands; ++b) out_values[b] = (float)(acc[b] / (double)count);
  free(acc);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* src/detectors_cpu.cpp:77-93 accumulate_moments_cpu, with
 * src/linalg.cpp:18-29 SpectralMoments::add per pixel. */
void orc_moments_accumulate(const float* cube, size_t plane_stride, size_t p0,
                            size_t p1, int bands, double* count, double* sum,
                            double* sum_outer) {
  float* narrow = (float*)malloc((size_t)bands * sizeof(float));
  if (!narrow) return;
  for (size_t p = p0; p < p1; ++p) {
    for (int b = 0; b < bands; ++b) narrow[b] = cube[(size_t)b * plane_stride + p];
    for (int i = 0; i < bands; ++i) {
      const double vi = narrow[i];
      sum[i] += vi;
      double* row = sum_outer + (size_t)i * (size_t)bands;
      for (int j = 0; j <= i; ++j) row[j] += vi * narrow[j];
    }
    *count += 1.0;
  }
  free(narrow);
}

/* src/linalg.cpp:81-103 moments_from_shifted. */
void orc_moments_from_shifted(const double* s1, const double* s2_lower,
                              double count, const double* reference, int bands,
                              double* out_sum, double* out_sum_outer) {
  memset(out_sum, 0, (size_t)bands * sizeof(double));
  memset(out_sum_outer, 0, (size_t)bands * (size_t)bands * sizeof(double));
  if (count <= 0.0) return;
  for (int i = 0; i < bands; ++i) out_sum[i] = s1[i] + count * reference[i];
  for (int i = 0; i < bands; ++i) {
    for (int j = 0; j <= i; ++j) {
      out_sum_outer[(size_t)i * bands + j] =
          s2_lower[(size_t)i * bands + j] + reference[i] * s1[j] +
          s1[i] * reference[j] + count * reference[i] * reference[j];
    }
  }
}

/* src/linalg.cpp:45-50 SpectralMoments::mean. */
void orc_moments_mean(double count, const double* sum, int bands, double* out) {
  for (int i = 0; i < bands; ++i) out[i] = sum[i] / count;
}

/* src/linalg.cpp:52-64 SpectralMoments::correlation: lower triangle / N,
 * mirrored, loading on the diagonal. */
void orc_moments_correlation(double count, const double* sum_outer, int bands,
                             double loading, double* out) {
  for (int i = 0; i < bands; ++i) {
    for (int j = 0; j <= i; ++j) {
      const double v = sum_outer[(size_t)i * bands + j] / count;
      out[(size_t)i * bands + j] = v;
      out[(size_t)j * bands + i] = v;
    }
    out[(size_t)i * bands + i] += loading;
  }
}

/* src/linalg.cpp:66-79 SpectralMoments::covariance: correlation(0) - mu mu^T,
 * loading on the diagonal. */
void orc_moments_covariance(double count, const double* sum,
                            const double* sum_outer, int bands, double loading,
                            double* out) {
  double* mu = (double*)malloc((size_t)bands * sizeof(double));
  if (!mu) return;
  orc_moments_correlation(count, sum_outer, bands, 0.0, out);
  orc_moments_mean(count, sum, bands, mu);
  for (int i = 0; i < bands; ++i) {
    for (int j = 0; j <= i; ++j) {
      const double v = out[(size_t)i * bands + j] - mu[i] * mu[j];
      out[(size_t)i * bands + j] = v;
      out[(size_t)j * bands + i] = v;
    }
    out[(size_t)i * bands + i] += loading;
  }
  free(mu);
}

/* ------------------------------------------------------------------------ */
/* src/linalg.cpp:105-122 cholesky: in place, lower, upper zeroed. */
int orc_cholesky(double* a, int n) {
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = a[(size_t)i * n + j];
      for (int k = 0; k < j; ++k) s -= a[(size_t)i * n + k] * a[(size_t)j * n + k];
      if (i == j) {
        if (!(s > 0.0)) return 0;
        a[(size_t)i * n + i] = sqrt(s);
      } else {
        a[(size_t)i * n + j] = s / a[(size_t)j * n + j];
      }
    }
    for (int j = i + 1; j < n; ++j) a[(size_t)i * n + j] = 0.0;
  }
  return 1;
}

/* src/linalg.cpp:124-137 cholesky_solve: forward then back substitution. */
void orc_cholesky_solve(const double* l, int n, const double* b, double* x) {
  for (int i = 0; i < n; ++i) {
    double s = b[i];
    for (int k = 0; k < i; ++k) s -= l[(size_t)i * n + k] * x[k];
    x[i] = s / l[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = x[i];
    for (int k = i + 1; k < n; ++k) s -= l[(size_t)k * n + i] * x[k];
    x[i] = s / l[(size_t)i * n + i];
  }
}

/* src/linalg.cpp:139-148 cholesky_inverse: one solve per unit vector. */
void orc_cholesky_inverse(const double* l, int n, double* inv) {
  double* e = (double*)calloc((size_t)n, sizeof(double));
  double* col = (double*)calloc((size_t)n, sizeof(double));
  if (!e || !col) {
    free(e);
    free(col);
    return;
  }
  for (int k = 0; k < n; ++k) {
    memset(e, 0, (size_t)n * sizeof(double));
    e[k] = 1.0;
    orc_cholesky_solve(l, n, e, col);
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + k] = col[i];
  }
  free(e);
  free(col);
}

/* src/linalg.cpp:150-162 triangular_inverse_lower: column by column. */
void orc_triangular_inverse_lower(const double* l, int n, double* inv) {
  memset(inv, 0, (size_t)n * (size_t)n * sizeof(double));
  for (int k = 0; k < n; ++k) {
    inv[(size_t)k * n + k] = 1.0 / l[(size_t)k * n + k];
    for (int i = k + 1; i < n; ++i) {
      double s = 0.0;
      for (int j = k; j < i; ++j) s -= l[(size_t)i * n + j] * inv[(size_t)j * n + k];
      inv[(size_t)i * n + k] = s / l[(size_t)i * n + i];
    }
  }
}

/* src/linalg.cpp:172-180 quadratic_form. */
double orc_quadratic_form(const double* m, const double* y, int n) {
  double total = 0.0;
  for (int i = 0; i < n; ++i) {
    double row = 0.0;
    for (int j = 0; j < n; ++j) row += m[(size_t)i * n + j] * y[j];
    total += y[i] * row;
  }
  return total;
}

/* ------------------------------------------------------------------------ */
/* std::to_string(double) prints with "%f". */
static void format_error(char* error, int error_len, const char* fmt,
                         double loading) {
  if (!error || error_len <= 0) return;
  char number[64];
  snprintf(number, sizeof number, "%f", loading);
  snprintf(error, (size_t)error_len, fmt, number);
}

/* src/detector.cpp:47-152 build_statistics. `library_bands` is the band
 * count of the targets (the reference checks it against the moments). */
int orc_build_statistics(double count, const double* sum,
                         const double* sum_outer, int bands,
                         const float* targets, int num_targets,
                         int library_bands, double loading, float* mean,
                         float* whitener, float* target_white,
                         float* target_norm2, float* cem_weight, char* error,
                         int error_len) {
  if (error && error_len > 0) error[0] = '\0';
  if (count <= 0.0) {
    if (error && error_len > 0)
      snprintf(error, (size_t)error_len, "no scene statistics accumulated yet");
    return ORC_STATS_EMPTY;
  }
  if (num_targets > 0 && library_bands != bands) {
    if (error && error_len > 0)
      snprintf(error, (size_t)error_len,
               "library has %d bands but the statistics cover %d",
               library_bands, bands);
    return ORC_STATS_BAND_MISMATCH;
  }
  const size_t nn = (size_t)bands * (size_t)bands;
  double* mu = (double*)malloc((size_t)bands * sizeof(double));
  double* cov = (double*)malloc(nn * sizeof(double));
  double* white = (double*)malloc(nn * sizeof(double));
  double* corr = (double*)malloc(nn * sizeof(double));
  double* centred = (double*)malloc((size_t)bands * sizeof(double));
  double* raw = (double*)malloc((size_t)bands * sizeof(double));
  double* solved = (double*)malloc((size_t)bands * sizeof(double));
  int rc = ORC_STATS_OK;
  if (!mu || !cov || !white || !corr || !centred || !raw || !solved) {
    rc = ORC_STATS_EMPTY;
    goto done;
  }

  orc_moments_mean(count, sum, bands, mu);
  orc_moments_covariance(count, sum, sum_outer, bands, loading, cov);
  if (!orc_cholesky(cov, bands)) {
    format_error(error, error_len,
                 "scene covariance is not positive definite at loading %s; "
                 "hyperspectral bands are strongly correlated, so try a larger "
                 "--loading",
                 loading);
    rc = ORC_STATS_COV_NOT_PD;
    goto done;
  }
  orc_triangular_inverse_lower(cov, bands, white);
  for (int i = 0; i < bands; ++i) mean[i] = (float)mu[i];
  for (size_t i = 0; i < nn; ++i) whitener[i] = (float)white[i];
  for (size_t i = 0; i < (size_t)num_targets * (size_t)bands; ++i) {
    target_white[i] = 0.0f;
    cem_weight[i] = 0.0f;
  }
  for (int t = 0; t < num_targets; ++t) target_norm2[t] = 0.0f;
  if (num_targets == 0) goto done;

  orc_moments_correlation(count, sum_outer, bands, loading, corr);
  if (!orc_cholesky(corr, bands)) {
    format_error(error, error_len,
                 "scene autocorrelation is not positive definite at loading "
                 "%s; try a larger --loading",
                 loading);
    rc = ORC_STATS_CORR_NOT_PD;
    goto done;
  }
  for (int t = 0; t < num_targets; ++t) {
    const float* d = targets + (size_t)t * (size_t)bands;
    for (int i = 0; i < bands; ++i) centred[i] = (double)d[i] - mu[i];
    double norm2 = 0.0;
    for (int i = 0; i < bands; ++i) {
      double s = 0.0;
      const double* row = white + (size_t)i * bands;
      for (int j = 0; j <= i; ++j) s += row[j] * centred[j];
      target_white[(size_t)t * bands + i] = (float)s;
      norm2 += s * s;
    }
    target_norm2[t] = (float)norm2;

    for (int i = 0; i < bands; ++i) raw[i] = d[i];
    orc_cholesky_solve(corr, bands, raw, solved);
    double gain = 0.0;
    for (int i = 0; i < bands; ++i) gain += raw[i] * solved[i];
    if (gain > 0.0) {
      for (int i = 0; i < bands; ++i)
        cem_weight[(size_t)t * bands + i] = (float)(solved[i] / gain);
    }
  }

done:
  free(mu);
  free(cov);
  free(white);
  free(corr);
  free(centred);
  free(raw);
  free(solved);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* src/detectors_cpu.cpp:26-36 whiten: z = L^-1 y with a float whitener,
 * double accumulation; returns ||z||^2. */
static double whiten(const float* whitener, int bands, const double* y,
                     double* z) {
  double norm2 = 0.0;
  for (int i = 0; i < bands; ++i) {
    const float* row = whitener + (size_t)i * bands;
    double s = 0.0;
    for (int j = 0; j <= i; ++j) s += (double)row[j] * y[j];
    z[i] = s;
    norm2 += s * s;
  }
  return norm2;
}

/* src/detectors_cpu.cpp:95-113 rx_cpu. */
void orc_rx(const float* cube, size_t plane_stride, size_t p0, size_t p1,
            int bands, const float* mean, const float* whitener,
            float* out_score) {
  double* y = (double*)malloc((size_t)bands * sizeof(double));
  double* z = (double*)malloc((size_t)bands * sizeof(double));
  if (!y || !z) {
    free(y);
    free(z);
    return;
  }
  for (size_t p = p0; p < p1; ++p) {
    for (int b = 0; b < bands; ++b)
      y[b] = (double)cube[(size_t)b * plane_stride + p] - (double)mean[b];
    out_score[p] = (float)whiten(whitener, bands, y, z);
  }
  free(y);
  free(z);
}

/* src/detectors_cpu.cpp:115-156 ace_best_cpu. */
void orc_ace_best(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                  int bands, const float* mean, const float* whitener,
                  const float* target_white, const float* target_norm2,
                  int num_targets, float* out_score, int32_t* out_target,
                  float* out_rx) {
  double* y = (double*)malloc((size_t)bands * sizeof(double));
  double* z = (double*)malloc((size_t)bands * sizeof(double));
  if (!y || !z) {
    free(y);
    free(z);
    return;
  }
  for (size_t p = p0; p < p1; ++p) {
    for (int b = 0; b < bands; ++b)
      y[b] = (double)cube[(size_t)b * plane_stride + p] - (double)mean[b];
    const double zz = whiten(whitener, bands, y, z);
    if (out_rx) out_rx[p] = (float)zz;
    double best = -1.0;
    int32_t best_target = -1;
    for (int t = 0; t < num_targets; ++t) {
      const float* u = target_white + (size_t)t * bands;
      const double uu = target_norm2[t];
      if (!(uu > 0.0) || !(zz > 0.0)) continue;
      double zu = 0.0;
      for (int b = 0; b < bands; ++b) zu += z[b] * (double)u[b];
      const double s = (zu * zu) / (uu * zz);
      if (s > best) {
        best = s;
        best_target = t;
      }
    }
    /* best < 0 ? 0 : std::min(best, 1.0), where std::min(a, b) is
     * (b < a) ? b : a. */
    out_score[p] = (float)(best < 0.0 ? 0.0 : (1.0 < best ? 1.0 : best));
    if (out_target) out_target[p] = best_target;
  }
  free(y);
  free(z);
}

/* src/detectors_cpu.cpp:158-187 cem_best_cpu: filter on the uncentred pixel,
 * strict > from -inf, so the first target wins a tie. */
void orc_cem_best(const float* cube, size_t plane_stride, size_t p0, size_t p1,
                  int bands, const float* cem_weight, int num_targets,
                  float* out_score, int32_t* out_target) {
  for (size_t p = p0; p < p1; ++p) {
    double best = -INFINITY;
    int32_t best_target = -1;
    for (int t = 0; t < num_targets; ++t) {
      const float* w = cem_weight + (size_t)t * bands;
      double response = 0.0;
      for (int b = 0; b < bands; ++b)
        response += (double)w[b] * (double)cube[(size_t)b * plane_stride + p];
      if (response > best) {
        best = response;
        best_target = t;
      }
    }
    out_score[p] = (float)best;
    if (out_target) out_target[p] = best_target;
  }
}

/* ------------------------------------------------------------------------ */
/* src/detectors_cpu.cpp:40-73 gamma_p: series below a+1, Lentz continued
 * fraction above. */
double orc_gamma_p(double a, double x) {
  if (x <= 0.0) return 0.0;
  const double log_gamma_a = lgamma(a);
  if (x < a + 1.0) {
    double term = 1.0 / a;
    double sum = term;
    for (int n = 1; n < 1000; ++n) {
      term *= x / (a + n);
      sum += term;
      if (fabs(term) < fabs(sum) * 1e-16) break;
    }
    return sum * exp(-x + a * log(x) - log_gamma_a);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - a;
  double c = 1.0 / tiny;
  double d = 1.0 / b;
  double h = d;
  for (int n = 1; n < 1000; ++n) {
    const double an = -n * (n - a);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double delta = d * c;
    h *= delta;
    if (fabs(delta - 1.0) < 1e-16) break;
  }
  const double q = exp(-x + a * log(x) - log_gamma_a) * h;
  return 1.0 - q;
}

/* src/detectors_cpu.cpp:189-206 rx_threshold_for_false_alarm_rate. Returns
 * NaN for an invalid argument (the reference throws). */
double orc_rx_threshold(int bands, double rate) {
  if (bands <= 0) return NAN;
  if (!(rate > 0.0) || !(rate < 1.0)) return NAN;
  const double a = 0.5 * bands;
  const double want = 1.0 - rate;
  double low = 0.0;
  double high = 4.0 * bands > 16.0 ? 4.0 * bands : 16.0;
  while (orc_gamma_p(a, 0.5 * high) < want && high < 1e9) high *= 2.0;
  for (int i = 0; i < 200; ++i) {
    const double mid = 0.5 * (low + high);
    if (orc_gamma_p(a, 0.5 * mid) < want) low = mid; else high = mid;
  }
  return 0.5 * (low + high);
}
