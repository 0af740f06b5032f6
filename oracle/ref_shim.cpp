// ref_shim.cpp - a C ABI over the compiled reference CPU scorer.
//
// TEST INFRASTRUCTURE ONLY. oracle/Makefile compiles this file together with
// the reference's own sources (src/types.cpp, linalg.cpp, detector.cpp,
// detectors_cpu.cpp, sam_cpu.cpp), read in place from the reference checkout,
// into oracle/_ref/libhsiref.so. tests/golden/make_golden.py calls it to write
// the golden fixtures that tests/test_oracle_golden.py holds the oracle to,
// bit for bit. The product never loads it.
//
// Each wrapper builds the reference's own types and calls the reference
// function named in its comment; nothing here re-implements arithmetic.
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "hsi/detector.hpp"
#include "hsi/detectors.hpp"
#include "hsi/linalg.hpp"
#include "hsi/sam.hpp"
#include "hsi/types.hpp"

namespace {

hsi::CubeShape shape_of(int h, int w, int b) {
  hsi::CubeShape s;
  s.height = h;
  s.width = w;
  s.bands = b;
  return s;
}

hsi::SpectralLibrary library_of(const float* targets, int num_targets, int bands) {
  hsi::SpectralLibrary lib;
  for (int t = 0; t < num_targets; ++t) {
    hsi::Spectrum s;
    s.name = "t" + std::to_string(t);
    s.values.assign(targets + static_cast<std::size_t>(t) * bands,
                    targets + static_cast<std::size_t>(t + 1) * bands);
    lib.targets.push_back(std::move(s));
  }
  return lib;
}

hsi::SpectralMoments moments_of(const float* cube, hsi::CubeShape shape) {
  hsi::SpectralMoments m;
  hsi::accumulate_moments_cpu(cube, shape, &m);
  return m;
}

}  // namespace

extern "C" {

// SpectralLibrary::norms (types.cpp:68-76)
void ref_library_norms(const float* targets, int num_targets, int bands, float* out) {
  const std::vector<float> n = library_of(targets, num_targets, bands).norms();
  std::memcpy(out, n.data(), n.size() * sizeof(float));
}

// sam_best_cpu (sam_cpu.cpp:48-77)
void ref_sam_best(const float* cube, int h, int w, int b, const float* targets,
                  const float* norms, int num_targets, float* score, int32_t* target) {
  hsi::sam_best_cpu(cube, shape_of(h, w, b), targets, norms, num_targets, score, target);
}

// sam_cpu (sam_cpu.cpp:22-46)
void ref_sam_all(const float* cube, int h, int w, int b, const float* targets,
                 const float* norms, int num_targets, float* out) {
  hsi::sam_cpu(cube, shape_of(h, w, b), targets, norms, num_targets, out);
}

// detect_cpu (sam_cpu.cpp:116-138). Writes x, y, target triples and scores;
// returns the number of detections (all of them; capacity bounds the copy).
int64_t ref_detect(const float* score, const int32_t* target, int h, int w,
                   float threshold, int nms_radius, int higher_is_better,
                   int32_t* out_xyt, float* out_score, int64_t capacity) {
  hsi::DetectionParams params;
  params.threshold = threshold;
  params.nms_radius = nms_radius;
  params.polarity = higher_is_better ? hsi::ScorePolarity::HigherIsBetter
                                     : hsi::ScorePolarity::LowerIsBetter;
  const std::vector<hsi::Detection> found =
      hsi::detect_cpu(score, target, shape_of(h, w, 1), params);
  for (std::size_t i = 0; i < found.size() && static_cast<int64_t>(i) < capacity; ++i) {
    out_xyt[3 * i + 0] = found[i].x;
    out_xyt[3 * i + 1] = found[i].y;
    out_xyt[3 * i + 2] = found[i].target;
    out_score[i] = found[i].score;
  }
  return static_cast<int64_t>(found.size());
}

// accumulate_moments_cpu (detectors_cpu.cpp:77-93): count, sum[B], lower
// triangle of sum_outer[B*B] as the reference leaves it.
void ref_moments(const float* cube, int h, int w, int b, double* count,
                 double* sum, double* sum_outer) {
  const hsi::SpectralMoments m = moments_of(cube, shape_of(h, w, b));
  *count = m.count;
  std::memcpy(sum, m.sum.data(), m.sum.size() * sizeof(double));
  std::memcpy(sum_outer, m.sum_outer.data(), m.sum_outer.size() * sizeof(double));
}

// build_statistics (detector.cpp:47-152) from the cube's own moments.
// Returns 1 on success, 0 on failure with the reference's message in `error`.
int ref_build_statistics(const float* cube, int h, int w, int b, const float* targets,
                         int num_targets, int library_bands, double loading,
                         float* mean, float* whitener, float* target_white,
                         float* target_norm2, float* cem_weight, char* error,
                         int error_len) {
  const hsi::SpectralMoments m = moments_of(cube, shape_of(h, w, b));
  const hsi::SpectralLibrary lib = library_of(targets, num_targets, library_bands);
  hsi::DetectorStatistics stats;
  std::string message;
  const bool ok = hsi::build_statistics(m, lib, loading, &stats, &message);
  if (error && error_len > 0) {
    std::strncpy(error, message.c_str(), static_cast<std::size_t>(error_len - 1));
    error[error_len - 1] = '\0';
  }
  if (!ok) return 0;
  std::memcpy(mean, stats.mean.data(), stats.mean.size() * sizeof(float));
  std::memcpy(whitener, stats.whitener.data(), stats.whitener.size() * sizeof(float));
  std::memcpy(target_white, stats.target_white.data(), stats.target_white.size() * sizeof(float));
  std::memcpy(target_norm2, stats.target_norm2.data(), stats.target_norm2.size() * sizeof(float));
  std::memcpy(cem_weight, stats.cem_weight.data(), stats.cem_weight.size() * sizeof(float));
  return 1;
}

namespace {
hsi::DetectorStatistics stats_of(int bands, int num_targets, const float* mean,
                                 const float* whitener, const float* target_white,
                                 const float* target_norm2, const float* cem_weight) {
  hsi::DetectorStatistics s;
  s.bands = bands;
  s.num_targets = num_targets;
  const std::size_t b = static_cast<std::size_t>(bands);
  const std::size_t t = static_cast<std::size_t>(num_targets);
  s.mean.assign(mean, mean + b);
  s.whitener.assign(whitener, whitener + b * b);
  if (num_targets > 0) {
    s.target_white.assign(target_white, target_white + t * b);
    s.target_norm2.assign(target_norm2, target_norm2 + t);
    s.cem_weight.assign(cem_weight, cem_weight + t * b);
  }
  s.valid = true;
  return s;
}
}  // namespace

// rx_cpu (detectors_cpu.cpp:95-113)
void ref_rx(const float* cube, int h, int w, int b, const float* mean,
            const float* whitener, float* out) {
  hsi::rx_cpu(cube, shape_of(h, w, b), stats_of(b, 0, mean, whitener, nullptr, nullptr, nullptr),
              out);
}

// ace_best_cpu (detectors_cpu.cpp:115-156)
void ref_ace_best(const float* cube, int h, int w, int b, const float* mean,
                  const float* whitener, const float* target_white,
                  const float* target_norm2, const float* cem_weight, int num_targets,
                  float* score, int32_t* target, float* rx) {
  hsi::ace_best_cpu(cube, shape_of(h, w, b),
                    stats_of(b, num_targets, mean, whitener, target_white, target_norm2,
                             cem_weight),
                    score, target, rx);
}

// cem_best_cpu (detectors_cpu.cpp:158-187)
void ref_cem_best(const float* cube, int h, int w, int b, const float* mean,
                  const float* whitener, const float* target_white,
                  const float* target_norm2, const float* cem_weight, int num_targets,
                  float* score, int32_t* target) {
  hsi::cem_best_cpu(cube, shape_of(h, w, b),
                    stats_of(b, num_targets, mean, whitener, target_white, target_norm2,
                             cem_weight),
                    score, target);
}

// rx_threshold_for_false_alarm_rate (detectors_cpu.cpp:189-206)
double ref_rx_threshold(int bands, double rate) {
  return hsi::rx_threshold_for_false_alarm_rate(bands, rate);
}

// cholesky / triangular_inverse_lower / cholesky_solve (linalg.cpp:105-162)
int ref_cholesky(double* a, int n) { return hsi::cholesky(a, n) ? 1 : 0; }
void ref_triangular_inverse_lower(const double* l, int n, double* inv) {
  hsi::triangular_inverse_lower(l, n, inv);
}
void ref_cholesky_solve(const double* l, int n, const double* b, double* x) {
  hsi::cholesky_solve(l, n, b, x);
}

// cholesky_inverse (linalg.cpp:139-148)
void ref_cholesky_inverse(const double* l, int n, double* inv) { hsi::cholesky_inverse(l, n, inv); }

// quadratic_form (linalg.cpp:172-180)
double ref_quadratic_form(const double* m, const double* y, int n) {
  return hsi::quadratic_form(m, y, n);
}

// moments_from_shifted (linalg.cpp:81-103): the rebuilt sum[B] and the full
// sum_outer[B*B] as SpectralMoments holds it (upper triangle zero).
void ref_moments_from_shifted(const double* s1, const double* s2_lower, double count,
                              const double* reference, int bands, double* out_sum,
                              double* out_sum_outer) {
  hsi::SpectralMoments m;
  hsi::moments_from_shifted(s1, s2_lower, count, reference, bands, &m);
  std::memcpy(out_sum, m.sum.data(), m.sum.size() * sizeof(double));
  std::memcpy(out_sum_outer, m.sum_outer.data(), m.sum_outer.size() * sizeof(double));
}

// mean_spectrum (sam_cpu.cpp:140-164). Returns 1, or 0 when the mask selects
// no pixel (the reference throws).
int ref_mean_spectrum(const float* cube, int h, int w, int b, const uint8_t* mask, float* out) {
  try {
    const hsi::Spectrum s = hsi::mean_spectrum(cube, shape_of(h, w, b), mask, "mean");
    std::memcpy(out, s.values.data(), s.values.size() * sizeof(float));
    return 1;
  } catch (const std::invalid_argument&) {
    return 0;
  }
}

}  // extern "C"
