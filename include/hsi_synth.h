/*
 * hsi_synth.h - seeded synthetic hyperspectral frames (host side).
 *
 * The frame source for benchmarks, tests and the CLIs. It plays the role of
 * the reference's SyntheticSource (include/hsi/frame_source.hpp:194-210,
 * src/frame_source.cpp:149-293) but draws the scenes the benchmark
 * configurations specify, which the reference generator does not produce:
 *
 *   - endmembers: a linear ramp plus six Gaussian bumps over the normalised
 *     wavelength, affinely mapped onto [0.05, 0.9];
 *   - background pixel: a Dirichlet(alpha = 1) mix of the first
 *     `num_endmembers` endmembers, with an optional per-frame drift of the
 *     mixing weights;
 *   - targets: library entry t (an endmember) planted in `plant_fraction` of
 *     the pixels at abundance U(0.6, 1.0), disjoint across targets;
 *   - anomalies: an extra endmember outside the library at abundance
 *     U(0.3, 1.0) in `anomaly_fraction` of the remaining pixels;
 *   - per-pixel gain U(0.3, 1.2), then additive N(0, noise_sigma) clipped at 0.
 *
 * Every random draw is a SplitMix64 hash of (seed, stream, index), so a frame
 * is a pure function of (config, frame index): identical bytes on any host,
 * thread count or call order. Generation runs on the host so the CPU oracle
 * and the GPU see the same floats.
 */
#ifndef HSI_SYNTH_H
#define HSI_SYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hsi_synth_config {
  int32_t height;
  int32_t width;
  int32_t bands;
  int32_t num_targets;       /* library size (endmembers 0.. are targets) */
  int32_t num_endmembers;    /* background endmembers mixed per pixel */
  int32_t drift_period;      /* frames per drift cycle */
  double plant_fraction;     /* fraction of pixels per planted target */
  double anomaly_fraction;   /* fraction of pixels holding the anomaly */
  double noise_sigma;        /* additive Gaussian noise */
  double drift;              /* relative amplitude of the per-frame mix drift */
  uint64_t seed;
} hsi_synth_config;

/* Defaults of the plain-method scene: 256x256x113, 8 endmembers, 4 targets
 * at 0.5%, no anomalies, sigma 0.005, no drift. */
void hsi_synth_default(hsi_synth_config* config);

/* Target library, [num_targets x bands] floats. Returns 0, or -1 on a bad
 * configuration. */
int hsi_synth_library(const hsi_synth_config* config, float* out);

/* The anomaly endmember, [bands] floats. */
int hsi_synth_anomaly(const hsi_synth_config* config, float* out);

/* Write frame `frame` as BSQ with `plane_stride` >= height*width elements
 * between band planes (padding zeroed). `truth`, if not NULL, receives one
 * byte per pixel: 0 background, t+1 for planted target t, 255 anomaly.
 * `threads` <= 0 uses every hardware thread. Returns 0 or -1. */
int hsi_synth_frame(const hsi_synth_config* config, uint64_t frame, float* dst,
                    size_t plane_stride, uint8_t* truth, int threads);

#ifdef __cplusplus
}
#endif

#endif /* HSI_SYNTH_H */
