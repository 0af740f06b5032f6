// hsi_host.hpp - C++ host layer over the hsi_mi355x C ABI.
//
// The reference's callers (its hsi_detect and hsi_score programs, its
// benchmarks) are written against a small C++ vocabulary: CubeShape,
// SpectralLibrary, Detection, DetectionParams, Detector, FrameSource and
// Pipeline (include/hsi/types.hpp, detector.hpp, frame_source.hpp,
// pipeline.hpp). This header offers the same names with the same meanings,
// implemented on HIP streams, pinned staging and hipGraph-captured frames,
// so those callers move over without a rewrite. Everything that touches the
// device goes through include/hsi_mi355x.h.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <string>
#include <vector>

#include "hsi_mi355x.h"
#include "hsi_synth.h"

namespace hsi {

// Cube geometry (types.hpp:27-46): int fields in (height, width, bands)
// order, size_t pixel count.
struct CubeShape {
  int height = 0;
  int width = 0;
  int bands = 0;
  std::size_t pixels() const { return static_cast<std::size_t>(height) * width; }
  std::size_t elements() const { return pixels() * static_cast<std::size_t>(bands); }
  bool valid() const { return height > 0 && width > 0 && bands > 0; }
  bool operator==(const CubeShape& o) const {
    return height == o.height && width == o.width && bands == o.bands;
  }
  bool operator!=(const CubeShape& o) const { return !(*this == o); }
};

enum class Interleave { Bsq, Bil, Bip };

// One named spectrum (types.hpp:49-60).
struct Spectrum {
  std::string name;
  std::vector<float> values;
  std::vector<float> wavelengths_nm;
  int bands() const { return static_cast<int>(values.size()); }
};

// Target library (types.hpp:62-80): target-major flattening and float norms
// computed as a double sum narrowed to float (types.cpp:53-80).
struct SpectralLibrary {
  std::vector<Spectrum> targets;
  int size() const { return static_cast<int>(targets.size()); }
  int bands() const { return targets.empty() ? 0 : targets.front().bands(); }
  std::vector<float> flatten() const;
  std::vector<float> norms() const;
};

// Detection record, byte-identical to hsi_detection (types.hpp:101-106).
struct Detection {
  int x = 0;
  int y = 0;
  int target = 0;
  float score = 0.0f;
};
static_assert(sizeof(Detection) == sizeof(hsi_detection), "Detection must stay a 16-byte record");

enum class ScorePolarity { LowerIsBetter, HigherIsBetter };  // types.hpp:114-117

// types.hpp:122-140.
struct DetectionParams {
  float threshold = 0.10f;
  int nms_radius = 2;
  int max_detections = 4096;
  ScorePolarity polarity = ScorePolarity::LowerIsBetter;
};

// detector.hpp:23-55.
enum class Detector { Sam, Rx, Ace, Cem };
struct DetectorTraits {
  const char* name;
  bool needs_targets;
  bool needs_statistics;
  ScorePolarity polarity;
  bool compute_bound;  // per-pixel cost grows with bands^2
};
DetectorTraits traits_of(Detector detector);
bool parse_detector(const std::string& name, Detector* out);
const char* to_string(Detector detector);
// Comma-separated list, e.g. "sam,cem".
bool parse_detector_list(const std::string& names, std::vector<Detector>* out);

// ---- files (envi.cpp, spectra_io.cpp) ---------------------------------------

struct HsiFrame {
  CubeShape shape;
  std::vector<float> data;  // BSQ, plane stride = shape.pixels()
  std::vector<float> wavelengths_nm;
};

// ENVI header + raw data; uint8/int16/uint16/int32/uint32 are scaled to
// [0, 1] by their type's maximum, float32/float64 are taken as they are
// (envi.cpp:195-258). Any interleave is returned as BSQ. Throws on error.
HsiFrame read_envi(const std::string& hdr_path);
// Writes a float32 BSQ cube with its header.
void write_envi(const std::string& hdr_path, const HsiFrame& frame);
// "name,v1,v2,..." rows, '#' comments, optional "wavelength" row
// (spectra_io.cpp:27-91). Throws on error.
SpectralLibrary read_spectra_csv(const std::string& path);
void write_spectra_csv(const std::string& path, const SpectralLibrary& library);

// A trained learned pre-screen for RX/ACE (no counterpart in the reference):
// the network's packed FP32 weights for `bands` inputs (hsi_mlp_param_count)
// and the logit cut at or above which pixels are rescored. File layout,
// little-endian: "HSIMLP01", int32 bands, float32 cut, uint32 count,
// float32 weights[count] (hsi_mi355x.prescreen.Prescreen.save writes it).
struct PrescreenModel {
  int bands = 0;
  float cut = 0.f;
  std::vector<float> weights;
};
PrescreenModel read_prescreen(const std::string& path);

// ---- frame sources (frame_source.hpp) ------------------------------------------

struct FrameMeta {
  std::uint64_t index = 0;
  double timestamp_s = 0.0;
};

class FrameSource {
 public:
  virtual ~FrameSource() = default;
  virtual CubeShape shape() const = 0;
  // Distinct frames before the stream repeats or ends; 0 = unbounded.
  virtual std::uint64_t length() const = 0;
  // Write the next frame as BSQ with `plane_stride` elements between band
  // planes. False once the stream is exhausted.
  virtual bool read_into(float* dst, std::size_t plane_stride, FrameMeta* meta) = 0;
  virtual void reset() = 0;
  virtual std::string describe() const = 0;
  virtual const std::uint8_t* truth_mask() const { return nullptr; }
  virtual const SpectralLibrary* library() const { return nullptr; }
};

// Seeded scenes from include/hsi_synth.h: frame i is a pure function of
// (config, i). `frames` = 0 streams forever.
struct SyntheticOptions {
  hsi_synth_config config{};
  std::uint64_t frames = 0;
  int threads = 0;
  bool keep_truth = false;
};
std::unique_ptr<FrameSource> make_synthetic_source(const SyntheticOptions& options);

// Replays ENVI cubes (all of one shape) from memory, cycling when `loop`.
// At most `max_resident` of the headers (the first ones) are loaded, as the
// reference's EnviReplayOptions::max_resident (frame_source.hpp:60-70).
std::unique_ptr<FrameSource> make_envi_replay_source(const std::vector<std::string>& hdr_paths,
                                                     bool loop, int max_resident = 4);

// ---- pipeline (pipeline.hpp:101-120) ----------------------------------------

// The reference's SamVariant (sam_cuda.hpp:19-36) as the pipeline selects it:
// Baseline and Optimized both score the FP32 cube (this build has one FP32
// kernel per detector, the tuned one); Half narrows each uploaded frame to
// FP16 on the device and scores that, as pipeline.cu:290-296 does for SAM.
enum class SamVariant { Baseline, Optimized, Half };
bool parse_sam_variant(const std::string& name, SamVariant* out);
const char* to_string(SamVariant variant);

// Detection threshold per detector. SAM uses DetectionParams::threshold (an
// angle in radians, as in the reference). RX uses `rx` when it is a number,
// else the chi-squared threshold of the frame's band count at
// `rx_false_alarm_rate` (rx_threshold_for_false_alarm_rate,
// detectors_cpu.cpp:189-206). ACE and CEM responses are unit-scaled.
struct DetectorThresholds {
  float ace = 0.5f;
  float cem = 0.5f;
  float rx = std::numeric_limits<float>::quiet_NaN();
  double rx_false_alarm_rate = 1e-4;
};

struct PipelineOptions {
  std::vector<Detector> detectors{Detector::Sam};
  hsi_precision precision = HSI_PLAIN;  // RX/ACE arithmetic form
  SamVariant variant = SamVariant::Optimized;
  DetectionParams detection;            // SAM threshold, NMS, capacity; polarity is per detector
  DetectorThresholds thresholds;        // RX, ACE, CEM thresholds
  // RX/ACE through the learned pre-screen: one pass accumulates the
  // (mixed-form) statistics and screens every pixel with `prescreen`, and
  // only the pixels at or above its cut are scored, in `precision`; every
  // other pixel scores 0. At most 127 bands.
  bool use_prescreen = false;
  PrescreenModel prescreen;
  double loading = 1e-6;                // diagonal loading for the statistics
  // Frames in flight; each slot owns a stream, pinned staging, a device cube,
  // score maps, a detection list and a context.
  int num_streams = 3;
  // Non-zero: fill the staging ring once and replay it (separates GPU speed
  // from host frame production), as the reference's prefill.
  int prefill = 0;
  // Capture each slot's device work (upload, statistics, scoring, detect,
  // download) once as a hipGraph and replay it per frame.
  bool use_graph = true;
  bool time_stages = true;  // per-stage events (eager mode only)
  bool return_score_map = false;
  int device = 0;
};

struct DetectorResult {
  Detector detector = Detector::Sam;
  std::vector<Detection> detections;
  unsigned detections_found = 0;  // may exceed detections.size() if saturated
  std::vector<float> score_map;   // empty unless return_score_map
};

struct FrameResult {
  FrameMeta meta;
  std::vector<DetectorResult> results;  // one per PipelineOptions::detectors
  bool statistics_ok = true;            // false: a factorisation failed; RX/ACE/CEM lists empty
  float ms_source = 0;
  float ms_upload = 0;
  float ms_stats = 0;
  float ms_score = 0;
  float ms_detect = 0;
  float ms_download = 0;
  float ms_gpu = 0;   // upload through download, device time
  float ms_wall = 0;  // wall clock attributable to this frame
};

struct PipelineStats {
  std::uint64_t frames = 0;
  double wall_s = 0;
  double fps = 0;
  double cube_gb_per_s = 0;
  double mpixel_per_s = 0;
  float ms_p50 = 0, ms_p95 = 0, ms_p99 = 0, ms_max = 0;
  float mean_source = 0, mean_upload = 0, mean_stats = 0, mean_score = 0, mean_detect = 0,
        mean_download = 0, mean_gpu = 0;
  std::uint64_t total_detections = 0;
};

// The threshold `options` give detector `detector` on frames of `bands`
// bands; throws std::invalid_argument for an RX false-alarm rate outside
// (0, 1).
float detection_threshold(const PipelineOptions& options, Detector detector, int bands);

// ---- scenes of any shape in one batch (hsi_detect --batch) -------------------

struct BatchSceneResult {
  CubeShape shape;
  bool statistics_ok = true;            // false: this scene's factorisation failed
  std::vector<DetectorResult> results;  // one per PipelineOptions::detectors
};
struct BatchRun {
  std::vector<BatchSceneResult> scenes;
  float ms_batch = 0;                   // statistics + scoring of every scene, device time
  std::string autotune_report;
};
// Uploads every scene, runs their statistics and scoring as one hsi_batch
// (replayed from a captured graph when options.use_graph; shapes tuned first
// when autotune), then compacts each scene's maps with each detector's own
// threshold. Uses options' detectors, thresholds, detection parameters,
// loading, use_graph and device; plain precision. SAM, ACE and CEM need
// library.bands() == every scene's band count. Throws on error.
BatchRun run_scene_batch(const std::vector<HsiFrame>& scenes, const SpectralLibrary& library,
                         const PipelineOptions& options, bool autotune);

// Frame-rate detector over a FrameSource. Slots run ahead of each other and
// results come back in frame order.
class Pipeline {
 public:
  virtual ~Pipeline() = default;
  // Null (and `error` filled) on failure.
  static std::unique_ptr<Pipeline> create(FrameSource* source, const SpectralLibrary& library,
                                          const PipelineOptions& options, std::string* error);
  // Process `frames` frames (0: until the source ends), calling `on_frame`
  // once per frame in order. `on_frame` may be empty.
  virtual bool run(std::uint64_t frames, const std::function<void(const FrameResult&)>& on_frame,
                   std::string* error) = 0;
  virtual PipelineStats stats() const = 0;
  virtual void reset_stats() = 0;
  virtual std::string describe() const = 0;
};

}  // namespace hsi
