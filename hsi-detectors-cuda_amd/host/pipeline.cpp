// pipeline.cpp - frame-rate detection pipeline on HIP streams and graphs.
//
// Role of the reference's CudaPipeline (pipeline.cu:60-438): frames flow from
// a FrameSource through pinned staging to the device, are scored and
// compacted there, and come back as detection lists in frame order. The
// MI355X design differs in these ways:
//   - every slot owns a whole hsi_ctx (library, statistics, scratch), so any
//     of the four detectors runs per frame, RX/ACE/CEM with statistics
//     recomputed from that frame on the device (moments -> FP64 build);
//   - a slot's device work - upload, statistics, scoring, compaction,
//     download - is captured once as a hipGraph and replayed per frame, so a
//     frame costs one launch call on the host;
//   - SAM and CEM share one read of the cube (hsi_score_sam_cem) and ACE
//     produces the RX map as a side output when both are requested;
//   - each detector thresholds with its own value (DetectorThresholds; RX
//     from a false-alarm rate by default), where the reference pipeline only
//     ever thresholds SAM angles.
// With SamVariant::Half each uploaded frame is narrowed to FP16 on the device
// and every detector scores the FP16 cube, as pipeline.cu:290-296 does for
// SAM; the statistics still accumulate from the FP32 upload.
// Host code only: device work goes through the hsi_mi355x C ABI; this file
// uses the HIP runtime for streams, events, pinned memory and graphs.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>

#include "hsi_host.hpp"

namespace hsi {
namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw DeviceError(std::string(what) + ": " + hipGetErrorString(e));
}

void hsi_check(hsi_ctx* ctx, hsi_status st, const char* what) {
  if (st != HSI_OK)
    throw DeviceError(std::string(what) + ": " + hsi_status_name(st) + " (" + hsi_ctx_last_error(ctx) + ")");
}

enum Ev { kStart, kUploaded, kStats, kScored, kDetected, kEnd, kEvents };

struct Slot {
  hipStream_t stream = nullptr;
  hsi_ctx* ctx = nullptr;
  float* staging = nullptr;  // pinned, stride * bands
  float* d_cube = nullptr;
  void* d_half = nullptr;  // FP16 copy of the frame (SamVariant::Half)
  std::vector<float*> d_score;
  std::vector<std::int32_t*> d_target;
  hsi_detection* d_dets = nullptr;   // [detectors][capacity]
  std::uint32_t* d_counts = nullptr;  // [detectors]
  hsi_detection* h_dets = nullptr;    // pinned mirror
  hsi_detection* d_cand = nullptr;    // pre-screen candidates [screen_cap_]
  std::uint32_t* d_cand_count = nullptr;
  float* d_spectra = nullptr;         // [screen_cap_][screen_pitch_]
  std::uint32_t* h_counts = nullptr;  // pinned: [detectors] counts + 2 status words
  float* h_scores = nullptr;          // pinned [detectors][pixels] if score maps are returned
  hipEvent_t ev[kEvents] = {};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int uses = 0;
  bool busy = false;
  FrameResult pending;
  Clock::time_point t_begin;
};

class HipPipeline final : public Pipeline {
 public:
  HipPipeline(FrameSource* source, const PipelineOptions& options)
      : source_(source), opt_(options), shape_(source->shape()) {}

  // Allocation and checks; on a throw the destructor releases what exists.
  void init(const SpectralLibrary& library) {
    if (!shape_.valid()) throw std::invalid_argument("frame source has an empty shape");
    if (opt_.detectors.empty()) throw std::invalid_argument("no detector selected");
    if (opt_.num_streams < 1) throw std::invalid_argument("num_streams must be at least 1");
    if (opt_.detection.max_detections < 0 || opt_.detection.nms_radius < 0)
      throw std::invalid_argument("detection capacity and NMS radius must be non-negative");
    bool need_lib = false;
    for (Detector d : opt_.detectors) {
      need_lib |= traits_of(d).needs_targets;
      need_stats_ |= traits_of(d).needs_statistics;
    }
    if (need_lib && library.size() == 0) throw std::invalid_argument("the selected detectors need a target library");
    if (library.size() > 0 && library.bands() != shape_.bands)
      throw std::invalid_argument("library has " + std::to_string(library.bands()) + " bands but frames have " +
                                  std::to_string(shape_.bands));
    lib_values_ = library.flatten();
    lib_norms_ = library.norms();
    lib_targets_ = library.size();
    half_ = opt_.variant == SamVariant::Half;
    if (opt_.use_prescreen) {
      bool rxace = false;
      for (Detector d : opt_.detectors) rxace |= d == Detector::Rx || d == Detector::Ace;
      if (!rxace) throw std::invalid_argument("the learned pre-screen screens RX and ACE: select one of them");
      if (opt_.prescreen.bands != shape_.bands)
        throw std::invalid_argument("pre-screen model is for " + std::to_string(opt_.prescreen.bands) +
                                    " bands but frames have " + std::to_string(shape_.bands));
      if (shape_.bands > 127) throw std::invalid_argument("the learned pre-screen supports at most 127 bands");
      // A quarter of the frame as candidates (split evenly over the screening waves).
      screen_cap_ = std::max<std::size_t>(shape_.pixels() / 4, 4096);
      screen_pitch_ = static_cast<std::size_t>((shape_.bands + 31) / 32 * 32);
    }
    for (Detector d : opt_.detectors) thresholds_.push_back(detection_threshold(opt_, d, shape_.bands));
    stride_ = hsi_plane_stride(shape_.height, shape_.width);
    cap_ = std::max(opt_.detection.max_detections, 1);
    hip_check(hipSetDevice(opt_.device), "select device");
    hipDeviceProp_t prop;
    hip_check(hipGetDeviceProperties(&prop, opt_.device), "device properties");
    device_name_ = prop.name;
    slots_.resize(static_cast<std::size_t>(opt_.num_streams));
    for (Slot& s : slots_) init_slot(s);
    if (opt_.prefill > 0) {
      for (std::size_t i = 0; i < slots_.size(); ++i) {
        FrameMeta meta;
        const auto t0 = Clock::now();
        if (!source_->read_into(slots_[i].staging, stride_, &meta))
          throw std::runtime_error("frame source ended during prefill");
        prefill_meta_.push_back(meta);
        prefill_ms_.push_back(static_cast<float>(ms_between(t0, Clock::now())));
      }
    }
  }

  ~HipPipeline() override {
    (void)hipSetDevice(opt_.device);
    for (Slot& s : slots_) {
      if (s.stream) (void)hipStreamSynchronize(s.stream);
      if (s.exec) (void)hipGraphExecDestroy(s.exec);
      if (s.graph) (void)hipGraphDestroy(s.graph);
      for (hipEvent_t e : s.ev)
        if (e) (void)hipEventDestroy(e);
      for (float* p : s.d_score) (void)hipFree(p);
      for (std::int32_t* p : s.d_target) (void)hipFree(p);
      (void)hipFree(s.d_cube);
      if (s.d_half) (void)hipFree(s.d_half);
      (void)hipFree(s.d_dets);
      (void)hipFree(s.d_counts);
      if (s.d_cand) (void)hipFree(s.d_cand);
      if (s.d_cand_count) (void)hipFree(s.d_cand_count);
      if (s.d_spectra) (void)hipFree(s.d_spectra);
      (void)hipHostFree(s.staging);
      (void)hipHostFree(s.h_dets);
      (void)hipHostFree(s.h_counts);
      if (s.h_scores) (void)hipHostFree(s.h_scores);
      if (s.ctx) hsi_ctx_destroy(s.ctx);
      if (s.stream) (void)hipStreamDestroy(s.stream);
    }
  }

  bool run(std::uint64_t frames, const std::function<void(const FrameResult&)>& on_frame,
           std::string* error) override {
    try {
      const std::uint64_t limit = frames ? frames : (source_->length() ? source_->length() : 0);
      if (!limit) throw std::invalid_argument("an unbounded source (or a prefilled ring) needs a frame count");
      const auto wall0 = Clock::now();
      std::uint64_t issued = 0;
      const std::size_t n = slots_.size();
      for (; !limit || issued < limit; ++issued) {
        Slot& s = slots_[issue_ % n];
        if (s.busy) finish(s, on_frame);
        if (!produce(s)) break;
        enqueue(s);
        ++issue_;
      }
      // drain in frame order
      for (std::size_t k = 0; k < n; ++k) {
        Slot& s = slots_[(issue_ + k) % n];
        if (s.busy) finish(s, on_frame);
      }
      wall_s_ += std::chrono::duration<double>(Clock::now() - wall0).count();
      return true;
    } catch (const std::exception& e) {
      if (error) *error = e.what();
      return false;
    }
  }

  PipelineStats stats() const override {
    PipelineStats st;
    st.frames = results_;
    st.wall_s = wall_s_;
    if (wall_s_ > 0) {
      st.fps = static_cast<double>(results_) / wall_s_;
      st.mpixel_per_s = st.fps * static_cast<double>(shape_.pixels()) / 1e6;
      st.cube_gb_per_s = st.fps * static_cast<double>(shape_.elements()) * sizeof(float) / 1e9;
    }
    if (!lat_.empty()) {
      std::vector<float> v = lat_;
      std::sort(v.begin(), v.end());
      auto pct = [&](double q) { return v[static_cast<std::size_t>(q * (v.size() - 1) + 0.5)]; };
      st.ms_p50 = pct(0.50);
      st.ms_p95 = pct(0.95);
      st.ms_p99 = pct(0.99);
      st.ms_max = v.back();
    }
    const double k = results_ ? 1.0 / static_cast<double>(results_) : 0.0;
    st.mean_source = static_cast<float>(sum_.ms_source * k);
    st.mean_upload = static_cast<float>(sum_.ms_upload * k);
    st.mean_stats = static_cast<float>(sum_.ms_stats * k);
    st.mean_score = static_cast<float>(sum_.ms_score * k);
    st.mean_detect = static_cast<float>(sum_.ms_detect * k);
    st.mean_download = static_cast<float>(sum_.ms_download * k);
    st.mean_gpu = static_cast<float>(sum_.ms_gpu * k);
    st.total_detections = total_dets_;
    return st;
  }

  void reset_stats() override {
    results_ = 0;
    wall_s_ = 0;
    lat_.clear();
    sum_ = FrameResult{};
    total_dets_ = 0;
  }

  std::string describe() const override {
    std::string d = "hsi pipeline on " + device_name_ + ": " + std::to_string(shape_.height) + "x" +
                    std::to_string(shape_.width) + "x" + std::to_string(shape_.bands) + ", detectors ";
    for (std::size_t i = 0; i < opt_.detectors.size(); ++i) d += (i ? "," : "") + std::string(to_string(opt_.detectors[i]));
    d += ", " + std::to_string(lib_targets_) + " targets, " + std::to_string(slots_.size()) + " streams, ";
    d += opt_.use_graph ? "hipGraph frames" : "eager launches";
    d += opt_.precision == HSI_MIXED ? ", mixed precision" : ", plain precision";
    d += std::string(", ") + (half_ ? "FP16" : "FP32") + " cube (variant " + to_string(opt_.variant) + ")";
    d += ", thresholds";
    for (std::size_t i = 0; i < opt_.detectors.size(); ++i) {
      char buf[64];
      std::snprintf(buf, sizeof buf, " %s=%.6g", to_string(opt_.detectors[i]), static_cast<double>(thresholds_[i]));
      d += buf;
    }
    if (opt_.use_prescreen) {
      char buf[64];
      std::snprintf(buf, sizeof buf, ", learned pre-screen for RX/ACE (cut %.6g)", static_cast<double>(opt_.prescreen.cut));
      d += buf;
    }
    if (opt_.prefill) d += ", prefilled ring";
    d += "; source: " + source_->describe();
    return d;
  }

 private:
  void init_slot(Slot& s) {
    const std::size_t P = shape_.pixels();
    const std::size_t bytes = stride_ * shape_.bands * sizeof(float);
    const std::size_t nd = opt_.detectors.size();
    hip_check(hipStreamCreateWithFlags(&s.stream, hipStreamNonBlocking), "create stream");
    for (hipEvent_t& e : s.ev) hip_check(hipEventCreate(&e), "create event");
    if (hsi_ctx_create(opt_.device, &s.ctx) != HSI_OK) throw DeviceError("cannot create a context");
    if (lib_targets_ > 0)
      hsi_check(s.ctx, hsi_set_library(s.ctx, lib_values_.data(), lib_norms_.data(), lib_targets_, shape_.bands),
                "install library");
    hsi_check(s.ctx, hsi_ctx_reserve(s.ctx, P, shape_.bands), "reserve scratch");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.staging), bytes, hipHostMallocDefault), "pinned staging");
    std::memset(s.staging, 0, bytes);
    hip_check(hipMalloc(reinterpret_cast<void**>(&s.d_cube), bytes), "device cube");
    if (half_) hip_check(hipMalloc(&s.d_half, stride_ * shape_.bands * sizeof(std::uint16_t)), "device FP16 cube");
    for (std::size_t i = 0; i < nd; ++i) {
      float* sc = nullptr;
      std::int32_t* tg = nullptr;
      hip_check(hipMalloc(reinterpret_cast<void**>(&sc), P * sizeof(float)), "score map");
      hip_check(hipMalloc(reinterpret_cast<void**>(&tg), P * sizeof(std::int32_t)), "target map");
      s.d_score.push_back(sc);
      s.d_target.push_back(tg);
    }
    if (opt_.use_prescreen) {
      hsi_check(s.ctx, hsi_mlp_set_weights(s.ctx, shape_.bands, opt_.prescreen.weights.data()), "install pre-screen");
      hip_check(hipMalloc(reinterpret_cast<void**>(&s.d_cand), screen_cap_ * sizeof(hsi_detection)), "candidates");
      hip_check(hipMalloc(reinterpret_cast<void**>(&s.d_cand_count), sizeof(std::uint32_t)), "candidate count");
      hip_check(hipMalloc(reinterpret_cast<void**>(&s.d_spectra), screen_cap_ * screen_pitch_ * sizeof(float)),
                "candidate spectra");
    }
    hip_check(hipMalloc(reinterpret_cast<void**>(&s.d_dets), nd * cap_ * sizeof(hsi_detection)), "detections");
    hip_check(hipMalloc(reinterpret_cast<void**>(&s.d_counts), nd * sizeof(std::uint32_t)), "counts");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_dets), nd * cap_ * sizeof(hsi_detection), hipHostMallocDefault),
              "pinned detections");
    hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_counts), (nd + 2) * sizeof(std::uint32_t), hipHostMallocDefault),
              "pinned counts");
    std::memset(s.h_counts, 0, (nd + 2) * sizeof(std::uint32_t));
    if (opt_.return_score_map)
      hip_check(hipHostMalloc(reinterpret_cast<void**>(&s.h_scores), nd * P * sizeof(float), hipHostMallocDefault),
                "pinned score maps");
  }

  int index_of(Detector d) const {
    for (std::size_t i = 0; i < opt_.detectors.size(); ++i)
      if (opt_.detectors[i] == d) return static_cast<int>(i);
    return -1;
  }

  bool produce(Slot& s) {
    s.pending = FrameResult{};
    s.t_begin = Clock::now();
    if (opt_.prefill > 0) {
      const std::size_t k = issue_ % slots_.size();
      s.pending.meta = prefill_meta_[k];
      s.pending.meta.index = issue_;
      return true;
    }
    FrameMeta meta;
    if (!source_->read_into(s.staging, stride_, &meta)) return false;
    s.pending.meta = meta;
    s.pending.ms_source = static_cast<float>(ms_between(s.t_begin, Clock::now()));
    return true;
  }

  // The device work of one frame, in stream order on s.stream.
  void device_work(Slot& s, bool events) {
    const int H = shape_.height, W = shape_.width, B = shape_.bands;
    const std::size_t P = shape_.pixels();
    const std::size_t nd = opt_.detectors.size();
    hsi_stream st = s.stream;
    if (events) hip_check(hipEventRecord(s.ev[kStart], s.stream), "event");
    hip_check(hipMemcpyAsync(s.d_cube, s.staging, stride_ * B * sizeof(float), hipMemcpyHostToDevice, s.stream),
              "upload frame");
    if (half_)
      hsi_check(s.ctx, hsi_narrow_to_half(s.d_cube, s.d_half, stride_ * static_cast<std::size_t>(B), st),
                "narrow to FP16");
    const void* cube = half_ ? s.d_half : static_cast<const void*>(s.d_cube);
    const hsi_dtype dtype = half_ ? HSI_F16 : HSI_F32;
    if (events) hip_check(hipEventRecord(s.ev[kUploaded], s.stream), "event");
    const bool screen = opt_.use_prescreen;
    if (need_stats_) {
      if (screen)
        hsi_check(s.ctx,
                  hsi_stats_accumulate_screen(s.ctx, s.d_cube, H, W, B, stride_, nullptr, opt_.prescreen.cut, s.d_cand,
                                              s.d_cand_count, static_cast<int>(screen_cap_), s.d_spectra,
                                              screen_pitch_, st),
                  "moments and screening");
      else
        hsi_check(s.ctx, hsi_stats_accumulate(s.ctx, s.d_cube, HSI_F32, H, W, B, stride_, opt_.precision, nullptr, st),
                  "moments");
      hsi_check(s.ctx, hsi_stats_build(s.ctx, opt_.loading, st), "statistics");
      hsi_check(s.ctx, hsi_stats_status_copy(s.ctx, s.h_counts + nd, st), "statistics status");
    }
    if (events) hip_check(hipEventRecord(s.ev[kStats], s.stream), "event");
    const int i_sam = index_of(Detector::Sam), i_cem = index_of(Detector::Cem);
    const int i_rx = index_of(Detector::Rx), i_ace = index_of(Detector::Ace);
    if (i_sam >= 0 || i_cem >= 0)
      hsi_check(s.ctx,
                hsi_score_sam_cem(s.ctx, cube, dtype, H, W, B, stride_, i_sam >= 0 ? s.d_score[i_sam] : nullptr,
                                  i_sam >= 0 ? s.d_target[i_sam] : nullptr, i_cem >= 0 ? s.d_score[i_cem] : nullptr,
                                  i_cem >= 0 ? s.d_target[i_cem] : nullptr, st),
                "score SAM/CEM");
    if (screen && (i_ace >= 0 || i_rx >= 0)) {
      const hsi_detector det = i_ace >= 0 ? HSI_DET_ACE : HSI_DET_RX;
      float* score = det == HSI_DET_ACE ? s.d_score[i_ace] : s.d_score[i_rx];
      hsi_check(s.ctx,
                hsi_score_candidate_spectra(s.ctx, det, opt_.precision, s.d_cube, H, W, B, stride_, s.d_cand,
                                            s.d_cand_count, static_cast<int>(screen_cap_), s.d_spectra,
                                            screen_pitch_, score, det == HSI_DET_ACE ? s.d_target[i_ace] : nullptr,
                                            det == HSI_DET_ACE && i_rx >= 0 ? s.d_score[i_rx] : nullptr, st),
                "rescore candidates");
    } else if (i_ace >= 0)
      hsi_check(s.ctx,
                hsi_score(s.ctx, HSI_DET_ACE, opt_.precision, cube, dtype, H, W, B, stride_, s.d_score[i_ace],
                          s.d_target[i_ace], i_rx >= 0 ? s.d_score[i_rx] : nullptr, st),
                "score ACE");
    else if (i_rx >= 0)
      hsi_check(s.ctx,
                hsi_score(s.ctx, HSI_DET_RX, opt_.precision, cube, dtype, H, W, B, stride_, s.d_score[i_rx],
                          nullptr, nullptr, st),
                "score RX");
    if (events) hip_check(hipEventRecord(s.ev[kScored], s.stream), "event");
    for (std::size_t i = 0; i < nd; ++i) {
      const Detector d = opt_.detectors[i];
      hsi_detection_params p;
      p.threshold = thresholds_[i];
      p.nms_radius = opt_.detection.nms_radius;
      p.max_detections = opt_.detection.max_detections;
      p.higher_is_better = traits_of(d).polarity == ScorePolarity::HigherIsBetter ? 1 : 0;
      hsi_check(s.ctx,
                hsi_detect(s.ctx, s.d_score[i], d == Detector::Rx ? nullptr : s.d_target[i], H, W, p,
                           s.d_dets + i * cap_, s.d_counts + i, st),
                "detect");
    }
    if (events) hip_check(hipEventRecord(s.ev[kDetected], s.stream), "event");
    hip_check(hipMemcpyAsync(s.h_counts, s.d_counts, nd * sizeof(std::uint32_t), hipMemcpyDeviceToHost, s.stream),
              "download counts");
    hip_check(hipMemcpyAsync(s.h_dets, s.d_dets, nd * cap_ * sizeof(hsi_detection), hipMemcpyDeviceToHost, s.stream),
              "download detections");
    if (opt_.return_score_map)
      for (std::size_t i = 0; i < nd; ++i)
        hip_check(hipMemcpyAsync(s.h_scores + i * P, s.d_score[i], P * sizeof(float), hipMemcpyDeviceToHost, s.stream),
                  "download score map");
    if (events) hip_check(hipEventRecord(s.ev[kEnd], s.stream), "event");
  }

  void enqueue(Slot& s) {
    hip_check(hipSetDevice(opt_.device), "select device");
    // A slot's first frame runs eagerly (code objects load, layouts settle);
    // from the second frame on, its work is one graph launch.
    if (opt_.use_graph && s.uses > 0) {
      if (!s.exec) {
        hip_check(hipStreamBeginCapture(s.stream, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
          device_work(s, false);
        } catch (...) {
          hipGraph_t g = nullptr;
          (void)hipStreamEndCapture(s.stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(s.stream, &s.graph), "end capture");
        hip_check(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0), "instantiate graph");
      }
      hip_check(hipEventRecord(s.ev[kStart], s.stream), "event");
      hip_check(hipGraphLaunch(s.exec, s.stream), "launch graph");
      hip_check(hipEventRecord(s.ev[kEnd], s.stream), "event");
    } else {
      device_work(s, true);
    }
    s.busy = true;
    ++s.uses;
  }

  void finish(Slot& s, const std::function<void(const FrameResult&)>& on_frame) {
    hip_check(hipStreamSynchronize(s.stream), "wait for frame");
    s.busy = false;
    FrameResult& r = s.pending;
    const std::size_t nd = opt_.detectors.size();
    const std::size_t P = shape_.pixels();
    const bool graphed = opt_.use_graph && s.exec && s.uses > 1;
    auto span = [&](int a, int b) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, s.ev[a], s.ev[b]);
      return ms;
    };
    r.ms_gpu = span(kStart, kEnd);
    if (!graphed && opt_.time_stages) {
      r.ms_upload = span(kStart, kUploaded);
      r.ms_stats = span(kUploaded, kStats);
      r.ms_score = span(kStats, kScored);
      r.ms_detect = span(kScored, kDetected);
      r.ms_download = span(kDetected, kEnd);
    }
    r.statistics_ok = !need_stats_ || (s.h_counts[nd] == 0 && s.h_counts[nd + 1] == 0);
    r.results.resize(nd);
    for (std::size_t i = 0; i < nd; ++i) {
      DetectorResult& dr = r.results[i];
      dr.detector = opt_.detectors[i];
      if (traits_of(dr.detector).needs_statistics && !r.statistics_ok) continue;
      dr.detections_found = s.h_counts[i];
      const std::size_t keep = std::min<std::size_t>(dr.detections_found, static_cast<std::size_t>(opt_.detection.max_detections));
      dr.detections.resize(keep);
      std::memcpy(static_cast<void*>(dr.detections.data()), s.h_dets + i * cap_, keep * sizeof(hsi_detection));
      if (opt_.return_score_map) dr.score_map.assign(s.h_scores + i * P, s.h_scores + (i + 1) * P);
      total_dets_ += keep;
    }
    r.ms_wall = static_cast<float>(ms_between(s.t_begin, Clock::now()));
    lat_.push_back(r.ms_wall);
    sum_.ms_source += r.ms_source;
    sum_.ms_upload += r.ms_upload;
    sum_.ms_stats += r.ms_stats;
    sum_.ms_score += r.ms_score;
    sum_.ms_detect += r.ms_detect;
    sum_.ms_download += r.ms_download;
    sum_.ms_gpu += r.ms_gpu;
    ++results_;
    if (on_frame) on_frame(r);
  }

  FrameSource* source_;
  PipelineOptions opt_;
  CubeShape shape_;
  std::size_t stride_ = 0;
  std::size_t cap_ = 1;
  bool need_stats_ = false;
  bool half_ = false;
  std::size_t screen_cap_ = 0;     // pre-screen candidates per frame
  std::size_t screen_pitch_ = 0;   // floats per candidate spectrum
  std::vector<float> thresholds_;  // per detector, in PipelineOptions::detectors order
  std::vector<float> lib_values_, lib_norms_;
  int lib_targets_ = 0;
  std::string device_name_;
  std::vector<Slot> slots_;
  std::vector<FrameMeta> prefill_meta_;
  std::vector<float> prefill_ms_;
  std::uint64_t issue_ = 0;
  std::uint64_t results_ = 0;
  double wall_s_ = 0;
  std::vector<float> lat_;
  FrameResult sum_;
  std::uint64_t total_dets_ = 0;
};

}  // namespace

std::unique_ptr<Pipeline> Pipeline::create(FrameSource* source, const SpectralLibrary& library,
                                           const PipelineOptions& options, std::string* error) {
  try {
    if (!source) throw std::invalid_argument("null frame source");
    std::unique_ptr<HipPipeline> p(new HipPipeline(source, options));
    p->init(library);
    return p;
  } catch (const std::exception& e) {
    if (error) *error = e.what();
    return nullptr;
  }
}

}  // namespace hsi
