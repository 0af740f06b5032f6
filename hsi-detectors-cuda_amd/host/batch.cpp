// batch.cpp - scenes of any shape scored together (hsi_detect --batch).
//
// No counterpart in the reference, whose hsi_detect scores one stream of
// same-shape frames: here every scene (its own size and band count) is
// uploaded once, the statistics and scoring of all of them run as one
// hsi_batch (moments per scene, every FP64 build in one launch per build
// form, each detector's scenes in one persistent launch per shape class),
// replayed from a captured hipGraph, and each scene is then compacted with
// its own thresholds by hsi_detect on the scene's context.
// Host code only: device work goes through the hsi_mi355x C ABI; this file
// uses the HIP runtime for memory, a stream and events.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "hsi_host.hpp"

namespace hsi {
namespace {

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// Device buffers of one scene, released with the run.
struct SceneMem {
  float* cube = nullptr;
  float* score[4] = {};
  std::int32_t* target[4] = {};
  hsi_detection* dets = nullptr;
  std::uint32_t* count = nullptr;
};

}  // namespace

BatchRun run_scene_batch(const std::vector<HsiFrame>& scenes, const SpectralLibrary& library,
                         const PipelineOptions& opt, bool autotune) {
  if (scenes.empty()) throw std::invalid_argument("a batch needs at least one scene");
  if (opt.detectors.empty()) throw std::invalid_argument("no detector selected");
  bool need_lib = false;
  for (Detector d : opt.detectors) need_lib |= traits_of(d).needs_targets;
  for (const HsiFrame& f : scenes) {
    if (!f.shape.valid()) throw std::invalid_argument("a batch scene has an empty shape");
    if (need_lib && (library.size() == 0 || library.bands() != f.shape.bands))
      throw std::invalid_argument("the library has " + std::to_string(library.bands()) + " bands, a scene has " +
                                  std::to_string(f.shape.bands) + ": SAM, ACE and CEM need one library per band count");
  }
  const std::vector<float> lib = library.flatten();
  const int cap = std::max(opt.detection.max_detections, 1);
  hip_ok(hipSetDevice(opt.device), "select device");
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hsi_batch* batch = nullptr;
  std::vector<SceneMem> mem(scenes.size());
  BatchRun out;
  auto release = [&]() {
    if (stream) (void)hipStreamSynchronize(stream);
    if (batch) hsi_batch_destroy(batch);
    for (SceneMem& m : mem) {
      (void)hipFree(m.cube);
      for (int i = 0; i < 4; ++i) {
        (void)hipFree(m.score[i]);
        (void)hipFree(m.target[i]);
      }
      (void)hipFree(m.dets);
      (void)hipFree(m.count);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (stream) (void)hipStreamDestroy(stream);
  };
  try {
    hip_ok(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "create stream");
    hip_ok(hipEventCreate(&e0), "create event");
    hip_ok(hipEventCreate(&e1), "create event");
    std::vector<hsi_scene> desc(scenes.size());
    const std::size_t nd = opt.detectors.size();
    for (std::size_t s = 0; s < scenes.size(); ++s) {
      const HsiFrame& f = scenes[s];
      const std::size_t P = f.shape.pixels(), B = static_cast<std::size_t>(f.shape.bands);
      const std::size_t stride = hsi_plane_stride(f.shape.height, f.shape.width);
      SceneMem& m = mem[s];
      hip_ok(hipMalloc(reinterpret_cast<void**>(&m.cube), stride * B * sizeof(float)), "scene cube");
      hip_ok(hipMemset(m.cube, 0, stride * B * sizeof(float)), "clear scene cube");
      hip_ok(hipMemcpy2D(m.cube, stride * sizeof(float), f.data.data(), P * sizeof(float), P * sizeof(float), B,
                         hipMemcpyHostToDevice),
             "upload scene");
      hsi_scene& d = desc[s];
      d = hsi_scene{};
      d.d_cube = m.cube;
      d.height = f.shape.height;
      d.width = f.shape.width;
      d.bands = f.shape.bands;
      d.plane_stride = stride;
      d.library = library.size() > 0 && library.bands() == f.shape.bands ? lib.data() : nullptr;
      d.num_targets = d.library ? static_cast<int>(library.size()) : 0;
      for (std::size_t i = 0; i < nd; ++i) {
        hip_ok(hipMalloc(reinterpret_cast<void**>(&m.score[i]), P * sizeof(float)), "score map");
        hip_ok(hipMalloc(reinterpret_cast<void**>(&m.target[i]), P * sizeof(std::int32_t)), "target map");
        switch (opt.detectors[i]) {
          case Detector::Sam: d.d_sam_score = m.score[i]; d.d_sam_target = m.target[i]; break;
          case Detector::Cem: d.d_cem_score = m.score[i]; d.d_cem_target = m.target[i]; break;
          case Detector::Rx: d.d_rx_score = m.score[i]; break;
          case Detector::Ace: d.d_ace_score = m.score[i]; d.d_ace_target = m.target[i]; break;
        }
      }
      hip_ok(hipMalloc(reinterpret_cast<void**>(&m.dets), nd * cap * sizeof(hsi_detection)), "detections");
      hip_ok(hipMalloc(reinterpret_cast<void**>(&m.count), nd * sizeof(std::uint32_t)), "counts");
    }
    if (hsi_batch_create(opt.device, &batch) != HSI_OK) throw std::runtime_error("cannot create a batch");
    auto bcheck = [&](hsi_status st, const char* what) {
      if (st != HSI_OK)
        throw std::runtime_error(std::string(what) + ": " + hsi_status_name(st) + " (" + hsi_batch_last_error(batch) + ")");
    };
    bcheck(hsi_batch_set_scenes(batch, desc.data(), static_cast<int>(desc.size())), "install scenes");
    if (autotune) {
      // The tuner times the scoring launches, which read the statistics: build them first.
      bcheck(hsi_batch_statistics(batch, opt.loading, stream), "batch statistics");
      char report[8192] = {};
      bcheck(hsi_batch_autotune(batch, stream, report, sizeof report), "autotune");
      out.autotune_report = report;
    }
    // The first run captures the graph (with use_graph); the second is timed.
    bcheck(hsi_batch_run(batch, opt.loading, opt.use_graph ? 1 : 0, stream), "batch run");
    hip_ok(hipEventRecord(e0, stream), "event");
    bcheck(hsi_batch_run(batch, opt.loading, opt.use_graph ? 1 : 0, stream), "batch run");
    hip_ok(hipEventRecord(e1, stream), "event");
    hip_ok(hipEventSynchronize(e1), "wait for the batch");
    float ms = 0.f;
    hip_ok(hipEventElapsedTime(&ms, e0, e1), "batch time");
    out.ms_batch = ms;
    int failed = -1;
    const bool stats_ok = hsi_batch_check(batch, &failed) == HSI_OK;
    out.scenes.resize(scenes.size());
    for (std::size_t s = 0; s < scenes.size(); ++s) {
      const HsiFrame& f = scenes[s];
      SceneMem& m = mem[s];
      hsi_ctx* ctx = hsi_batch_context(batch, static_cast<int>(s));
      BatchSceneResult& r = out.scenes[s];
      r.shape = f.shape;
      r.statistics_ok = stats_ok || static_cast<int>(s) != failed;
      for (std::size_t i = 0; i < nd; ++i) {
        const Detector det = opt.detectors[i];
        hsi_detection_params p;
        p.threshold = detection_threshold(opt, det, f.shape.bands);
        p.nms_radius = opt.detection.nms_radius;
        p.max_detections = cap;
        p.higher_is_better = traits_of(det).polarity == ScorePolarity::HigherIsBetter ? 1 : 0;
        const hsi_status st = hsi_detect(ctx, m.score[i], det == Detector::Rx ? nullptr : m.target[i],
                                         f.shape.height, f.shape.width, p, m.dets + i * cap, m.count + i, stream);
        if (st != HSI_OK)
          throw std::runtime_error(std::string("detect: ") + hsi_status_name(st) + " (" + hsi_ctx_last_error(ctx) + ")");
      }
      std::vector<std::uint32_t> counts(nd);
      std::vector<hsi_detection> dets(nd * cap);
      hip_ok(hipMemcpyAsync(counts.data(), m.count, nd * sizeof(std::uint32_t), hipMemcpyDeviceToHost, stream),
             "download counts");
      hip_ok(hipMemcpyAsync(dets.data(), m.dets, nd * cap * sizeof(hsi_detection), hipMemcpyDeviceToHost, stream),
             "download detections");
      hip_ok(hipStreamSynchronize(stream), "wait for detections");
      for (std::size_t i = 0; i < nd; ++i) {
        DetectorResult dr;
        dr.detector = opt.detectors[i];
        dr.detections_found = counts[i];
        const std::size_t n = std::min<std::size_t>(counts[i], static_cast<std::size_t>(cap));
        for (std::size_t k = 0; k < n; ++k) {
          const hsi_detection& d = dets[i * cap + k];
          dr.detections.push_back(Detection{d.x, d.y, d.target, d.score});
        }
        r.results.push_back(std::move(dr));
      }
    }
  } catch (...) {
    release();
    throw;
  }
  release();
  return out;
}

}  // namespace hsi
