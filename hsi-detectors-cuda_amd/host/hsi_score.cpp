// hsi_score - score one hyperspectral cube against a target library.
//
// Counterpart of the reference's hsi_score (score_main.cpp:66-175), which
// runs the CPU SAM scorer over an ENVI cube and a CSV library and dumps the
// raw maps. This one runs any of the four detectors on the GPU through the
// C ABI, with the same flags and the same raw f32 / i32 dump formats, so its
// output diffs directly against a reference run on the same files.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "hsi_host.hpp"

namespace {

void usage() {
  std::puts(R"(hsi_score - score one hyperspectral cube on an MI355X

  hsi_score --cube=FILE.hdr --library=FILE.csv [options]

  --cube=FILE.hdr        ENVI cube to score
  --library=FILE.csv     target spectra (not needed for --detector=rx)
  --detector=NAME        sam, rx, ace or cem                          [sam]
  --target=NAME          score against this one target only
  --precision=plain|mixed  RX/ACE arithmetic form                   [plain]
  --loading=F            diagonal loading of the statistics          [1e-6]
  --threshold=F          detection threshold of the selected detector
                         [SAM 0.10 rad; RX from --rx-false-alarm; ACE, CEM 0.5]
  --rx-false-alarm=RATE  RX threshold as a chi-squared false-alarm rate [1e-4]
  --nms=N                non-maximum suppression radius, 0 off          [2]
  --max-detections=N     detection list capacity                     [4096]
  --dump-score=FILE      write the score map as raw float32
  --dump-target=FILE     write the winning target index as raw int32
  --print=N              print the first N detections                  [10]
  --device=N             HIP device                                     [0]
  --help)");
}

// Every flag this program reads; anything else is refused (exit 2).
const std::set<std::string> kFlags = {"cube", "library", "detector", "target", "precision", "loading",
                                      "threshold", "rx-false-alarm", "nms", "max-detections", "dump-score", "dump-target",
                                      "print", "device", "help", "h"};

std::map<std::string, std::string> parse(int argc, char** argv, std::string* bad) {
  std::map<std::string, std::string> a;
  for (int i = 1; i < argc; ++i) {
    std::string t = argv[i];
    if (t.rfind("--", 0) != 0) {
      if (bad->empty()) *bad = "unexpected argument '" + t + "'";
      continue;
    }
    t = t.substr(2);
    const std::size_t eq = t.find('=');
    const std::string key = eq == std::string::npos ? t : t.substr(0, eq);
    if (!kFlags.count(key) && bad->empty()) *bad = "unknown option --" + key;
    if (eq == std::string::npos) a[key] = "";
    else a[key] = t.substr(eq + 1);
  }
  return a;
}

struct Fail : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Fail(std::string(what) + ": " + hipGetErrorString(e));
}

void hsi_ok(hsi_ctx* ctx, hsi_status st, const char* what) {
  if (st != HSI_OK) throw Fail(std::string(what) + ": " + hsi_ctx_last_error(ctx));
}

template <typename T>
void dump(const std::string& path, const std::vector<T>& v) {
  std::ofstream out(path, std::ios::binary);
  if (!out) throw Fail("cannot write " + path);
  out.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

}  // namespace

int main(int argc, char** argv) {
  std::string bad;
  const auto args = parse(argc, argv, &bad);
  if (!bad.empty() && !args.count("help") && !args.count("h")) {
    std::fprintf(stderr, "hsi_score: %s (see --help)\n", bad.c_str());
    return 2;
  }
  auto get = [&](const char* k, const std::string& d) {
    auto it = args.find(k);
    return it == args.end() || it->second.empty() ? d : it->second;
  };
  hsi::Detector det = hsi::Detector::Sam;
  if (!hsi::parse_detector(get("detector", "sam"), &det)) {
    std::fputs("--detector must be sam, rx, ace or cem\n", stderr);
    return 2;
  }
  const bool need_lib = hsi::traits_of(det).needs_targets;
  if (args.count("help") || args.count("h") || !args.count("cube") || (need_lib && !args.count("library"))) {
    usage();
    return args.count("help") || args.count("h") ? 0 : 2;
  }
  hsi_ctx* ctx = nullptr;
  try {
    const hsi::HsiFrame frame = hsi::read_envi(get("cube", ""));
    hsi::SpectralLibrary library;
    if (args.count("library")) library = hsi::read_spectra_csv(get("library", ""));
    if (args.count("target")) {
      const std::string want = get("target", "");
      hsi::SpectralLibrary one;
      for (const auto& t : library.targets)
        if (t.name == want) one.targets.push_back(t);
      if (one.targets.empty()) throw Fail("no target named '" + want + "' in the library");
      library = one;
    }
    const hsi::CubeShape shape = frame.shape;
    if (library.size() > 0 && library.bands() != shape.bands)
      throw Fail("library has " + std::to_string(library.bands()) + " bands but the cube has " +
                 std::to_string(shape.bands));
    const int device = std::stoi(get("device", "0"));
    if (hsi_ctx_create(device, &ctx) != HSI_OK) throw Fail("cannot create a context on the device");
    const std::size_t P = shape.pixels();
    const std::size_t stride = hsi_plane_stride(shape.height, shape.width);
    float* d_cube = nullptr;
    float *d_score = nullptr, *d_aux = nullptr;
    std::int32_t* d_target = nullptr;
    hsi_detection* d_dets = nullptr;
    std::uint32_t* d_count = nullptr;
    const int cap = std::stoi(get("max-detections", "4096"));
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_cube), stride * shape.bands * sizeof(float)), "device cube");
    hip_ok(hipMemset(d_cube, 0, stride * shape.bands * sizeof(float)), "clear cube");
    hip_ok(hipMemcpy2D(d_cube, stride * sizeof(float), frame.data.data(), P * sizeof(float), P * sizeof(float),
                       static_cast<std::size_t>(shape.bands), hipMemcpyHostToDevice),
           "upload cube");
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_score), P * sizeof(float)), "score map");
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_aux), P * sizeof(float)), "aux map");
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_target), P * sizeof(std::int32_t)), "target map");
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_dets), static_cast<std::size_t>(cap > 0 ? cap : 1) * sizeof(hsi_detection)),
           "detections");
    hip_ok(hipMalloc(reinterpret_cast<void**>(&d_count), sizeof(std::uint32_t)), "count");
    hip_ok(hipMemset(d_target, 0, P * sizeof(std::int32_t)), "clear targets");
    if (library.size() > 0) {
      const std::vector<float> values = library.flatten();
      const std::vector<float> norms = library.norms();
      hsi_ok(ctx, hsi_set_library(ctx, values.data(), norms.data(), library.size(), library.bands()), "library");
    }
    const hsi_precision prec = get("precision", "plain") == "mixed" ? HSI_MIXED : HSI_PLAIN;
    if (hsi::traits_of(det).needs_statistics) {
      hsi_ok(ctx, hsi_stats_accumulate(ctx, d_cube, HSI_F32, shape.height, shape.width, shape.bands, stride, prec,
                                       nullptr, nullptr),
             "moments");
      hsi_ok(ctx, hsi_stats_build(ctx, std::stod(get("loading", "1e-6")), nullptr), "statistics");
      hsi_ok(ctx, hsi_stats_check(ctx, nullptr), "statistics");
    }
    const hsi_detector hd = static_cast<hsi_detector>(static_cast<int>(det));
    hsi_ok(ctx,
           hsi_score(ctx, hd, det == hsi::Detector::Rx || det == hsi::Detector::Ace ? prec : HSI_PLAIN, d_cube, HSI_F32,
                     shape.height, shape.width, shape.bands, stride, d_score,
                     det == hsi::Detector::Rx ? nullptr : d_target, det == hsi::Detector::Ace ? d_aux : nullptr,
                     nullptr),
           "score");
    hsi_detection_params p;
    // --threshold applies to the selected detector; without it each detector
    // takes the pipeline's default (SAM 0.10 rad, RX from --rx-false-alarm,
    // ACE and CEM 0.5).
    hsi::PipelineOptions defaults;
    defaults.thresholds.rx_false_alarm_rate = std::stod(get("rx-false-alarm", "1e-4"));
    p.threshold = args.count("threshold") ? std::stof(get("threshold", "0.10"))
                                          : hsi::detection_threshold(defaults, det, shape.bands);
    p.nms_radius = std::stoi(get("nms", "2"));
    p.max_detections = cap;
    p.higher_is_better = hsi::traits_of(det).polarity == hsi::ScorePolarity::HigherIsBetter ? 1 : 0;
    hsi_ok(ctx,
           hsi_detect(ctx, d_score, det == hsi::Detector::Rx ? nullptr : d_target, shape.height, shape.width, p,
                      d_dets, d_count, nullptr),
           "detect");
    hip_ok(hipDeviceSynchronize(), "score");
    std::vector<float> score(P);
    std::vector<std::int32_t> target(P);
    std::uint32_t found = 0;
    hip_ok(hipMemcpy(score.data(), d_score, P * sizeof(float), hipMemcpyDeviceToHost), "download scores");
    hip_ok(hipMemcpy(target.data(), d_target, P * sizeof(std::int32_t), hipMemcpyDeviceToHost), "download targets");
    hip_ok(hipMemcpy(&found, d_count, sizeof(found), hipMemcpyDeviceToHost), "download count");
    const std::size_t keep = std::min<std::size_t>(found, static_cast<std::size_t>(cap > 0 ? cap : 0));
    std::vector<hsi_detection> dets(keep);
    if (keep) hip_ok(hipMemcpy(dets.data(), d_dets, keep * sizeof(hsi_detection), hipMemcpyDeviceToHost), "download detections");
    std::printf("%s: %dx%dx%d, %d targets, %u detections\n", hsi::to_string(det), shape.height, shape.width,
                shape.bands, library.size(), found);
    const int show = std::stoi(get("print", "10"));
    for (int i = 0; i < show && i < static_cast<int>(keep); ++i) {
      const hsi_detection& d = dets[static_cast<std::size_t>(i)];
      const std::string name = d.target >= 0 && d.target < library.size() ? library.targets[d.target].name : "-";
      std::printf("  x=%d y=%d target=%s score=%.6g\n", d.x, d.y, name.c_str(), d.score);
    }
    if (args.count("dump-score")) dump(get("dump-score", ""), score);
    if (args.count("dump-target")) dump(get("dump-target", ""), target);
    (void)hipFree(d_cube);
    (void)hipFree(d_score);
    (void)hipFree(d_aux);
    (void)hipFree(d_target);
    (void)hipFree(d_dets);
    (void)hipFree(d_count);
    hsi_ctx_destroy(ctx);
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "hsi_score: %s\n", e.what());
    if (ctx) hsi_ctx_destroy(ctx);
    return 1;
  }
}
