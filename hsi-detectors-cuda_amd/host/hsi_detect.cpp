// hsi_detect - frame-rate target detection on one MI355X.
//
// Counterpart of the reference's hsi_detect (main.cpp:125-292): frames from a
// synthetic or ENVI-replay source go through the Pipeline, and the run reports
// throughput, latency percentiles and per-stage device times. Any of the four
// detectors (or several) score each frame, each against its own threshold.
// The reference's flags keep their names and meanings (--source, --envi,
// --resident, --width/--height/--bands, --targets, --noise, --library,
// --threshold, --nms, --max-detections, --variant, --memory, --streams,
// --frames, --warmup, --prefill, --csv, --print-detections, --dump-score,
// --quiet); an unknown flag or value is an error (exit 2), never ignored.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "hsi_host.hpp"

namespace {

void usage() {
  std::puts(R"(hsi_detect - hyperspectral target detection on an MI355X

  hsi_detect [options]

Source
  --source=synthetic|envi       where frames come from               [synthetic]
  --envi=A.hdr,B.hdr            ENVI cubes to replay (cycled)
  --resident=N                  ENVI cubes held in RAM (the first N)       [4]
  --width=N --height=N          synthetic frame size                 [512x512]
  --bands=N                     synthetic band count                     [113]
  --targets=N                   synthetic library size                     [4]
  --noise=F                     synthetic noise sigma                  [0.005]
  --drift=F                     synthetic per-frame drift of the mix       [0]
  --seed=N                      synthetic seed                      [20261203]

Detection
  --library=FILE.csv            target spectra; the synthetic source supplies
                                its own when this is absent
  --detector=LIST               sam, rx, ace, cem, comma-separated       [sam]
  --threshold=RAD               SAM spectral angle threshold            [0.10]
  --rx-threshold=F              RX threshold (Mahalanobis squared); default:
                                the chi-squared quantile of --rx-false-alarm
  --rx-false-alarm=RATE         per-pixel false-alarm rate for RX       [1e-4]
  --ace-threshold=F             ACE threshold                            [0.5]
  --cem-threshold=F             CEM threshold                            [0.5]
  --nms=N                       non-maximum suppression radius, 0 off      [2]
  --max-detections=N            detection list capacity per detector   [4096]
  --loading=F                   diagonal loading of the statistics      [1e-6]

Execution
  --variant=baseline|optimized|half   baseline and optimized score the FP32
                                cube; half narrows each frame to FP16 on the
                                device and scores that        [optimized]
  --memory=copy                 frames are copied to HBM (the only mode on a
                                discrete GPU; zerocopy is refused)      [copy]
  --precision=plain|mixed       RX/ACE arithmetic form                 [plain]
  --prescreen=MODEL             RX/ACE through a trained learned pre-screen
                                (hsi_mi355x.prescreen.Prescreen.save): one
                                pass builds the statistics (mixed form) and
                                screens every pixel; only pixels at or above
                                the model's cut are scored, the rest score 0
  --batch                       score the scenes together instead of as a
                                stream: each --envi cube (any size; bands
                                matching the library for SAM/ACE/CEM), or the
                                first --frames synthetic frames [4], uploaded
                                once, their statistics and scoring run as one
                                batch replayed from a captured hipGraph; lists
                                are dumped with frame = scene index
  --autotune                    with --batch: time the launch shapes first
  --streams=N                   frames in flight                           [3]
  --frames=N                    frames to process                        [300]
  --warmup=N                    unmeasured frames first                   [30]
  --prefill                     stage the ring once and replay it
  --no-graph                    launch kernels one by one instead of replaying
                                a captured hipGraph per frame
  --device=N                    HIP device                                 [0]

Output
  --csv=FILE                    write per-frame timings
  --print-detections=N          print the first N detections of the last frame
  --dump-score=FILE             write the last frame's first score map (raw f32)
  --dump-detections=FILE        write every measured frame's detection lists
                                (CSV: frame,detector,x,y,target,score)
  --json                        print one JSON summary line
  --quiet
  --help)");
}

// Every flag this program reads; anything else is refused.
const std::set<std::string> kFlags = {
    "source", "envi", "resident", "width", "height", "bands", "targets", "noise", "drift", "seed",
    "library", "detector", "threshold", "rx-threshold", "rx-false-alarm", "ace-threshold",
    "cem-threshold", "nms", "max-detections", "loading", "variant", "memory", "precision", "prescreen", "batch",
    "autotune", "streams",
    "frames", "warmup", "prefill", "no-graph", "device", "csv", "print-detections", "dump-score",
    "dump-detections", "json", "quiet", "help", "h"};

struct Args {
  std::map<std::string, std::string> v;
  bool has(const std::string& k) const { return v.count(k) != 0; }
  std::string str(const std::string& k, const std::string& d) const {
    auto it = v.find(k);
    return it == v.end() || it->second.empty() ? d : it->second;
  }
  long long integer(const std::string& k, long long d) const {
    auto it = v.find(k);
    return it == v.end() || it->second.empty() ? d : std::stoll(it->second);
  }
  double real(const std::string& k, double d) const {
    auto it = v.find(k);
    return it == v.end() || it->second.empty() ? d : std::stod(it->second);
  }
};

// --key=value and --flag tokens; a stray positional argument or an unknown
// key is reported in `bad`.
Args parse(int argc, char** argv, std::string* bad) {
  Args a;
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
    if (eq == std::string::npos) a.v[key] = "";
    else a.v[key] = t.substr(eq + 1);
  }
  return a;
}

std::vector<std::string> commas(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == ',') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

// --batch: the scenes scored together (hsi::run_scene_batch).
int run_batch(const Args& args, const hsi::PipelineOptions& opt, const std::string& kind,
              const hsi_synth_config& synth, const hsi::SpectralLibrary& library, bool quiet) {
  for (const char* flag : {"prescreen", "prefill"})
    if (args.has(flag)) {
      std::fprintf(stderr, "--%s does not apply to --batch\n", flag);
      return 2;
    }
  if (opt.precision != HSI_PLAIN || opt.variant == hsi::SamVariant::Half) {
    std::fputs("--batch scores in plain precision on FP32 cubes: drop --precision=mixed and --variant=half\n",
               stderr);
    return 2;
  }
  std::vector<hsi::HsiFrame> scenes;
  if (kind == "envi") {
    for (const std::string& path : commas(args.str("envi", ""))) scenes.push_back(hsi::read_envi(path));
  } else {
    const long long n = args.integer("frames", 4);
    if (n < 1 || n > 256) {
      std::fputs("--batch takes 1..256 synthetic frames\n", stderr);
      return 2;
    }
    for (long long k = 0; k < n; ++k) {
      hsi::HsiFrame f;
      f.shape = hsi::CubeShape{synth.height, synth.width, synth.bands};
      f.data.resize(f.shape.elements());
      if (hsi_synth_frame(&synth, static_cast<std::uint64_t>(k), f.data.data(), f.shape.pixels(), nullptr, 0) != 0)
        throw std::runtime_error("synthetic frame generation failed");
      scenes.push_back(std::move(f));
    }
  }
  const hsi::BatchRun run = hsi::run_scene_batch(scenes, library, opt, args.has("autotune"));
  double pixels = 0;
  unsigned long long total = 0;
  bool stats_ok = true;
  for (const hsi::BatchSceneResult& r : run.scenes) {
    pixels += static_cast<double>(r.shape.pixels());
    stats_ok &= r.statistics_ok;
    for (const auto& dr : r.results) total += dr.detections_found;
  }
  if (!quiet) {
    std::printf("batch of %zu scenes, %.0f pixels: %.3f ms (statistics + scoring%s), %.1f Mpix/s\n",
                run.scenes.size(), pixels, static_cast<double>(run.ms_batch), opt.use_graph ? ", graph replay" : "",
                pixels / (static_cast<double>(run.ms_batch) * 1e3));
    if (!run.autotune_report.empty()) std::printf("%s", run.autotune_report.c_str());
    for (std::size_t s = 0; s < run.scenes.size(); ++s) {
      const hsi::BatchSceneResult& r = run.scenes[s];
      std::printf("scene %zu %dx%dx%d%s:", s, r.shape.height, r.shape.width, r.shape.bands,
                  r.statistics_ok ? "" : " (statistics failed)");
      for (const auto& dr : r.results) std::printf(" %s %u", hsi::to_string(dr.detector), dr.detections_found);
      std::printf("\n");
    }
  }
  if (args.has("dump-detections")) {
    std::ofstream out(args.str("dump-detections", ""));
    if (!out) throw std::runtime_error("cannot write " + args.str("dump-detections", ""));
    out << "frame,detector,x,y,target,score\n";
    char line[160];
    for (std::size_t s = 0; s < run.scenes.size(); ++s)
      for (const auto& dr : run.scenes[s].results)
        for (const hsi::Detection& d : dr.detections) {
          std::snprintf(line, sizeof line, "%zu,%s,%d,%d,%d,%.9g\n", s, hsi::to_string(dr.detector), d.x, d.y,
                        d.target, static_cast<double>(d.score));
          out << line;
        }
  }
  if (args.has("json"))
    std::printf("{\"scenes\": %zu, \"pixels\": %.0f, \"batch_ms\": %.4f, \"mpix_s\": %.3f, \"graph\": %s, "
                "\"detections\": %llu, \"statistics_ok\": %s}\n",
                run.scenes.size(), pixels, static_cast<double>(run.ms_batch),
                pixels / (static_cast<double>(run.ms_batch) * 1e3), opt.use_graph ? "true" : "false", total,
                stats_ok ? "true" : "false");
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  std::string bad;
  const Args args = parse(argc, argv, &bad);
  if (args.has("help") || args.has("h")) {
    usage();
    return 0;
  }
  if (!bad.empty()) {
    std::fprintf(stderr, "hsi_detect: %s (see --help)\n", bad.c_str());
    return 2;
  }
  try {
    // ---- execution flags that can be refused before any work --------------
    hsi::SamVariant variant = hsi::SamVariant::Optimized;
    if (!hsi::parse_sam_variant(args.str("variant", "optimized"), &variant)) {
      std::fputs("--variant must be baseline, optimized or half\n", stderr);
      return 2;
    }
    const std::string memory = args.str("memory", "copy");
    if (memory == "zerocopy") {
      std::fputs("--memory=zerocopy maps host memory into every kernel read; it is meant for integrated "
                 "GPUs, and the MI355X reads frames from its own HBM: use --memory=copy\n", stderr);
      return 2;
    }
    if (memory != "copy") {
      std::fputs("--memory must be copy or zerocopy\n", stderr);
      return 2;
    }
    const bool quiet = args.has("quiet");
    // ---- source -----------------------------------------------------------
    std::unique_ptr<hsi::FrameSource> source;
    const std::string kind = args.str("source", "synthetic");
    hsi_synth_config synth{};
    if (kind == "synthetic") {
      hsi::SyntheticOptions so;
      hsi_synth_default(&so.config);
      so.config.width = static_cast<int>(args.integer("width", 512));
      so.config.height = static_cast<int>(args.integer("height", 512));
      so.config.bands = static_cast<int>(args.integer("bands", 113));
      so.config.num_targets = static_cast<int>(args.integer("targets", 4));
      so.config.noise_sigma = args.real("noise", 0.005);
      so.config.seed = static_cast<std::uint64_t>(args.integer("seed", 20261203));
      so.config.drift = args.real("drift", 0.0);
      so.frames = 0;
      synth = so.config;
      source = hsi::make_synthetic_source(so);
    } else if (kind == "envi") {
      if (!args.has("envi")) {
        std::fputs("--source=envi needs --envi=<header>[,<header>...]\n", stderr);
        return 2;
      }
      // A batch reads each cube whole (scenes may differ in shape); a stream
      // replays cubes of one shape.
      if (!args.has("batch"))
        source = hsi::make_envi_replay_source(commas(args.str("envi", "")), true,
                                              static_cast<int>(args.integer("resident", 4)));
    } else if (kind == "hot") {
      std::fputs("--source=hot (folders of snapshot-mosaic PNG frames) is not part of this build; "
                 "decode the frames and demosaic them on the device with hsi_demosaic\n", stderr);
      return 2;
    } else {
      std::fprintf(stderr, "unknown --source=%s\n", kind.c_str());
      return 2;
    }
    // ---- library ----------------------------------------------------------
    hsi::SpectralLibrary library;
    if (args.has("library")) library = hsi::read_spectra_csv(args.str("library", ""));
    else if (source && source->library()) library = *source->library();
    // ---- pipeline ---------------------------------------------------------
    hsi::PipelineOptions opt;
    if (!hsi::parse_detector_list(args.str("detector", "sam"), &opt.detectors)) {
      std::fputs("--detector must list sam, rx, ace or cem\n", stderr);
      return 2;
    }
    const std::string prec = args.str("precision", "plain");
    if (prec != "plain" && prec != "mixed") {
      std::fputs("--precision must be plain or mixed\n", stderr);
      return 2;
    }
    opt.precision = prec == "mixed" ? HSI_MIXED : HSI_PLAIN;
    if (args.has("prescreen")) {
      if (args.str("prescreen", "").empty()) {
        std::fputs("--prescreen needs a model file\n", stderr);
        return 2;
      }
      opt.use_prescreen = true;
      opt.prescreen = hsi::read_prescreen(args.str("prescreen", ""));
    }
    opt.variant = variant;
    opt.loading = args.real("loading", 1e-6);
    opt.detection.threshold = static_cast<float>(args.real("threshold", 0.10));
    opt.thresholds.ace = static_cast<float>(args.real("ace-threshold", 0.5));
    opt.thresholds.cem = static_cast<float>(args.real("cem-threshold", 0.5));
    opt.thresholds.rx = static_cast<float>(args.real("rx-threshold", std::nan("")));
    opt.thresholds.rx_false_alarm_rate = args.real("rx-false-alarm", 1e-4);
    opt.detection.nms_radius = static_cast<int>(args.integer("nms", 2));
    opt.detection.max_detections = static_cast<int>(args.integer("max-detections", 4096));
    opt.num_streams = static_cast<int>(args.integer("streams", 3));
    opt.prefill = args.has("prefill") ? 1 : 0;
    opt.use_graph = !args.has("no-graph");
    opt.device = static_cast<int>(args.integer("device", 0));
    opt.return_score_map = args.has("dump-score");
    if (args.has("batch"))
      return run_batch(args, opt, kind, synth, library, quiet);
    if (args.has("autotune")) {
      std::fputs("--autotune applies to --batch\n", stderr);
      return 2;
    }
    std::string error;
    auto pipeline = hsi::Pipeline::create(source.get(), library, opt, &error);
    if (!pipeline) {
      std::fprintf(stderr, "pipeline: %s\n", error.c_str());
      return 1;
    }
    if (!quiet) std::printf("%s\n", pipeline->describe().c_str());
    const long long warmup = args.integer("warmup", 30), frames = args.integer("frames", 300);
    if (warmup > 0 && !pipeline->run(static_cast<std::uint64_t>(warmup), nullptr, &error)) {
      std::fprintf(stderr, "warm-up: %s\n", error.c_str());
      return 1;
    }
    pipeline->reset_stats();
    std::ofstream csv;
    if (args.has("csv")) {
      csv.open(args.str("csv", ""));
      csv << "frame,ms_wall,ms_gpu,ms_source,ms_upload,ms_stats,ms_score,ms_detect,ms_download,statistics_ok";
      for (hsi::Detector d : opt.detectors) csv << ",found_" << hsi::to_string(d);
      csv << "\n";
    }
    std::ofstream dets_out;
    if (args.has("dump-detections")) {
      dets_out.open(args.str("dump-detections", ""));
      if (!dets_out) throw std::runtime_error("cannot write " + args.str("dump-detections", ""));
      dets_out << "frame,detector,x,y,target,score\n";
    }
    hsi::FrameResult last;
    bool stats_failed = false;
    const bool ok = pipeline->run(
        static_cast<std::uint64_t>(frames),
        [&](const hsi::FrameResult& r) {
          if (csv.is_open()) {
            csv << r.meta.index << "," << r.ms_wall << "," << r.ms_gpu << "," << r.ms_source << "," << r.ms_upload
                << "," << r.ms_stats << "," << r.ms_score << "," << r.ms_detect << "," << r.ms_download << ","
                << (r.statistics_ok ? 1 : 0);
            for (const auto& dr : r.results) csv << "," << dr.detections_found;
            csv << "\n";
          }
          if (dets_out.is_open()) {
            char line[160];
            for (const auto& dr : r.results)
              for (const hsi::Detection& d : dr.detections) {
                std::snprintf(line, sizeof line, "%llu,%s,%d,%d,%d,%.9g\n",
                              static_cast<unsigned long long>(r.meta.index), hsi::to_string(dr.detector), d.x, d.y,
                              d.target, static_cast<double>(d.score));
                dets_out << line;
              }
          }
          stats_failed |= !r.statistics_ok;
          last = r;
        },
        &error);
    if (!ok) {
      std::fprintf(stderr, "run: %s\n", error.c_str());
      return 1;
    }
    const hsi::PipelineStats st = pipeline->stats();
    if (!quiet) {
      std::printf("frames %llu  %.1f fps  %.1f Mpix/s  %.2f GB/s of cube\n",
                  static_cast<unsigned long long>(st.frames), st.fps, st.mpixel_per_s, st.cube_gb_per_s);
      std::printf("latency ms  p50 %.3f  p95 %.3f  p99 %.3f  max %.3f\n", st.ms_p50, st.ms_p95, st.ms_p99, st.ms_max);
      std::printf("mean ms  source %.3f  upload %.3f  stats %.3f  score %.3f  detect %.3f  download %.3f  gpu %.3f\n",
                  st.mean_source, st.mean_upload, st.mean_stats, st.mean_score, st.mean_detect, st.mean_download,
                  st.mean_gpu);
      for (const auto& dr : last.results)
        std::printf("last frame %s: %u found\n", hsi::to_string(dr.detector), dr.detections_found);
      if (stats_failed) std::puts("warning: statistics failed on some frames; try a larger --loading");
    }
    const long long show = args.integer("print-detections", 0);
    for (const auto& dr : last.results)
      for (long long i = 0; i < show && i < static_cast<long long>(dr.detections.size()); ++i) {
        const hsi::Detection& d = dr.detections[static_cast<std::size_t>(i)];
        std::printf("%s x=%d y=%d target=%d score=%.6g\n", hsi::to_string(dr.detector), d.x, d.y, d.target, d.score);
      }
    if (args.has("dump-score") && !last.results.empty()) {
      std::ofstream out(args.str("dump-score", ""), std::ios::binary);
      const auto& m = last.results.front().score_map;
      out.write(reinterpret_cast<const char*>(m.data()), static_cast<std::streamsize>(m.size() * sizeof(float)));
    }
    if (args.has("json")) {
      std::printf("{\"frames\": %llu, \"fps\": %.3f, \"mpix_s\": %.3f, \"cube_gb_s\": %.3f, \"p50_ms\": %.4f, "
                  "\"p99_ms\": %.4f, \"mean_gpu_ms\": %.4f, \"graph\": %s, \"detections\": %llu, "
                  "\"statistics_ok\": %s}\n",
                  static_cast<unsigned long long>(st.frames), st.fps, st.mpixel_per_s, st.cube_gb_per_s, st.ms_p50,
                  st.ms_p99, st.mean_gpu, opt.use_graph ? "true" : "false",
                  static_cast<unsigned long long>(st.total_detections), stats_failed ? "false" : "true");
    }
    return 0;
  } catch (const std::exception& e) {
    std::fprintf(stderr, "hsi_detect: %s\n", e.what());
    return 1;
  }
}
