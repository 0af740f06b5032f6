// host_io_checks.cpp - CPU checks of the C++ host layer's file readers and
// frame sources, restating the reference's own tests of the same interfaces:
//   spectral library CSV     tests/test_spectra.cpp:41-117
//   ENVI cubes               tests/test_io.cpp:108-225
//   synthetic / ENVI replay  tests/test_sources.cpp:29-81, 173-229 (the synthetic
//                            source is the workload generator: see below)
// Built and run by tests/test_host_io.py; argv[1] is a scratch directory.
// Prints one line per failed check and a summary; exits non-zero on failure.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "hsi_host.hpp"

namespace {

int g_checks = 0, g_failed = 0;
std::string g_dir;

void expect(bool ok, const char* what, int line) {
  ++g_checks;
  if (!ok) {
    ++g_failed;
    std::printf("FAILED line %d: %s\n", line, what);
  }
}
#define EXPECT(cond) expect((cond), #cond, __LINE__)
#define EXPECT_NEAR(a, b, tol) expect(std::fabs(double(a) - double(b)) <= (tol), #a " ~ " #b, __LINE__)

std::string path_of(const std::string& name) { return g_dir + "/" + name; }

std::string put_text(const std::string& name, const std::string& text) {
  const std::string p = path_of(name);
  std::ofstream(p) << text;
  return p;
}

template <typename T>
void put_raw(const std::string& name, const std::vector<T>& v) {
  std::ofstream out(path_of(name), std::ios::binary);
  out.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

// The message of what `fn` throws, or "" when it returns.
std::string thrown_by(const std::function<void()>& fn) {
  try {
    fn();
  } catch (const std::exception& e) {
    return e.what();
  }
  return "";
}
bool mentions(const std::string& message, const char* needle) {
  return !message.empty() && message.find(needle) != std::string::npos;
}

// ---- spectral library CSV ---------------------------------------------------

void spectra_round_trip() {
  hsi::SpectralLibrary lib;
  lib.targets.push_back({"blood", {0.11f, 0.19f, 0.24f}, {400.f, 410.f, 420.f}});
  lib.targets.push_back({"tomato_concentrate", {0.31f, 0.28f, 0.22f}, {400.f, 410.f, 420.f}});
  const std::string p = path_of("library.csv");
  hsi::write_spectra_csv(p, lib);
  const hsi::SpectralLibrary back = hsi::read_spectra_csv(p);
  EXPECT(back.size() == 2 && back.bands() == 3);
  if (back.size() != 2 || back.bands() != 3) return;
  EXPECT(back.targets[0].name == "blood" && back.targets[1].name == "tomato_concentrate");
  for (int t = 0; t < 2; ++t)
    for (int b = 0; b < 3; ++b) {
      EXPECT_NEAR(back.targets[t].values[b], lib.targets[t].values[b], 1e-6);
      EXPECT(back.targets[t].wavelengths_nm.size() == 3);
      if (back.targets[t].wavelengths_nm.size() == 3) EXPECT_NEAR(back.targets[t].wavelengths_nm[b], 400.0 + 10 * b, 1e-3);
    }
}

void spectra_comments_blanks_whitespace() {
  const std::string p = put_text("messy.csv",
                                 "# hand-edited\n\nwavelength, 400.0 , 410.0\n   blood , 0.5 , 0.25\n\n"
                                 "# another comment\nketchup,0.1,0.9\n");
  const hsi::SpectralLibrary lib = hsi::read_spectra_csv(p);
  EXPECT(lib.size() == 2);
  if (lib.size() != 2) return;
  EXPECT(lib.targets[0].name == "blood");
  EXPECT(lib.targets[1].name == "ketchup");
  EXPECT_NEAR(lib.targets[0].values[0], 0.5, 1e-7);
  EXPECT_NEAR(lib.targets[0].values[1], 0.25, 1e-7);
  EXPECT_NEAR(lib.targets[1].values[1], 0.9, 1e-7);
}

void spectra_flatten_and_norms() {
  hsi::SpectralLibrary lib;
  lib.targets.push_back({"a", {3.f, 4.f}, {}});
  lib.targets.push_back({"b", {0.f, 2.f}, {}});
  const std::vector<float> flat = lib.flatten();
  EXPECT(flat.size() == 4);
  if (flat.size() == 4) EXPECT(flat[0] == 3.f && flat[1] == 4.f && flat[2] == 0.f && flat[3] == 2.f);  // target-major
  const std::vector<float> n = lib.norms();
  EXPECT(n.size() == 2);
  if (n.size() == 2) EXPECT(n[0] == 5.f && n[1] == 2.f);  // exact: sqrt of exact double sums
}

void spectra_errors_name_the_problem() {
  auto read = [](const std::string& p) { return [p] { hsi::read_spectra_csv(p); }; };
  EXPECT(mentions(thrown_by(read(put_text("ragged.csv", "blood,0.1,0.2\nketchup,0.3\n"))), "ketchup"));
  EXPECT(mentions(thrown_by(read(put_text("nan.csv", "blood,0.1,oops\n"))), "oops"));
  EXPECT(mentions(thrown_by(read(put_text("empty.csv", "# only a comment\n"))), "no target spectra"));
  EXPECT(mentions(thrown_by(read(put_text("wl.csv", "wavelength,400,410,420\nblood,0.1,0.2\n"))), "wavelength row"));
  EXPECT(mentions(thrown_by(read(path_of("no_such_library.csv"))), "cannot open"));
}

// ---- ENVI -------------------------------------------------------------------

float pattern(int b, int y, int x) { return static_cast<float>(b * 100 + y * 10 + x); }

// Element (b, y, x) of an H x W x B cube at its offset in each interleave.
std::size_t offset_of(hsi::Interleave il, int H, int W, int B, int b, int y, int x) {
  (void)H;
  switch (il) {
    case hsi::Interleave::Bsq: return (static_cast<std::size_t>(b) * H + y) * W + x;
    case hsi::Interleave::Bil: return (static_cast<std::size_t>(y) * B + b) * W + x;
    default: return (static_cast<std::size_t>(y) * W + x) * B + b;
  }
}

bool frame_matches(const hsi::HsiFrame& f, int H, int W, int B, float offset) {
  if (f.shape.height != H || f.shape.width != W || f.shape.bands != B) return false;
  if (f.data.size() != static_cast<std::size_t>(H) * W * B) return false;
  for (int b = 0; b < B; ++b)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x)
        if (f.data[(static_cast<std::size_t>(b) * H + y) * W + x] != offset + pattern(b, y, x)) return false;
  return true;
}

void envi_bil_with_a_real_world_header() {
  const int H = 3, W = 5, B = 4;
  std::vector<float> disk(static_cast<std::size_t>(H) * W * B);
  for (int b = 0; b < B; ++b)
    for (int y = 0; y < H; ++y)
      for (int x = 0; x < W; ++x) disk[offset_of(hsi::Interleave::Bil, H, W, B, b, y, x)] = pattern(b, y, x);
  put_raw("bil.float", disk);
  // Attribution comments (one shaped like a field), a braced multi-line
  // description holding '=' and a wavelength list wrapped over lines.
  const std::string hdr = put_text(
      "bil.hdr",
      "ENVI\n;test cube, licence text at http://example.org/x?a=b\n;samples = 9999\n"
      "description = {\n  a description = with signs\n}\nsamples = 5\nlines = 3\nbands = 4\n"
      "header offset = 0\nfile type = ENVI Standard\ndata type = 4\ninterleave = bil\nbyte order = 0\n"
      "wavelength = {\n 400.0, 450.0,\n 500.0, 550.0}\n");
  const hsi::HsiFrame f = hsi::read_envi(hdr);
  EXPECT(frame_matches(f, H, W, B, 0.f));
  EXPECT(f.wavelengths_nm.size() == 4);
  if (f.wavelengths_nm.size() == 4) {
    EXPECT_NEAR(f.wavelengths_nm[0], 400.0, 1e-3);
    EXPECT_NEAR(f.wavelengths_nm[1], 450.0, 1e-3);
    EXPECT_NEAR(f.wavelengths_nm[3], 550.0, 1e-3);
  }
}

void envi_integer_types_are_rescaled() {
  put_raw("u16.raw", std::vector<std::uint16_t>{0, 32767, 65535, 16383});
  const std::string hdr = put_text("u16.hdr",
                                   "ENVI\nsamples = 2\nlines = 2\nbands = 1\ndata type = 12\n"
                                   "interleave = bsq\nbyte order = 0\n");
  const hsi::HsiFrame f = hsi::read_envi(hdr);
  EXPECT(f.data.size() == 4);
  if (f.data.size() == 4) {
    EXPECT(f.data[0] == 0.f);
    EXPECT_NEAR(f.data[1], 0.5, 1e-4);
    EXPECT(f.data[2] == 1.f);
    EXPECT_NEAR(f.data[3], 0.25, 1e-4);
  }
  put_raw("u8.raw", std::vector<std::uint8_t>{0, 51, 255, 102});
  const hsi::HsiFrame g = hsi::read_envi(put_text(
      "u8.hdr", "ENVI\nsamples = 2\nlines = 2\nbands = 1\ndata type = 1\ninterleave = bsq\nbyte order = 0\n"));
  EXPECT(g.data.size() == 4);
  if (g.data.size() == 4) {
    EXPECT(g.data[0] == 0.f && g.data[2] == 1.f);
    EXPECT_NEAR(g.data[1], 0.2, 1e-6);
    EXPECT_NEAR(g.data[3], 0.4, 1e-6);
  }
}

void envi_every_interleave_reads_as_bsq() {
  const int H = 5, W = 7, B = 3;
  const struct { hsi::Interleave il; const char* name; } layouts[] = {
      {hsi::Interleave::Bsq, "bsq"}, {hsi::Interleave::Bil, "bil"}, {hsi::Interleave::Bip, "bip"}};
  for (const auto& l : layouts) {
    std::vector<float> disk(static_cast<std::size_t>(H) * W * B);
    for (int b = 0; b < B; ++b)
      for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) disk[offset_of(l.il, H, W, B, b, y, x)] = pattern(b, y, x);
    // float64 on disk for one layout, so the second float type is read too.
    const bool f64 = l.il == hsi::Interleave::Bip;
    const std::string stem = std::string("layout_") + l.name;
    if (f64) put_raw(stem + ".img", std::vector<double>(disk.begin(), disk.end()));
    else put_raw(stem + ".img", disk);
    const std::string hdr = put_text(stem + ".hdr", "ENVI\nsamples = 7\nlines = 5\nbands = 3\ndata type = " +
                                                        std::string(f64 ? "5" : "4") + "\ninterleave = " +
                                                        l.name + "\nbyte order = 0\n");
    EXPECT(frame_matches(hsi::read_envi(hdr), H, W, B, 0.f));
  }
}

void envi_write_read_round_trip() {
  hsi::HsiFrame f;
  f.shape = {4, 6, 5};
  f.data.resize(f.shape.elements());
  for (int b = 0; b < 5; ++b)
    for (int y = 0; y < 4; ++y)
      for (int x = 0; x < 6; ++x) f.data[(static_cast<std::size_t>(b) * 4 + y) * 6 + x] = pattern(b, y, x) + 0.125f;
  f.wavelengths_nm = {401.5f, 402.5f, 403.5f, 404.5f, 405.5f};
  const std::string hdr = path_of("written.hdr");
  hsi::write_envi(hdr, f);
  const hsi::HsiFrame g = hsi::read_envi(hdr);
  EXPECT(g.shape == f.shape && g.data == f.data);
  EXPECT(g.wavelengths_nm == f.wavelengths_nm);
}

void envi_errors_name_the_problem() {
  EXPECT(mentions(thrown_by([] { hsi::read_envi(path_of("no_such.hdr")); }), "cannot open"));
  const std::string missing = put_text("nobands.hdr", "ENVI\nsamples = 2\nlines = 2\ndata type = 4\ninterleave = bsq\n");
  EXPECT(mentions(thrown_by([&] { hsi::read_envi(missing); }), "bands"));
  put_raw("short.img", std::vector<float>(3, 1.f));
  const std::string shorter = put_text(
      "short.hdr", "ENVI\nsamples = 2\nlines = 2\nbands = 1\ndata type = 4\ninterleave = bsq\nbyte order = 0\n");
  EXPECT(!thrown_by([&] { hsi::read_envi(shorter); }).empty());
}

// ---- frame sources ----------------------------------------------------------

hsi::SyntheticOptions small_scene() {
  hsi::SyntheticOptions o;
  hsi_synth_default(&o.config);
  o.config.height = 96;
  o.config.width = 128;
  o.config.bands = 24;
  o.config.num_targets = 3;
  o.config.seed = 4242;
  o.threads = 2;
  return o;
}

void synthetic_shape_and_determinism() {
  const hsi::SyntheticOptions o = small_scene();
  auto a = hsi::make_synthetic_source(o);
  auto b = hsi::make_synthetic_source(o);
  const hsi::CubeShape want{96, 128, 24};
  EXPECT(a->shape() == want);
  EXPECT(a->library() != nullptr && a->library()->size() == 3 && a->library()->bands() == 24);
  std::vector<float> fa(want.elements()), fb(want.elements());
  hsi::FrameMeta ma, mb;
  EXPECT(a->read_into(fa.data(), want.pixels(), &ma));
  hsi::SyntheticOptions more = o;
  more.threads = 5;  // the frame does not depend on how the work is split
  auto c = hsi::make_synthetic_source(more);
  EXPECT(c->read_into(fb.data(), want.pixels(), &mb));
  EXPECT(fa == fb);
  EXPECT(ma.index == 0 && mb.index == 0);
  EXPECT(b->read_into(fb.data(), want.pixels(), &mb) && fa == fb);
  EXPECT(a->read_into(fb.data(), want.pixels(), &ma) && ma.index == 1 && fa != fb);
}

void synthetic_length_bounds_the_stream() {
  hsi::SyntheticOptions o = small_scene();
  o.frames = 3;
  auto s = hsi::make_synthetic_source(o);
  EXPECT(s->length() == 3);
  std::vector<float> first(s->shape().elements()), f(s->shape().elements());
  int n = 0;
  while (s->read_into(n == 0 ? first.data() : f.data(), s->shape().pixels(), nullptr)) ++n;
  EXPECT(n == 3);
  s->reset();
  EXPECT(s->read_into(f.data(), s->shape().pixels(), nullptr) && f == first);
}

// The reference's demo source moves its targets between frames
// (test_sources.cpp:65-81). This source is the seeded workload generator,
// whose scene keeps each planted pixel in place for a seed and drifts the
// background mix from frame to frame: the truth mask is the same for every
// frame, every target is planted, and the frames differ.
void synthetic_targets_stay_while_the_scene_drifts() {
  hsi::SyntheticOptions o = small_scene();
  o.keep_truth = true;
  o.config.drift = 0.2;
  auto s = hsi::make_synthetic_source(o);
  const std::size_t P = s->shape().pixels();
  std::vector<float> f0(s->shape().elements()), f1(s->shape().elements());
  EXPECT(s->read_into(f0.data(), P, nullptr) && s->truth_mask() != nullptr);
  const std::vector<std::uint8_t> first(s->truth_mask(), s->truth_mask() + P);
  EXPECT(s->read_into(f1.data(), P, nullptr));
  int planted[3] = {0, 0, 0};
  bool same = true;
  for (std::size_t p = 0; p < P; ++p) {
    if (first[p] >= 1 && first[p] <= 3) ++planted[first[p] - 1];
    same &= first[p] == s->truth_mask()[p];
  }
  EXPECT(planted[0] > 0 && planted[1] > 0 && planted[2] > 0);
  EXPECT(same);
  EXPECT(f0 != f1);
}

std::string replay_cube(const std::string& name, hsi::CubeShape shape, float offset) {
  hsi::HsiFrame f;
  f.shape = shape;
  f.data.resize(shape.elements());
  for (int b = 0; b < shape.bands; ++b)
    for (int y = 0; y < shape.height; ++y)
      for (int x = 0; x < shape.width; ++x)
        f.data[(static_cast<std::size_t>(b) * shape.height + y) * shape.width + x] = offset + pattern(b, y, x);
  const std::string hdr = path_of(name + ".hdr");
  hsi::write_envi(hdr, f);
  return hdr;
}

void envi_replay_cycles_and_honours_the_stride() {
  const hsi::CubeShape shape{3, 5, 4};  // 15 pixels: a padded stride differs from it
  const std::vector<std::string> paths = {replay_cube("one", shape, 0.f), replay_cube("two", shape, 1000.f)};
  auto s = hsi::make_envi_replay_source(paths, true);
  EXPECT(s->shape() == shape && s->length() == 0);
  const std::size_t P = shape.pixels(), stride = P + 3;
  std::vector<float> buf(stride * shape.bands, -1.f);
  for (int round = 0; round < 2; ++round)
    for (float offset : {0.f, 1000.f}) {
      EXPECT(s->read_into(buf.data(), stride, nullptr));
      bool planes = true, gaps = true;
      for (int b = 0; b < shape.bands; ++b) {
        for (std::size_t p = 0; p < P; ++p)
          planes &= buf[b * stride + p] == offset + pattern(b, static_cast<int>(p) / shape.width,
                                                            static_cast<int>(p) % shape.width);
        for (std::size_t p = P; p < stride; ++p) gaps &= buf[b * stride + p] == -1.f;  // the caller's, untouched
      }
      EXPECT(planes);
      EXPECT(gaps);
    }
  auto once = hsi::make_envi_replay_source(paths, false);
  EXPECT(once->length() == 2);
  EXPECT(once->read_into(buf.data(), stride, nullptr) && once->read_into(buf.data(), stride, nullptr));
  EXPECT(!once->read_into(buf.data(), stride, nullptr));
  // A cube of another shape would break every buffer sized for the first.
  std::vector<std::string> odd = paths;
  odd.push_back(replay_cube("odd", hsi::CubeShape{3, 5, 6}, 0.f));
  EXPECT(mentions(thrown_by([&] { hsi::make_envi_replay_source(odd, true); }), "shape"));
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s SCRATCH_DIR\n", argv[0]);
    return 2;
  }
  g_dir = argv[1];
  const std::pair<const char*, void (*)()> checks[] = {
      {"spectra_round_trip", spectra_round_trip},
      {"spectra_comments_blanks_whitespace", spectra_comments_blanks_whitespace},
      {"spectra_flatten_and_norms", spectra_flatten_and_norms},
      {"spectra_errors_name_the_problem", spectra_errors_name_the_problem},
      {"envi_bil_with_a_real_world_header", envi_bil_with_a_real_world_header},
      {"envi_integer_types_are_rescaled", envi_integer_types_are_rescaled},
      {"envi_every_interleave_reads_as_bsq", envi_every_interleave_reads_as_bsq},
      {"envi_write_read_round_trip", envi_write_read_round_trip},
      {"envi_errors_name_the_problem", envi_errors_name_the_problem},
      {"synthetic_shape_and_determinism", synthetic_shape_and_determinism},
      {"synthetic_length_bounds_the_stream", synthetic_length_bounds_the_stream},
      {"synthetic_targets_stay_while_the_scene_drifts", synthetic_targets_stay_while_the_scene_drifts},
      {"envi_replay_cycles_and_honours_the_stride", envi_replay_cycles_and_honours_the_stride},
  };
  for (const auto& c : checks) {
    const int before = g_failed;
    const std::string error = thrown_by(c.second);
    if (!error.empty()) {
      ++g_failed;
      std::printf("FAILED %s: threw %s\n", c.first, error.c_str());
    }
    std::printf("%s %s\n", g_failed == before ? "ok    " : "FAILED", c.first);
  }
  std::printf("%d checks, %d failed\n", g_checks, g_failed);
  return g_failed ? 1 : 0;
}
