// io.cpp - ENVI cubes, CSV spectral libraries, detector names and library
// flattening for the C++ host layer. Host-only code (no HIP).
//
// File formats follow the reference's readers so the same files load:
// ENVI header keys and sample scaling (envi.cpp:80-258), library CSV rows
// (spectra_io.cpp:27-91).
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

#include "hsi_host.hpp"

namespace hsi {
namespace {

std::string strip(const std::string& s) {
  std::size_t a = 0, b = s.size();
  while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
  while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
  return s.substr(a, b - a);
}

std::string lowercase(std::string s) {
  for (char& c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

std::vector<std::string> comma_fields(const std::string& s) {
  std::vector<std::string> out;
  std::string cur;
  for (char c : s) {
    if (c == ',' || c == '\n' || c == '\r') {
      out.push_back(strip(cur));
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  out.push_back(strip(cur));
  return out;
}

struct EnviHeader {
  std::map<std::string, std::string> fields;
  int samples = 0, lines = 0, bands = 0, data_type = 0;
  long long offset = 0;
  bool big_endian = false;
  Interleave interleave = Interleave::Bsq;
  std::vector<float> wavelengths;
};

int int_field(const EnviHeader& h, const std::string& key) {
  auto it = h.fields.find(key);
  if (it == h.fields.end()) throw std::runtime_error("ENVI header is missing required field '" + key + "'");
  try {
    std::size_t used = 0;
    const int v = std::stoi(it->second, &used);
    return v;
  } catch (const std::exception&) {
    throw std::runtime_error("ENVI field '" + key + "' is not an integer: " + it->second);
  }
}

std::size_t sample_bytes(int data_type) {
  switch (data_type) {
    case 1: return 1;
    case 2: case 12: return 2;
    case 3: case 13: case 4: return 4;
    case 5: return 8;
  }
  throw std::runtime_error("unsupported ENVI data type " + std::to_string(data_type));
}

EnviHeader parse_header(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open ENVI header: " + path);
  std::stringstream all;
  all << in.rdbuf();
  const std::string text = all.str();
  EnviHeader h;
  std::size_t pos = 0;
  bool first = true;
  while (pos < text.size()) {
    std::size_t eol = text.find('\n', pos);
    if (eol == std::string::npos) eol = text.size();
    std::string line = strip(text.substr(pos, eol - pos));
    pos = eol + 1;
    if (first) {
      first = false;
      if (lowercase(line) == "envi") continue;
    }
    const std::size_t eq = line.find('=');
    if (line.empty() || eq == std::string::npos) continue;
    const std::string key = lowercase(strip(line.substr(0, eq)));
    std::string value = strip(line.substr(eq + 1));
    if (!value.empty() && value[0] == '{') {
      // a brace list may run over several lines
      while (value.find('}') == std::string::npos) {
        if (pos >= text.size()) throw std::runtime_error("unterminated '{' in ENVI header: " + path);
        eol = text.find('\n', pos);
        if (eol == std::string::npos) eol = text.size();
        value += "\n" + text.substr(pos, eol - pos);
        pos = eol + 1;
      }
      value = value.substr(1, value.find('}') - 1);
    }
    h.fields[key] = value;
  }
  h.samples = int_field(h, "samples");
  h.lines = int_field(h, "lines");
  h.bands = int_field(h, "bands");
  h.data_type = int_field(h, "data type");
  if (h.fields.count("header offset")) h.offset = int_field(h, "header offset");
  if (h.fields.count("byte order")) h.big_endian = int_field(h, "byte order") == 1;
  if (h.fields.count("interleave")) {
    const std::string v = lowercase(strip(h.fields["interleave"]));
    if (v == "bsq") h.interleave = Interleave::Bsq;
    else if (v == "bil") h.interleave = Interleave::Bil;
    else if (v == "bip") h.interleave = Interleave::Bip;
    else throw std::runtime_error("unknown ENVI interleave '" + v + "'");
  }
  if (h.samples <= 0 || h.lines <= 0 || h.bands <= 0)
    throw std::runtime_error("ENVI header has a non-positive dimension: " + path);
  sample_bytes(h.data_type);
  if (h.fields.count("wavelength")) {
    for (const std::string& f : comma_fields(h.fields["wavelength"]))
      if (!f.empty()) h.wavelengths.push_back(std::stof(f));
    if (static_cast<int>(h.wavelengths.size()) != h.bands) h.wavelengths.clear();
  }
  return h;
}

std::string data_path_for(const std::string& hdr) {
  std::vector<std::string> candidates;
  const std::size_t dot = hdr.rfind('.');
  const std::string stem = dot == std::string::npos ? hdr : hdr.substr(0, dot);
  // The extensions the reference tries, in its order (envi.cpp:184-191), so
  // both pick the same file when several lie beside one header; then a few
  // upper-case spellings.
  for (const char* ext : {"", ".float", ".dat", ".img", ".bin", ".raw", ".bsq", ".bil", ".bip", ".cube", ".IMG",
                          ".RAW"})
    candidates.push_back(stem + ext);
  candidates.push_back(hdr + ".img");
  for (const std::string& c : candidates) {
    if (c == hdr) continue;
    std::ifstream probe(c, std::ios::binary);
    if (probe) return c;
  }
  throw std::runtime_error("no ENVI data file found alongside " + hdr);
}

template <typename T>
float load_sample(const unsigned char* p, bool swap) {
  unsigned char b[sizeof(T)];
  std::memcpy(b, p, sizeof(T));
  if (swap) std::reverse(b, b + sizeof(T));
  T v;
  std::memcpy(&v, b, sizeof(T));
  return static_cast<float>(v);
}

}  // namespace

std::vector<float> SpectralLibrary::flatten() const {
  const int b = bands();
  std::vector<float> out;
  out.reserve(static_cast<std::size_t>(size()) * b);
  for (const Spectrum& t : targets) {
    if (t.bands() != b)
      throw std::invalid_argument("SpectralLibrary: target '" + t.name + "' has " +
                                  std::to_string(t.bands()) + " bands, expected " + std::to_string(b));
    out.insert(out.end(), t.values.begin(), t.values.end());
  }
  return out;
}

std::vector<float> SpectralLibrary::norms() const {
  std::vector<float> out;
  for (const Spectrum& t : targets) {
    double acc = 0.0;
    for (float v : t.values) acc += static_cast<double>(v) * static_cast<double>(v);
    out.push_back(static_cast<float>(std::sqrt(acc)));
  }
  return out;
}

DetectorTraits traits_of(Detector d) {
  switch (d) {
    case Detector::Sam: return {"sam", true, false, ScorePolarity::LowerIsBetter, false};
    case Detector::Rx: return {"rx", false, true, ScorePolarity::HigherIsBetter, true};
    case Detector::Ace: return {"ace", true, true, ScorePolarity::HigherIsBetter, true};
    case Detector::Cem: return {"cem", true, true, ScorePolarity::HigherIsBetter, false};
  }
  throw std::invalid_argument("unknown detector");
}

const char* to_string(Detector d) { return traits_of(d).name; }

bool parse_detector(const std::string& name, Detector* out) {
  const std::string n = lowercase(strip(name));
  for (Detector d : {Detector::Sam, Detector::Rx, Detector::Ace, Detector::Cem}) {
    if (n == traits_of(d).name) {
      *out = d;
      return true;
    }
  }
  return false;
}

bool parse_sam_variant(const std::string& name, SamVariant* out) {
  const std::string n = lowercase(strip(name));
  if (n == "baseline") *out = SamVariant::Baseline;
  else if (n == "optimized") *out = SamVariant::Optimized;
  else if (n == "half") *out = SamVariant::Half;
  else return false;
  return true;
}

const char* to_string(SamVariant v) {
  switch (v) {
    case SamVariant::Baseline: return "baseline";
    case SamVariant::Optimized: return "optimized";
    case SamVariant::Half: return "half";
  }
  return "?";
}

float detection_threshold(const PipelineOptions& o, Detector d, int bands) {
  switch (d) {
    case Detector::Sam: return o.detection.threshold;
    case Detector::Ace: return o.thresholds.ace;
    case Detector::Cem: return o.thresholds.cem;
    case Detector::Rx: {
      if (!std::isnan(o.thresholds.rx)) return o.thresholds.rx;
      const double t = hsi_rx_threshold(bands, o.thresholds.rx_false_alarm_rate);
      if (std::isnan(t))
        throw std::invalid_argument("the RX false-alarm rate must lie in (0, 1), got " +
                                    std::to_string(o.thresholds.rx_false_alarm_rate));
      return static_cast<float>(t);
    }
  }
  throw std::invalid_argument("unknown detector");
}

bool parse_detector_list(const std::string& names, std::vector<Detector>* out) {
  std::vector<Detector> list;
  for (const std::string& f : comma_fields(names)) {
    Detector d;
    if (!parse_detector(f, &d)) return false;
    if (std::find(list.begin(), list.end(), d) == list.end()) list.push_back(d);
  }
  if (list.empty()) return false;
  *out = list;
  return true;
}

HsiFrame read_envi(const std::string& hdr_path) {
  const EnviHeader h = parse_header(hdr_path);
  const std::string data = data_path_for(hdr_path);
  const std::size_t sb = sample_bytes(h.data_type);
  const std::size_t count = static_cast<std::size_t>(h.samples) * h.lines * h.bands;
  std::ifstream in(data, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open ENVI data file: " + data);
  in.seekg(0, std::ios::end);
  const long long size = static_cast<long long>(in.tellg());
  if (size < h.offset + static_cast<long long>(count * sb))
    throw std::runtime_error("ENVI data file " + data + " is " + std::to_string(size) +
                             " bytes, header promises " + std::to_string(h.offset + count * sb));
  in.seekg(h.offset);
  std::vector<unsigned char> raw(count * sb);
  in.read(reinterpret_cast<char*>(raw.data()), static_cast<std::streamsize>(raw.size()));
  if (!in) throw std::runtime_error("short read on ENVI data file: " + data);

  const bool host_big = [] {
    const std::uint16_t one = 1;
    unsigned char c;
    std::memcpy(&c, &one, 1);
    return c == 0;
  }();
  const bool swap = h.big_endian != host_big;
  // float arithmetic throughout, as the reference's reader: the sample
  // widened to float, times a float full-scale reciprocal.
  float scale = 1.0f;
  switch (h.data_type) {
    case 1: scale = 1.0f / 255.0f; break;
    case 2: scale = 1.0f / 32767.0f; break;
    case 12: scale = 1.0f / 65535.0f; break;
    case 3: case 13: scale = 1.0f / 2147483647.0f; break;
    default: break;
  }
  HsiFrame f;
  f.shape = {h.lines, h.samples, h.bands};
  f.wavelengths_nm = h.wavelengths;
  f.data.resize(count);
  const std::size_t P = f.shape.pixels();
  const int W = h.samples, B = h.bands;
  for (std::size_t i = 0; i < count; ++i) {
    const unsigned char* p = raw.data() + i * sb;
    float v = 0.0f;
    switch (h.data_type) {
      case 1: v = static_cast<float>(*p); break;
      case 2: v = load_sample<std::int16_t>(p, swap); break;
      case 12: v = load_sample<std::uint16_t>(p, swap); break;
      case 3: v = load_sample<std::int32_t>(p, swap); break;
      case 13: v = load_sample<std::uint32_t>(p, swap); break;
      case 4: v = load_sample<float>(p, swap); break;
      case 5: v = load_sample<double>(p, swap); break;
    }
    const float value = v * scale;
    // position of element i in the file's interleave -> BSQ index
    std::size_t b, pix;
    if (h.interleave == Interleave::Bsq) {
      b = i / P;
      pix = i % P;
    } else if (h.interleave == Interleave::Bip) {
      pix = i / B;
      b = i % B;
    } else {  // BIL: line, band, sample
      const std::size_t line = i / (static_cast<std::size_t>(B) * W);
      const std::size_t rem = i % (static_cast<std::size_t>(B) * W);
      b = rem / W;
      pix = line * W + rem % W;
    }
    f.data[b * P + pix] = value;
  }
  return f;
}

void write_envi(const std::string& hdr_path, const HsiFrame& frame) {
  const std::size_t dot = hdr_path.rfind('.');
  const std::string data = (dot == std::string::npos ? hdr_path : hdr_path.substr(0, dot)) + ".img";
  std::ofstream h(hdr_path);
  if (!h) throw std::runtime_error("cannot write ENVI header: " + hdr_path);
  h << "ENVI\nsamples = " << frame.shape.width << "\nlines = " << frame.shape.height
    << "\nbands = " << frame.shape.bands << "\nheader offset = 0\nfile type = ENVI Standard\n"
    << "data type = 4\ninterleave = bsq\nbyte order = 0\n";
  if (!frame.wavelengths_nm.empty()) {
    h << "wavelength = {";
    for (std::size_t i = 0; i < frame.wavelengths_nm.size(); ++i)
      h << (i ? ", " : "") << frame.wavelengths_nm[i];
    h << "}\n";
  }
  std::ofstream d(data, std::ios::binary);
  if (!d) throw std::runtime_error("cannot write ENVI data: " + data);
  d.write(reinterpret_cast<const char*>(frame.data.data()),
          static_cast<std::streamsize>(frame.data.size() * sizeof(float)));
}

SpectralLibrary read_spectra_csv(const std::string& path) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open spectral library: " + path);
  SpectralLibrary lib;
  std::vector<float> wavelengths;
  std::string line;
  int n = 0;
  while (std::getline(in, line)) {
    ++n;
    const std::string t = strip(line);
    if (t.empty() || t[0] == '#') continue;
    const std::vector<std::string> f = comma_fields(t);
    if (f.size() < 2)
      throw std::runtime_error(path + ":" + std::to_string(n) + ": expected a name followed by at least one value");
    std::vector<float> values;
    for (std::size_t i = 1; i < f.size(); ++i) {
      try {
        std::size_t used = 0;
        values.push_back(std::stof(f[i], &used));
      } catch (const std::exception&) {
        throw std::runtime_error(path + ":" + std::to_string(n) + ": '" + f[i] + "' is not a number");
      }
    }
    if (f[0] == "wavelength" || f[0] == "wavelengths") {
      wavelengths = values;
      continue;
    }
    lib.targets.push_back({f[0], values, {}});
  }
  if (lib.targets.empty()) throw std::runtime_error(path + ": no target spectra found");
  const int b = lib.bands();
  for (const Spectrum& s : lib.targets)
    if (s.bands() != b)
      throw std::runtime_error(path + ": target '" + s.name + "' has " + std::to_string(s.bands()) +
                               " bands, expected " + std::to_string(b));
  if (!wavelengths.empty()) {
    if (static_cast<int>(wavelengths.size()) != b)
      throw std::runtime_error(path + ": wavelength row has " + std::to_string(wavelengths.size()) +
                               " entries but the spectra have " + std::to_string(b) + " bands");
    for (Spectrum& s : lib.targets) s.wavelengths_nm = wavelengths;
  }
  return lib;
}

void write_spectra_csv(const std::string& path, const SpectralLibrary& library) {
  std::ofstream out(path);
  if (!out) throw std::runtime_error("cannot write spectral library: " + path);
  out.precision(9);
  if (!library.targets.empty() && !library.targets.front().wavelengths_nm.empty()) {
    out << "wavelength";
    for (float w : library.targets.front().wavelengths_nm) out << "," << w;
    out << "\n";
  }
  for (const Spectrum& s : library.targets) {
    out << s.name;
    for (float v : s.values) out << "," << v;
    out << "\n";
  }
}

PrescreenModel read_prescreen(const std::string& path) {
  std::ifstream in(path, std::ios::binary);
  if (!in) throw std::runtime_error("cannot open pre-screen model: " + path);
  char magic[8] = {};
  in.read(magic, sizeof magic);
  if (!in || std::memcmp(magic, "HSIMLP01", 8) != 0)
    throw std::runtime_error("not a pre-screen model (bad magic): " + path);
  std::int32_t bands = 0;
  float cut = 0.f;
  std::uint32_t count = 0;
  in.read(reinterpret_cast<char*>(&bands), sizeof bands);
  in.read(reinterpret_cast<char*>(&cut), sizeof cut);
  in.read(reinterpret_cast<char*>(&count), sizeof count);
  if (!in) throw std::runtime_error("truncated pre-screen model: " + path);
  const std::size_t want = hsi_mlp_param_count(bands);
  if (want == 0 || count != want)
    throw std::runtime_error("pre-screen model " + path + ": " + std::to_string(count) + " weights for " +
                             std::to_string(bands) + " bands (expected " + std::to_string(want) + ")");
  if (!std::isfinite(cut)) throw std::runtime_error("pre-screen model " + path + ": the cut is not finite");
  PrescreenModel m;
  m.bands = bands;
  m.cut = cut;
  m.weights.resize(count);
  in.read(reinterpret_cast<char*>(m.weights.data()), static_cast<std::streamsize>(count * sizeof(float)));
  if (!in) throw std::runtime_error("truncated pre-screen model: " + path);
  for (float v : m.weights)
    if (!std::isfinite(v)) throw std::runtime_error("pre-screen model " + path + " holds a non-finite weight");
  return m;
}

}  // namespace hsi
