// synth.cpp - seeded synthetic frames (include/hsi_synth.h).
#include "hsi_synth.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr double kTwoPi = 6.283185307179586476925;

std::uint64_t mix64(std::uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Independent random streams, one per purpose.
enum Stream : std::uint64_t {
  kEndmember = 1,
  kWeights = 2,
  kPlant = 3,
  kAbundance = 4,
  kGain = 5,
  kNoise = 6,
  kAnomaly = 7,
};

struct Rng {
  std::uint64_t key;
  Rng(std::uint64_t seed, std::uint64_t stream, std::uint64_t frame)
      : key(mix64(mix64(seed ^ mix64(stream)) + frame)) {}
  // Uniform in [0, 1) with 53 random bits.
  double uniform(std::uint64_t index) const {
    return static_cast<double>(mix64(key + index * 0xD1B54A32D192ED03ull) >> 11) * 0x1.0p-53;
  }
  // Uniform in (0, 1]: safe under log.
  double open_uniform(std::uint64_t index) const { return 1.0 - uniform(index); }
};

bool valid(const hsi_synth_config* c) {
  return c && c->height > 0 && c->width > 0 && c->bands > 1 && c->num_targets >= 0 &&
         c->num_endmembers > 0 && c->plant_fraction >= 0.0 && c->anomaly_fraction >= 0.0 &&
         c->noise_sigma >= 0.0 &&
         c->plant_fraction * c->num_targets + c->anomaly_fraction <= 1.0;
}

// Endmember k: ramp plus six Gaussian bumps, mapped onto [0.05, 0.9].
std::vector<double> endmember(const hsi_synth_config& c, int k) {
  const Rng rng(c.seed, kEndmember, 0);
  const std::uint64_t base = static_cast<std::uint64_t>(k) * 64;
  const double offset = 0.1 + 0.4 * rng.uniform(base + 0);
  const double slope = -0.3 + 0.6 * rng.uniform(base + 1);
  double centre[6], width[6], height[6];
  for (int j = 0; j < 6; ++j) {
    centre[j] = rng.uniform(base + 2 + 3 * j);
    width[j] = 0.03 + 0.17 * rng.uniform(base + 3 + 3 * j);
    height[j] = -0.3 + 0.9 * rng.uniform(base + 4 + 3 * j);
  }
  std::vector<double> s(static_cast<std::size_t>(c.bands));
  for (int b = 0; b < c.bands; ++b) {
    const double lam = static_cast<double>(b) / static_cast<double>(c.bands - 1);
    double v = offset + slope * lam;
    for (int j = 0; j < 6; ++j) {
      const double d = (lam - centre[j]) / width[j];
      v += height[j] * std::exp(-0.5 * d * d);
    }
    s[static_cast<std::size_t>(b)] = v;
  }
  const auto [lo, hi] = std::minmax_element(s.begin(), s.end());
  const double mn = *lo, span = (*hi - *lo) > 0.0 ? (*hi - *lo) : 1.0;
  for (double& v : s) v = 0.05 + 0.85 * (v - mn) / span;
  return s;
}

// Endmember indices: background 0..E-1 (the first ones double as targets),
// extra targets E..T-1 when the library is larger, then the anomaly.
int spectra_count(const hsi_synth_config& c) {
  return std::max(c.num_endmembers, c.num_targets) + 1;
}

}  // namespace

extern "C" {

void hsi_synth_default(hsi_synth_config* c) {
  if (!c) return;
  c->height = 256;
  c->width = 256;
  c->bands = 113;
  c->num_targets = 4;
  c->num_endmembers = 8;
  c->drift_period = 300;
  c->plant_fraction = 0.005;
  c->anomaly_fraction = 0.0;
  c->noise_sigma = 0.005;
  c->drift = 0.0;
  c->seed = 20261203;
}

int hsi_synth_library(const hsi_synth_config* c, float* out) {
  if (!valid(c) || !out) return -1;
  for (int t = 0; t < c->num_targets; ++t) {
    const std::vector<double> s = endmember(*c, t);
    for (int b = 0; b < c->bands; ++b)
      out[static_cast<std::size_t>(t) * c->bands + b] = static_cast<float>(s[static_cast<std::size_t>(b)]);
  }
  return 0;
}

int hsi_synth_anomaly(const hsi_synth_config* c, float* out) {
  if (!valid(c) || !out) return -1;
  const std::vector<double> s = endmember(*c, spectra_count(*c) - 1);
  for (int b = 0; b < c->bands; ++b) out[b] = static_cast<float>(s[static_cast<std::size_t>(b)]);
  return 0;
}

int hsi_synth_frame(const hsi_synth_config* cp, uint64_t frame, float* dst, size_t stride,
                    uint8_t* truth, int threads) {
  if (!valid(cp) || !dst) return -1;
  const hsi_synth_config c = *cp;
  const std::size_t pixels = static_cast<std::size_t>(c.height) * c.width;
  if (stride < pixels) return -1;
  const int B = c.bands;
  const int E = c.num_endmembers;
  const int S = spectra_count(c);
  std::vector<double> spec(static_cast<std::size_t>(S) * B);
  for (int k = 0; k < S; ++k) {
    const std::vector<double> s = endmember(c, k);
    std::copy(s.begin(), s.end(), spec.begin() + static_cast<std::ptrdiff_t>(k) * B);
  }
  // Per-frame drift of the mixing weights: endmember k is scaled by
  // 1 + drift * sin(2 pi (frame / period + k / E)).
  std::vector<double> drift(static_cast<std::size_t>(E), 1.0);
  if (c.drift != 0.0 && c.drift_period > 0) {
    for (int k = 0; k < E; ++k)
      drift[static_cast<std::size_t>(k)] =
          1.0 + c.drift * std::sin(kTwoPi * (static_cast<double>(frame) / c.drift_period +
                                             static_cast<double>(k) / E));
  }
  const Rng weights(c.seed, kWeights, frame), abundance(c.seed, kAbundance, frame),
      gain_rng(c.seed, kGain, frame), noise(c.seed, kNoise, frame);
  // Where targets and anomalies sit is fixed across frames of one seed.
  const Rng plant(c.seed, kPlant, 0), anomaly(c.seed, kAnomaly, 0);
  const double plant_total = c.plant_fraction * c.num_targets;

  auto work = [&](std::size_t p0, std::size_t p1) {
    std::vector<double> w(static_cast<std::size_t>(E));
    std::vector<double> x(static_cast<std::size_t>(B));
    for (std::size_t p = p0; p < p1; ++p) {
      double total = 0.0;
      for (int k = 0; k < E; ++k) {
        const double e = -std::log(weights.open_uniform(p * E + k)) * drift[static_cast<std::size_t>(k)];
        w[static_cast<std::size_t>(k)] = e;
        total += e;
      }
      for (int b = 0; b < B; ++b) {
        double v = 0.0;
        for (int k = 0; k < E; ++k) v += w[static_cast<std::size_t>(k)] * spec[static_cast<std::size_t>(k) * B + b];
        x[static_cast<std::size_t>(b)] = v / total;
      }
      std::uint8_t label = 0;
      const double u = plant.uniform(p);
      if (u < plant_total && c.plant_fraction > 0.0) {
        int t = static_cast<int>(u / c.plant_fraction);
        if (t >= c.num_targets) t = c.num_targets - 1;
        const double a = 0.6 + 0.4 * abundance.uniform(2 * p);
        const double* s = spec.data() + static_cast<std::size_t>(t) * B;
        for (int b = 0; b < B; ++b) x[static_cast<std::size_t>(b)] = a * s[b] + (1.0 - a) * x[static_cast<std::size_t>(b)];
        label = static_cast<std::uint8_t>(t + 1);
      } else if (c.anomaly_fraction > 0.0 && anomaly.uniform(p) < c.anomaly_fraction) {
        const double a = 0.3 + 0.7 * abundance.uniform(2 * p + 1);
        const double* s = spec.data() + static_cast<std::size_t>(S - 1) * B;
        for (int b = 0; b < B; ++b) x[static_cast<std::size_t>(b)] = a * s[b] + (1.0 - a) * x[static_cast<std::size_t>(b)];
        label = 255;
      }
      if (truth) truth[p] = label;
      const double g = 0.3 + 0.9 * gain_rng.uniform(p);
      const std::uint64_t nbase = p * static_cast<std::uint64_t>(B + 1);
      for (int b = 0; b < B; b += 2) {
        // Box-Muller: one pair of normals per two bands.
        const double r = std::sqrt(-2.0 * std::log(noise.open_uniform(nbase + b)));
        const double th = kTwoPi * noise.uniform(nbase + b + 1);
        const double n0 = c.noise_sigma * r * std::cos(th);
        const double n1 = c.noise_sigma * r * std::sin(th);
        double v0 = g * x[static_cast<std::size_t>(b)] + n0;
        dst[static_cast<std::size_t>(b) * stride + p] = static_cast<float>(v0 > 0.0 ? v0 : 0.0);
        if (b + 1 < B) {
          double v1 = g * x[static_cast<std::size_t>(b + 1)] + n1;
          dst[static_cast<std::size_t>(b + 1) * stride + p] = static_cast<float>(v1 > 0.0 ? v1 : 0.0);
        }
      }
    }
  };

  unsigned n = threads > 0 ? static_cast<unsigned>(threads) : std::thread::hardware_concurrency();
  if (n == 0) n = 1;
  if (n > pixels / 1024 + 1) n = static_cast<unsigned>(pixels / 1024 + 1);
  if (n <= 1) {
    work(0, pixels);
  } else {
    std::vector<std::thread> pool;
    const std::size_t chunk = (pixels + n - 1) / n;
    for (unsigned i = 0; i < n; ++i) {
      const std::size_t a = i * chunk, b = std::min(pixels, a + chunk);
      if (a >= b) break;
      pool.emplace_back(work, a, b);
    }
    for (std::thread& t : pool) t.join();
  }
  if (stride > pixels) {
    for (int b = 0; b < B; ++b)
      std::memset(dst + static_cast<std::size_t>(b) * stride + pixels, 0, (stride - pixels) * sizeof(float));
  }
  return 0;
}

}  // extern "C"
