// sources.cpp - frame sources for the C++ host layer (frame_source.hpp's
// FrameSource contract: write BSQ frames into caller memory with a given
// plane stride). Host-only code.
#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "hsi_host.hpp"

namespace hsi {
namespace {

class SyntheticSource final : public FrameSource {
 public:
  explicit SyntheticSource(const SyntheticOptions& o) : opt_(o) {
    std::vector<float> lib(static_cast<std::size_t>(o.config.num_targets) * o.config.bands);
    if (o.config.num_targets > 0 && hsi_synth_library(&opt_.config, lib.data()) != 0)
      throw std::invalid_argument("synthetic source: bad configuration");
    for (int t = 0; t < o.config.num_targets; ++t) {
      Spectrum s;
      s.name = "endmember_" + std::to_string(t);
      s.values.assign(lib.begin() + static_cast<std::ptrdiff_t>(t) * o.config.bands,
                      lib.begin() + static_cast<std::ptrdiff_t>(t + 1) * o.config.bands);
      library_.targets.push_back(std::move(s));
    }
    if (o.keep_truth) truth_.resize(shape().pixels());
  }

  CubeShape shape() const override { return {opt_.config.height, opt_.config.width, opt_.config.bands}; }
  std::uint64_t length() const override { return opt_.frames; }

  bool read_into(float* dst, std::size_t plane_stride, FrameMeta* meta) override {
    if (opt_.frames && next_ >= opt_.frames) return false;
    if (hsi_synth_frame(&opt_.config, next_, dst, plane_stride, truth_.empty() ? nullptr : truth_.data(),
                        opt_.threads) != 0)
      throw std::runtime_error("synthetic source: frame generation failed");
    if (meta) {
      meta->index = next_;
      meta->timestamp_s = static_cast<double>(next_) / 30.0;
    }
    ++next_;
    return true;
  }

  void reset() override { next_ = 0; }
  std::string describe() const override {
    const CubeShape s = shape();
    return "synthetic " + std::to_string(s.height) + "x" + std::to_string(s.width) + "x" +
           std::to_string(s.bands) + ", " + std::to_string(opt_.config.num_targets) + " targets, seed " +
           std::to_string(opt_.config.seed);
  }
  const std::uint8_t* truth_mask() const override { return truth_.empty() ? nullptr : truth_.data(); }
  const SpectralLibrary* library() const override { return &library_; }

 private:
  SyntheticOptions opt_;
  SpectralLibrary library_;
  std::vector<std::uint8_t> truth_;
  std::uint64_t next_ = 0;
};

class EnviReplaySource final : public FrameSource {
 public:
  EnviReplaySource(const std::vector<std::string>& paths, bool loop, int max_resident) : loop_(loop) {
    if (paths.empty()) throw std::invalid_argument("ENVI replay needs at least one header");
    const std::size_t keep = static_cast<std::size_t>(max_resident < 1 ? 1 : max_resident);
    for (const std::string& p : paths) {
      if (frames_.size() >= keep) break;
      frames_.push_back(read_envi(p));
      if (frames_.back().shape != frames_.front().shape)
        throw std::invalid_argument("ENVI replay: " + p + " has a different shape from " + paths.front());
    }
    names_.assign(paths.begin(), paths.begin() + static_cast<std::ptrdiff_t>(frames_.size()));
  }
  CubeShape shape() const override { return frames_.front().shape; }
  std::uint64_t length() const override { return loop_ ? 0 : frames_.size(); }
  bool read_into(float* dst, std::size_t plane_stride, FrameMeta* meta) override {
    if (!loop_ && next_ >= frames_.size()) return false;
    // Only the band planes are written; the elements between them are the
    // caller's and stay as they were (frame_source.cpp's replay does the same,
    // test_sources.cpp:196-214).
    const HsiFrame& f = frames_[next_ % frames_.size()];
    const std::size_t P = f.shape.pixels();
    for (int b = 0; b < f.shape.bands; ++b)
      std::memcpy(dst + static_cast<std::size_t>(b) * plane_stride, f.data.data() + static_cast<std::size_t>(b) * P,
                  P * sizeof(float));
    if (meta) {
      meta->index = next_;
      meta->timestamp_s = 0.0;
    }
    ++next_;
    return true;
  }
  void reset() override { next_ = 0; }
  std::string describe() const override {
    return "envi replay of " + std::to_string(frames_.size()) + " cube(s), first " + names_.front();
  }

 private:
  std::vector<HsiFrame> frames_;
  std::vector<std::string> names_;
  bool loop_;
  std::uint64_t next_ = 0;
};

}  // namespace

std::unique_ptr<FrameSource> make_synthetic_source(const SyntheticOptions& options) {
  return std::unique_ptr<FrameSource>(new SyntheticSource(options));
}

std::unique_ptr<FrameSource> make_envi_replay_source(const std::vector<std::string>& hdr_paths, bool loop,
                                                     int max_resident) {
  return std::unique_ptr<FrameSource>(new EnviReplaySource(hdr_paths, loop, max_resident));
}

}  // namespace hsi
