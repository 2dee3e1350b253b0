// learned.h — learned direction schedule: a 16-64-64-2 MLP (bf16 weights,
// fp32 accumulation) that reads the per-iteration features of the level
// loop and picks push or pull for the next step.  Beamer's fixed rule stays
// the fallback (no model loaded) and levels never depend on the choice.
#ifndef SCAPH_HOST_LEARNED_H
#define SCAPH_HOST_LEARNED_H

#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "engine.h"

namespace scaph {

constexpr int kMlpIn = 16, kMlpHidden = 64, kMlpOut = 2;
constexpr int kMlpParams = kMlpHidden * kMlpIn + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           kMlpOut * kMlpHidden + kMlpOut;

// bf16 <-> fp32 from the bf16 definition (top 16 bits of an IEEE binary32;
// round-to-nearest-even on the dropped 16 bits, NaN kept quiet).
float bf16_to_f32(uint16_t h);
uint16_t f32_to_bf16(float f);

// The 16 features of one step: log2(1+n_f), log2(1+m_f), log2(1+m_u), the
// step index, and the 12-bin page-activity histogram (fractions of pages
// whose active-vertex share falls in each bin).
std::array<float, kMlpIn> level_features(const SchedState& s, const LevelRecord* last);

class LearnedSchedule {
 public:
  bool ready() const { return ready_; }
  void set(const uint16_t* params, const float* mean, const float* scale, float prefetch_gain);
  void load_file(const std::string& path);
  // fp32-accumulate forward pass; out[0] = pull logit, out[1] = prefetch threshold
  void forward(const float* x, float* out) const;
  Dir decide(const Engine& eng, const SchedState& s) const;
  const uint16_t* params() const { return p_.data(); }
  const float* mean() const { return mean_.data(); }
  // the offsets for a graph of n vertices and m edges: mean(), plus log2 n,
  // log2 m, log2 m on features 0..2 when the model reads shares (flag bit 0)
  std::array<float, kMlpIn> mean_for(uint64_t n, uint64_t m) const;
  // binds the offsets of the loaded graph (decide() and graph_mean() use them)
  void set_graph(uint64_t n, uint64_t m) { eff_ = mean_for(n, m); }
  const float* graph_mean() const { return eff_.data(); }
  const float* scale() const { return scale_.data(); }

 private:
  bool ready_ = false;
  std::vector<uint16_t> p_;
  std::array<float, kMlpIn> mean_{}, scale_{}, eff_{};
  float prefetch_gain_ = 0.f;
  uint32_t flags_ = 0;
};

}  // namespace scaph

#endif
