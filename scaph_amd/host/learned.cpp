// learned.cpp — host side of the learned direction schedule.
#include "learned.h"

#include <cmath>
#include <cstring>
#include <fstream>

namespace scaph {

float bf16_to_f32(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

uint16_t f32_to_bf16(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return (uint16_t)((u >> 16) | 0x40);
  const uint32_t lsb = (u >> 16) & 1u;
  u += 0x7FFFu + lsb;
  return (uint16_t)(u >> 16);
}

std::array<float, kMlpIn> level_features(const SchedState& s, const LevelRecord* last) {
  std::array<float, kMlpIn> x{};
  x[0] = (float)std::log2(1.0 + (double)s.n_f);
  x[1] = (float)std::log2(1.0 + (double)s.m_f);
  x[2] = (float)std::log2(1.0 + (double)s.m_u);
  x[3] = (float)s.level;
  if (last)
    for (int i = 0; i < 12; ++i) x[4 + i] = last->page_hist[i];
  return x;
}

void LearnedSchedule::set(const uint16_t* params, const float* mean, const float* scale,
                          float prefetch_gain) {
  p_.assign(params, params + kMlpParams);
  for (int i = 0; i < kMlpIn; ++i) {
    mean_[i] = mean[i];
    scale_[i] = scale[i] == 0.f ? 1.f : scale[i];
  }
  prefetch_gain_ = prefetch_gain;
  flags_ = 0;  // the caller's offsets are used as given
  eff_ = mean_;
  ready_ = true;
}

void LearnedSchedule::load_file(const std::string& path) {
  // raw little-endian: kMlpParams u16, 16 f32 mean, 16 f32 scale, 1 f32
  // gain, then (newer files) 1 u32 of flags
  std::ifstream in(path, std::ios::binary);
  if (!in) throw Error(1, "cannot open model file " + path);
  std::vector<uint16_t> p(kMlpParams);
  std::array<float, kMlpIn> m{}, s{};
  float g = 0;
  in.read(reinterpret_cast<char*>(p.data()), p.size() * 2);
  in.read(reinterpret_cast<char*>(m.data()), sizeof(m));
  in.read(reinterpret_cast<char*>(s.data()), sizeof(s));
  in.read(reinterpret_cast<char*>(&g), 4);
  if (!in) throw Error(1, "model file " + path + " is truncated");
  uint32_t flags = 0;
  if (!in.read(reinterpret_cast<char*>(&flags), 4)) flags = 0;
  set(p.data(), m.data(), s.data(), g);
  flags_ = flags;
}

std::array<float, kMlpIn> LearnedSchedule::mean_for(uint64_t n, uint64_t m) const {
  std::array<float, kMlpIn> out = mean_;
  if (flags_ & 1u) {
    out[0] += (float)std::log2((double)std::max<uint64_t>(n, 1));
    out[1] += (float)std::log2((double)std::max<uint64_t>(m, 1));
    out[2] += (float)std::log2((double)std::max<uint64_t>(m, 1));
  }
  return out;
}

void LearnedSchedule::forward(const float* x, float* out) const {
  const uint16_t* W1 = p_.data();
  const uint16_t* b1 = W1 + kMlpHidden * kMlpIn;
  const uint16_t* W2 = b1 + kMlpHidden;
  const uint16_t* b2 = W2 + kMlpHidden * kMlpHidden;
  const uint16_t* W3 = b2 + kMlpHidden;
  const uint16_t* b3 = W3 + kMlpOut * kMlpHidden;
  float h1[kMlpHidden], h2[kMlpHidden];
  for (int j = 0; j < kMlpHidden; ++j) {
    float a = bf16_to_f32(b1[j]);
    for (int i = 0; i < kMlpIn; ++i) a += bf16_to_f32(W1[j * kMlpIn + i]) * x[i];
    h1[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < kMlpHidden; ++j) {
    float a = bf16_to_f32(b2[j]);
    for (int i = 0; i < kMlpHidden; ++i) a += bf16_to_f32(W2[j * kMlpHidden + i]) * h1[i];
    h2[j] = a > 0.f ? a : 0.f;
  }
  for (int j = 0; j < kMlpOut; ++j) {
    float a = bf16_to_f32(b3[j]);
    for (int i = 0; i < kMlpHidden; ++i) a += bf16_to_f32(W3[j * kMlpHidden + i]) * h2[i];
    out[j] = a;
  }
}

Dir LearnedSchedule::decide(const Engine& eng, const SchedState& s) const {
  if (!ready_ || !eng.has_pull()) return eng.beamer(s);
  std::array<float, kMlpIn> x = level_features(s, s.last);
  for (int i = 0; i < kMlpIn; ++i) x[i] = (x[i] - eff_[i]) / scale_[i];
  float y[kMlpOut];
  forward(x.data(), y);
  return y[0] > 0.f ? Dir::kPull : Dir::kPush;
}

}  // namespace scaph
