// gen.cpp — host generators (multi-threaded generation and sort).
#include "gen.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <map>
#include <sstream>
#include <thread>

#include "../csrc/rmat_spec.h"

namespace scaph {

namespace {

int pick_threads(int t) {
  if (t > 0) return t;
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 4;
  return (int)std::min(16u, h);
}

template <typename F>
void parallel_ranges(uint64_t n, int threads, F f) {
  std::vector<std::thread> ts;
  for (int t = 0; t < threads; ++t) {
    const uint64_t b = n * t / threads, e = n * (t + 1) / threads;
    ts.emplace_back([=] { f(b, e); });
  }
  for (auto& th : ts) th.join();
}

void parallel_sort_unique(std::vector<uint64_t>* keys, int threads) {
  std::vector<uint64_t>& k = *keys;
  const uint64_t n = k.size();
  if (n < (1u << 16) || threads <= 1) {
    std::sort(k.begin(), k.end());
  } else {
    std::vector<uint64_t> bounds;
    for (int t = 0; t <= threads; ++t) bounds.push_back(n * t / threads);
    parallel_ranges(threads, threads, [&](uint64_t b, uint64_t e) {
      for (uint64_t t = b; t < e; ++t) std::sort(k.begin() + bounds[t], k.begin() + bounds[t + 1]);
    });
    // pairwise merge rounds
    for (size_t width = 1; width < (size_t)threads; width *= 2) {
      std::vector<std::thread> ts;
      for (size_t i = 0; i + width < (size_t)threads; i += 2 * width) {
        const uint64_t a = bounds[i], m = bounds[i + width],
                       e = bounds[std::min(i + 2 * width, (size_t)threads)];
        ts.emplace_back([&k, a, m, e] {
          std::inplace_merge(k.begin() + a, k.begin() + m, k.begin() + e);
        });
      }
      for (auto& th : ts) th.join();
    }
  }
  k.erase(std::unique(k.begin(), k.end()), k.end());
}

uint32_t to_thr(double x) {
  if (x <= 0) return 0;
  const double t = std::floor(x * 4294967296.0);
  if (t >= 4294967295.0) return 0xFFFFFFFFu;
  return (uint32_t)t;
}

}  // namespace

uint32_t zipf_degree(uint64_t seed, uint32_t v, double x_min, double alpha, uint32_t cap) {
  const double u = (double)(scaph_spec::word(seed, v) >> 11) * (1.0 / 9007199254740992.0);
  const double x = x_min * std::pow(1.0 - u, -1.0 / (alpha - 1.0));
  return x >= (double)cap ? cap : (uint32_t)x;
}

void rmat_thresholds(const RmatParams& p, uint32_t* thr_a, uint32_t* thr_ab, uint32_t* thr_abc) {
  *thr_a = to_thr(p.a);
  *thr_ab = to_thr(p.a + p.b);
  *thr_abc = to_thr(p.a + p.b + p.c);
}

HostCsr csr_from_sorted_keys(const std::vector<uint64_t>& keys, uint32_t n) {
  HostCsr g;
  g.n = n;
  g.offsets.assign((size_t)n + 1, 0);
  g.ngh.reserve(keys.size());
  for (uint64_t k : keys) {
    const uint32_t u = (uint32_t)(k >> 32), v = (uint32_t)k;
    if (u == v) continue;  // self-loop
    g.offsets[(size_t)u + 1]++;
    g.ngh.push_back(v);
  }
  for (uint32_t v = 0; v < n; ++v) g.offsets[v + 1] += g.offsets[v];
  return g;
}

HostCsr generate_rmat(const RmatParams& p, int threads) {
  threads = pick_threads(threads);
  uint32_t ta, tab, tabc;
  rmat_thresholds(p, &ta, &tab, &tabc);
  const uint64_t n = 1ull << p.scale;
  const uint64_t m = (uint64_t)p.edge_factor * n;
  const scaph_spec::PermKey pk = scaph_spec::make_perm_key(p.seed, p.scale);
  std::vector<uint64_t> keys(m);
  parallel_ranges(m, threads, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      uint32_t u, v;
      scaph_spec::rmat_edge(p.seed, i, p.scale, ta, tab, tabc, &u, &v);
      u = (uint32_t)scaph_spec::perm_range(pk, u, n);
      v = (uint32_t)scaph_spec::perm_range(pk, v, n);
      keys[i] = ((uint64_t)u << 32) | v;
    }
  });
  parallel_sort_unique(&keys, threads);
  return csr_from_sorted_keys(keys, (uint32_t)n);
}

HostCsr generate_er(uint32_t n, uint64_t m, uint64_t seed, int threads) {
  threads = pick_threads(threads);
  std::vector<uint64_t> keys(m);
  parallel_ranges(m, threads, [&](uint64_t b, uint64_t e) {
    for (uint64_t i = b; i < e; ++i) {
      const uint64_t wu = scaph_spec::word(seed, 2 * i), wv = scaph_spec::word(seed, 2 * i + 1);
      const uint32_t u = (uint32_t)(((unsigned __int128)wu * n) >> 64);
      const uint32_t v = (uint32_t)(((unsigned __int128)wv * n) >> 64);
      keys[i] = ((uint64_t)u << 32) | v;
    }
  });
  parallel_sort_unique(&keys, threads);
  return csr_from_sorted_keys(keys, n);
}

uint32_t max_degree_vertex(const HostCsr& g) {
  uint32_t best = 0;
  uint64_t bd = 0;
  for (uint32_t v = 0; v < g.n; ++v) {
    const uint64_t d = g.offsets[(size_t)v + 1] - g.offsets[v];
    if (d > bd) {
      bd = d;
      best = v;
    }
  }
  return best;
}

void fill_uniform_weights(HostCsr* g, uint32_t wmax, uint64_t seed) {
  constexpr uint64_t kWeightSalt = 0x7765696768747321ull;  // "weights!"
  const uint64_t key = scaph_spec::fmix64(seed ^ kWeightSalt);
  const uint64_t m = g->ngh.size();
  g->wgh.resize(m);
  uint32_t* w = g->wgh.data();
  parallel_ranges(m, m < (1u << 20) ? 1 : pick_threads(0), [=](uint64_t b, uint64_t e) {
    for (uint64_t k = b; k < e; ++k) w[k] = 1u + (uint32_t)(scaph_spec::word(key, k) % wmax);
  });
}

bool generate_from_spec(const std::string& spec, HostCsr* out, std::string* err) {
  const size_t colon = spec.find(':');
  const std::string kind = spec.substr(0, colon);
  std::map<std::string, std::string> kv;
  if (colon != std::string::npos) {
    std::stringstream ss(spec.substr(colon + 1));
    std::string item;
    while (std::getline(ss, item, ',')) {
      const size_t eq = item.find('=');
      if (eq == std::string::npos) {
        if (err) *err = "malformed generator item '" + item + "'";
        return false;
      }
      kv[item.substr(0, eq)] = item.substr(eq + 1);
    }
  }
  auto num = [&](const char* k, double d) { return kv.count(k) ? std::atof(kv[k].c_str()) : d; };
  auto u64 = [&](const char* k, uint64_t d) {
    return kv.count(k) ? (uint64_t)std::strtoull(kv[k].c_str(), nullptr, 0) : d;
  };
  const uint64_t wmax = u64("wmax", 0);
  if (wmax > 0xFFFFFFFFull) {
    if (err) *err = "wmax must fit in 32 bits";
    return false;
  }
  if (kind == "rmat") {
    RmatParams p;
    p.scale = (uint32_t)u64("scale", 20);
    p.edge_factor = (uint32_t)u64("ef", 16);
    p.a = num("a", 0.57);
    p.b = num("b", 0.19);
    p.c = num("c", 0.19);
    p.seed = u64("seed", 1);
    if (p.scale == 0 || p.scale > 31) {
      if (err) *err = "rmat scale must be in 1..31";
      return false;
    }
    *out = generate_rmat(p);
    if (wmax) fill_uniform_weights(out, (uint32_t)wmax, p.seed);
    return true;
  }
  if (kind == "er") {
    const uint32_t n = (uint32_t)u64("n", 1000);
    const double deg = num("deg", 8);
    const uint64_t seed = u64("seed", 1);
    if (n == 0) {
      if (err) *err = "er n must be > 0";
      return false;
    }
    *out = generate_er(n, (uint64_t)std::llround(deg * n), seed);
    if (wmax) fill_uniform_weights(out, (uint32_t)wmax, seed);
    return true;
  }
  if (err) *err = "unknown generator '" + kind + "'";
  return false;
}

}  // namespace scaph
