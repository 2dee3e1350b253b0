// batch.cpp — see batch.h.
#include "batch.h"

#include <algorithm>
#include <numeric>

#include "gen.h"
#include "../csrc/rmat_spec.h"

namespace scaph {

Batch::Batch(uint32_t count, uint32_t min_scale, uint32_t max_scale, uint32_t min_ef,
             uint32_t max_ef, uint64_t seed, int device) {
  if (count == 0 || min_scale == 0 || min_scale > max_scale || max_scale > 16 || min_ef == 0 ||
      min_ef > max_ef)
    throw Error(SCAPH_STATUS_INVALID, "batch: need count > 0, 1 <= scales <= 16, 1 <= edge factors");
  H_ERR(scaph_init(device));
  scaph_device_props props;
  H_ERR(scaph_get_device_props(device, &props));
  cus_ = (uint32_t)std::max(1, props.compute_units);
  st_.reset(new Stream());
  graphs_.resize(count);
  uint64_t vb = 0, eb = 0;
  for (uint32_t g = 0; g < count; ++g) {
    scaph_batch_entry& G = graphs_[g];
    G.scale = min_scale + (uint32_t)(scaph_spec::word(seed, 4ull * g) % (max_scale - min_scale + 1));
    G.edge_factor = min_ef + (uint32_t)(scaph_spec::word(seed, 4ull * g + 1) % (max_ef - min_ef + 1));
    G.seed = scaph_spec::word(seed, 4ull * g + 2);
    G.n = 1u << G.scale;
    G.src = (uint32_t)(scaph_spec::word(seed, 4ull * g + 3) % G.n);
    G.vbase = vb;
    G.ebase = eb;
    vb += G.n;
    eb += (uint64_t)G.edge_factor * G.n;
    max_n_ = std::max(max_n_, G.n);
  }
  total_v_ = vb;
  if (total_v_ >= 0x7FFFFFFFull) throw Error(SCAPH_STATUS_INVALID, "batch too large");
  // largest graphs first
  order_.resize(count);
  std::iota(order_.begin(), order_.end(), 0u);
  std::stable_sort(order_.begin(), order_.end(), [&](uint32_t a, uint32_t b) {
    return (uint64_t)graphs_[a].edge_factor * graphs_[a].n >
           (uint64_t)graphs_[b].edge_factor * graphs_[b].n;
  });
  void* s = st_->get();
  d_graphs_.alloc(count * sizeof(scaph_batch_entry));
  d_order_.alloc(count * 4);
  H_ERR(scaph_copy_h2d_async(d_graphs_.get(), graphs_.data(), count * sizeof(scaph_batch_entry), s));
  H_ERR(scaph_copy_h2d_async(d_order_.get(), order_.data(), count * 4, s));
  // keys -> union CSR
  RmatParams rp;  // a, b, c of the standard R-MAT
  uint32_t ta, tab, tabc;
  rmat_thresholds(rp, &ta, &tab, &tabc);
  const uint64_t raw = eb;
  {
    DevBuf keys(raw * 8), tmp(raw * 8), perm(scaph_batch_perm_ws_bytes(count)),
        deg((size_t)total_v_ * 4), cnt(8);
    offsets_.alloc(((size_t)total_v_ + 1) * 8);
    ngh_.alloc(std::max<uint64_t>(raw, 1) * 4);
    H_ERR(scaph_batch_gen_rmat_keys(keys.get<uint64_t>(), d_graphs_.get<scaph_batch_entry>(), count,
                                    raw, ta, tab, tabc, perm.get(), s));
    uint64_t m = 0;
    H_ERR(scaph_csr_from_keys(keys.get<uint64_t>(), tmp.get<uint64_t>(), raw, (uint32_t)total_v_, 0,
                              nullptr, offsets_.get<uint64_t>(), ngh_.get<uint32_t>(),
                              deg.get<uint32_t>(), &m, s));
    total_e_ = m;
    st_->sync();
  }
  levels_.alloc((size_t)total_v_ * 4);
  work_.alloc(256);
}

Batch::~Batch() {
  for (auto& c : captured_) scaph_graph_destroy(c.second);
}

void* Batch::graph_for(const BatchTune& t) {
  for (auto& c : captured_)
    if (c.first == t) return c.second;
  void* s = st_->get();
  const uint32_t grid = cus_ * t.wgs_per_cu;
  qbuf_.ensure((size_t)grid * max_n_ * 4);
  H_ERR(scaph_graph_capture_begin(s));
  H_ERR(scaph_batch_bfs(offsets_.get<uint64_t>(), ngh_.get<uint32_t>(),
                        d_graphs_.get<scaph_batch_entry>(), d_order_.get<uint32_t>(), count(),
                        levels_.get<uint32_t>(), work_.get<uint32_t>(), qbuf_.get<uint32_t>(),
                        max_n_, t.block, grid, s));
  void* g = nullptr;
  H_ERR(scaph_graph_capture_end(s, &g));
  captured_.emplace_back(t, g);
  return g;
}

double Batch::run(const BatchTune& t) {
  if (t.wgs_per_cu == 0 || t.wgs_per_cu > 64) throw Error(SCAPH_STATUS_INVALID, "batch: wgs_per_cu");
  void* s = st_->get();
  const uint32_t grid = cus_ * t.wgs_per_cu;
  void* g = t.use_graph ? graph_for(t) : nullptr;
  qbuf_.ensure((size_t)grid * max_n_ * 4);
  Event e0, e1;
  e0.record(s);
  if (g) {
    H_ERR(scaph_graph_launch(g, s));
  } else {
    H_ERR(scaph_batch_bfs(offsets_.get<uint64_t>(), ngh_.get<uint32_t>(),
                          d_graphs_.get<scaph_batch_entry>(), d_order_.get<uint32_t>(), count(),
                          levels_.get<uint32_t>(), work_.get<uint32_t>(), qbuf_.get<uint32_t>(),
                          max_n_, t.block, grid, s));
  }
  e1.record(s);
  e1.sync();
  return Event::elapsed_ms(e0, e1);
}

void Batch::levels(uint32_t g, uint32_t* host_out) {
  const scaph_batch_entry& G = graph(g);
  H_ERR(scaph_copy_d2h_async(host_out, levels_.get<uint32_t>() + G.vbase, (size_t)G.n * 4,
                             st_->get()));
  st_->sync();
}

void Batch::csr(uint32_t g, uint64_t* offsets, uint32_t* ngh, uint64_t cap, uint64_t* count) {
  const scaph_batch_entry& G = graph(g);
  std::vector<uint64_t> off(G.n + 1);
  H_ERR(scaph_copy_d2h_async(off.data(), offsets_.get<uint64_t>() + G.vbase, off.size() * 8,
                             st_->get()));
  st_->sync();
  const uint64_t m = off[G.n] - off[0];
  if (count) *count = m;
  if (offsets)
    for (uint32_t i = 0; i <= G.n; ++i) offsets[i] = off[i] - off[0];
  if (ngh && m) {
    if (cap < m) throw Error(SCAPH_STATUS_INVALID, "batch csr: neighbour buffer too small");
    H_ERR(scaph_copy_d2h_async(ngh, ngh_.get<uint32_t>() + off[0], m * 4, st_->get()));
    st_->sync();
    for (uint64_t i = 0; i < m; ++i) ngh[i] -= (uint32_t)G.vbase;
  }
}

uint64_t Batch::graph_edges(uint32_t g) {
  uint64_t m = 0;
  csr(g, nullptr, nullptr, 0, &m);
  return m;
}

BatchTune Batch::autotune(std::vector<std::pair<BatchTune, double>>* log) {
  BatchTune best;
  double best_ms = 1e300;
  for (uint32_t b : {256u, 512u, 1024u})
    for (uint32_t w : {1u, 2u, 4u, 8u})
      for (uint32_t gr : {0u, 1u}) {
        BatchTune t;
        t.block = b;
        t.wgs_per_cu = w;
        t.use_graph = gr;
        run(t);  // untimed: warms and captures
        double ms = 1e300;
        for (int r = 0; r < 2; ++r) ms = std::min(ms, run(t));
        if (log) log->emplace_back(t, ms);
        if (ms < best_ms) {
          best_ms = ms;
          best = t;
        }
      }
  tuned_ = best;
  return best;
}

}  // namespace scaph
