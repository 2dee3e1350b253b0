// msbfs.cpp — host side of the multi-source bitset BFS (see msbfs.h).
#include "msbfs.h"

#include <algorithm>
#include <cstring>

namespace scaph {

MultiSource::MultiSource(Engine& e, uint32_t max_levels)
    : e_(e), gen_(e.graph_generation()), max_levels_(std::max<uint32_t>(max_levels, 4)) {
  if (!e_.resident() || !e_.has_pull())
    throw Error(SCAPH_STATUS_INVALID,
                "multi-source BFS needs a resident graph with pull pages (-build_pull)");
  const uint32_t n = e_.num_vertices();
  const uint64_t m = e_.num_edges();
  state_.alloc(scaph_ms_state_bytes(n, m, max_levels_));
  H_ERR(scaph_ms_state_carve(state_.get(), n, m, max_levels_, &st_));
  src_.alloc(SCAPH_MS_MAX_QUERIES * 4);
  lev_.alloc((size_t)n * 4);
  hctr_.alloc(((size_t)max_levels_ + 1) * sizeof(scaph_ms_counters));
  hsrc_.alloc(SCAPH_MS_MAX_QUERIES * 4);
}

MultiSource::~MultiSource() { drop_graphs(); }

void MultiSource::drop_graphs() {
  for (auto& s : segs_)
    for (void* g : s.graphs) scaph_graph_destroy(g);
  segs_.clear();
}

void MultiSource::launch_levels(uint32_t first, uint32_t last, const scaph_ms_params& p) {
  void* st = e_.stream();
  for (uint32_t L = first; L < last; ++L)
    H_ERR(scaph_ms_level(&e_.out_desc(), &e_.in_desc(), e_.d_hub_list(), e_.d_degree(), &st_, L,
                         &p, st));
}

MultiSource::Segments& MultiSource::segments(uint32_t count, const MsTune& t,
                                             const scaph_ms_params& p) {
  for (auto& s : segs_)
    if (s.count == count && s.tune == t) return s;
  void* st = e_.stream();
  Segments seg;
  seg.count = count;
  seg.tune = t;
  const uint32_t last = max_levels_ - 1;  // levels 0 .. max_levels-2 exist
  for (uint32_t first = 0; first < last; first += t.segment) {
    const uint32_t end = std::min(last, first + t.segment);
    H_ERR(scaph_graph_capture_begin(st));
    // the source list is read from src_ at replay time
    if (first == 0) H_ERR(scaph_ms_init(&st_, src_.get<uint32_t>(), count, e_.d_degree(), st));
    launch_levels(first, end, p);
    void* g = nullptr;
    H_ERR(scaph_graph_capture_end(st, &g));
    seg.graphs.push_back(g);
  }
  segs_.push_back(std::move(seg));
  return segs_.back();
}

MsStats MultiSource::run(const uint32_t* sources, uint32_t count, const MsTune& t) {
  if (stale()) throw Error(SCAPH_STATUS_INVALID, "graph changed since the multi-source state");
  if (count == 0 || count > SCAPH_MS_MAX_QUERIES)
    throw Error(SCAPH_STATUS_INVALID, "multi-source BFS takes 1..256 sources");
  const uint32_t n = e_.num_vertices();
  for (uint32_t q = 0; q < count; ++q)
    if (sources[q] >= n) throw Error(SCAPH_STATUS_INVALID, "source vertex out of range");
  if (t.push_div == 0 || t.push_mdiv == 0 || t.segment == 0 ||
      (t.pull_block != 256 && t.pull_block != 512 && t.pull_block != 1024))
    throw Error(SCAPH_STATUS_INVALID, "invalid multi-source tuning");
  scaph_ms_params p{};
  p.num_queries = count;
  p.push_div = t.push_div;
  p.push_mdiv = t.push_mdiv;
  p.pull_block = t.pull_block;
  p.pull_grid = t.pull_grid;
  p.num_hubs = e_.num_hubs();
  p.m = e_.num_edges();
  void* st = e_.stream();
  std::memcpy(hsrc_.get(), sources, (size_t)count * 4);
  H_ERR(scaph_copy_h2d_async(src_.get(), hsrc_.get(), (size_t)count * 4, st));
  auto* hc = hctr_.get<scaph_ms_counters>();
  const size_t rec = sizeof(scaph_ms_counters);
  Event e0, e1;
  uint32_t depth = 0;
  if (t.persistent) {
    e0.record(st);
    H_ERR(scaph_ms_init(&st_, src_.get<uint32_t>(), count, e_.d_degree(), st));
    H_ERR(scaph_ms_run_persistent(&e_.out_desc(), &e_.in_desc(), e_.d_hub_list(), e_.d_degree(),
                                  &st_, 0, &p, st));
    e1.record(st);
    const uint32_t last = max_levels_ - 1;
    H_ERR(scaph_copy_d2h_async(hc + last, reinterpret_cast<char*>(st_.ctr) + (size_t)last * rec,
                               rec, st));
    e1.sync();
    if (hc[last].n_active != 0)
      throw Error(SCAPH_STATUS_INVALID, "multi-source BFS deeper than its level buffers");
  } else if (!t.use_graph) {
    e0.record(st);
    H_ERR(scaph_ms_init(&st_, src_.get<uint32_t>(), count, e_.d_degree(), st));
    for (uint32_t L = 0;; ++L) {
      if (L + 1 >= max_levels_)
        throw Error(SCAPH_STATUS_INVALID, "multi-source BFS deeper than its level buffers");
      launch_levels(L, L + 1, p);
      H_ERR(scaph_copy_d2h_async(hc + L + 1, reinterpret_cast<char*>(st_.ctr) + (L + 1) * rec,
                                 rec, st));
      e1.record(st);
      e1.sync();
      if (hc[L + 1].n_active == 0) {
        depth = L + 1;
        break;
      }
    }
  } else {
    Segments& seg = segments(count, t, p);
    e0.record(st);
    const uint32_t last = max_levels_ - 1;
    for (size_t k = 0; k < seg.graphs.size(); ++k) {
      H_ERR(scaph_graph_launch(seg.graphs[k], st));
      const uint32_t end = std::min<uint32_t>(last, (uint32_t)(k + 1) * t.segment);
      H_ERR(scaph_copy_d2h_async(hc + end, reinterpret_cast<char*>(st_.ctr) + (size_t)end * rec,
                                 rec, st));
      e1.record(st);
      e1.sync();
      if (hc[end].n_active == 0) break;
      if (k + 1 == seg.graphs.size())
        throw Error(SCAPH_STATUS_INVALID, "multi-source BFS deeper than its level buffers");
    }
  }
  // the counters of every level, then the depth from them
  H_ERR(scaph_copy_d2h_async(hc, st_.ctr, ((size_t)max_levels_ + 1) * rec, st));
  H_ERR(scaph_stream_sync(st));
  MsStats out;
  out.ms = Event::elapsed_ms(e0, e1);
  out.queries = count;
  if (t.use_graph || t.persistent) {
    depth = 0;
    while (depth < max_levels_ && hc[depth].n_active) ++depth;
  }
  for (uint32_t L = 0; L <= depth && L < max_levels_; ++L)
    if (hc[L].overflow) throw Error(SCAPH_STATUS_INVALID, "multi-source workspace overflow");
  for (uint32_t L = 0; L < depth; ++L) out.pairs += hc[L].bits;
  out.depth = depth;
  depth_ = depth;
  count_ = count;
  return out;
}

void MultiSource::levels(uint32_t query, uint32_t* host_out) {
  if (query >= count_) throw Error(SCAPH_STATUS_INVALID, "no such query in the last run");
  void* st = e_.stream();
  H_ERR(scaph_ms_levels_u32(&st_, depth_, query, lev_.get<uint32_t>(), st));
  H_ERR(scaph_copy_d2h_async(host_out, lev_.get(), (size_t)e_.num_vertices() * 4, st));
  H_ERR(scaph_stream_sync(st));
}

std::vector<scaph_ms_counters> MultiSource::counters() {
  std::vector<scaph_ms_counters> v(depth_ + 1);
  std::memcpy(v.data(), hctr_.get(), v.size() * sizeof(scaph_ms_counters));
  return v;
}

MsTune MultiSource::autotune(const uint32_t* sources, uint32_t count,
                             std::vector<std::pair<MsTune, double>>* log) {
  std::vector<MsTune> cand;
  const uint32_t thr[][2] = {{64, 16}, {256, 64}, {16, 8}};
  scaph_device_props props;
  const uint32_t cus =
      scaph_get_device_props(0, &props) == 0 && props.compute_units > 0 ? props.compute_units : 256;
  for (uint32_t pb : {256u, 512u, 1024u}) {
    // pull grid: one block per vertex block, or as many as stay resident
    uint32_t per_cu = 1;
    H_ERR(scaph_ms_pull_blocks_per_cu(pb, e_.in_desc().num_pages, e_.num_hubs(), &per_cu));
    for (const auto& th : thr)
      for (uint32_t g : {0u, 1u})
        for (uint32_t pg : {0u, per_cu * cus}) {
          MsTune t;
          t.pull_block = pb;
          t.push_div = th[0];
          t.push_mdiv = th[1];
          t.use_graph = g;
          t.pull_grid = pg;
          cand.push_back(t);
        }
  }
  // every level in one cooperative launch, for each direction threshold
  for (const auto& th : thr) {
    MsTune t;
    t.push_div = th[0];
    t.push_mdiv = th[1];
    t.persistent = 1;
    cand.push_back(t);
  }
  MsTune best;
  double best_ms = 1e300;
  for (const MsTune& t : cand) {
    run(sources, count, t);  // untimed: warms caches and captures graphs
    double ms = 1e300;
    for (int r = 0; r < 2; ++r) ms = std::min(ms, run(sources, count, t).ms);
    if (log) log->emplace_back(t, ms);
    if (ms < best_ms) {
      best_ms = ms;
      best = t;
    }
  }
  return best;
}

}  // namespace scaph
