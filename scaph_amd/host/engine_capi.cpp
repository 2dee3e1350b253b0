// engine_capi.cpp — extern "C" wrappers of the host engine
// (include/scaph_engine.h).  Exceptions never cross the boundary.
#include "scaph_engine.h"

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "engine.h"
#include "msbfs.h"
#include "batch.h"
#include "learned.h"
#include "pagehandle.h"

struct scaph_engine {
  scaph::Config cfg;
  scaph::Engine* eng = nullptr;
  scaph::RunStats last;
  scaph::LearnedSchedule learned;
  std::unique_ptr<scaph::MultiSource> ms;
  scaph::MsTune ms_tune;
  ~scaph_engine() {
    ms.reset();
    delete eng;
  }
  scaph::MultiSource& multi() {
    if (!ms || ms->stale()) {
      ms.reset();
      ms.reset(new scaph::MultiSource(*eng, 32));
    }
    return *ms;
  }
};

namespace {

void put_err(char* err, size_t len, const std::string& msg) {
  if (!err || len == 0) return;
  const size_t k = std::min(msg.size(), len - 1);
  std::memcpy(err, msg.data(), k);
  err[k] = '\0';
}

template <typename F>
int guarded(char* err, size_t len, F f) {
  try {
    f();
    put_err(err, len, "");
    return 0;
  } catch (const scaph::Error& e) {
    put_err(err, len, e.what());
    return e.status ? e.status : 1;
  } catch (const std::exception& e) {
    put_err(err, len, e.what());
    return 1;
  }
}

// The learned schedule decides on the device (scaph_mlp_decide inside the
// level loop; the fixed rule takes the first level and uncertain calls);
// the host-side model is the schedule of an engine without a device model
// and serves the CPU-side checks.
// A model read from a file whose features are shares of the graph takes
// the loaded graph's offsets; called after every graph load.
void bind_model(scaph_engine* e) {
  if (!e->learned.ready() || e->cfg.model.empty()) return;
  e->learned.set_graph(e->eng->num_vertices(), e->eng->num_edges());
  e->eng->set_device_model(e->learned.params(), e->learned.graph_mean(), e->learned.scale());
}

void apply_schedule(scaph_engine* e) {
  if (e->cfg.schedule == "learned") {
    scaph::LearnedSchedule* ls = &e->learned;
    scaph::Engine* eng = e->eng;
    eng->set_schedule([ls, eng](const scaph::SchedState& s) { return ls->decide(*eng, s); });
  } else {
    e->eng->set_schedule(nullptr);
  }
}

}  // namespace

extern "C" {

int scaph_engine_create(int argc, const char* const* argv, scaph_engine** out, char* err,
                        size_t errlen) {
  if (!out) return 1;
  *out = nullptr;
  return guarded(err, errlen, [&] {
    std::vector<std::string> args;
    for (int i = 1; i < argc; ++i) args.push_back(argv[i] ? argv[i] : "");
    auto* e = new scaph_engine();
    std::string msg;
    if (!scaph::parse_config(args, &e->cfg, nullptr, &msg)) {
      delete e;
      throw scaph::Error(1, msg);
    }
    try {
      e->eng = new scaph::Engine(e->cfg);
      if (!e->cfg.model.empty()) {
        e->learned.load_file(e->cfg.model);
        e->eng->set_device_model(e->learned.params(), e->learned.mean(), e->learned.scale());
      }
      apply_schedule(e);
    } catch (...) {
      delete e;
      throw;
    }
    *out = e;
  });
}

void scaph_engine_destroy(scaph_engine* e) { delete e; }

int scaph_engine_load(scaph_engine* e, char* err, size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    e->eng->load_from_config();
    bind_model(e);
  });
}

int scaph_engine_load_csr(scaph_engine* e, uint32_t n, const uint64_t* offsets, const uint32_t* ngh,
                          int keep_host, char* err, size_t errlen) {
  if (!e || !offsets) return 1;
  return guarded(err, errlen, [&] {
    scaph::HostCsr g;
    g.n = n;
    g.offsets.assign(offsets, offsets + (size_t)n + 1);
    const uint64_t m = g.offsets[n];
    if (m && !ngh) throw scaph::Error(1, "null neighbour array");
    g.ngh.assign(ngh, ngh + m);
    for (uint64_t i = 0; i < m; ++i)
      if (g.ngh[i] >= n) throw scaph::Error(1, "neighbour id out of range");
    for (uint32_t v = 0; v < n; ++v)
      if (g.offsets[v] > g.offsets[(size_t)v + 1]) throw scaph::Error(1, "offsets decrease");
    e->eng->load_csr(std::move(g), keep_host != 0);
    bind_model(e);
  });
}

int scaph_engine_gen_rmat_device(scaph_engine* e, uint32_t scale, uint32_t edge_factor, double a,
                                 double b, double c, uint64_t seed, int keep_host, char* err,
                                 size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    scaph::RmatParams p;
    p.scale = scale;
    p.edge_factor = edge_factor;
    p.a = a;
    p.b = b;
    p.c = c;
    p.seed = seed;
    e->eng->gen_rmat_device(p, keep_host != 0);
    bind_model(e);
  });
}

int scaph_engine_gen_er_device(scaph_engine* e, uint32_t n, uint64_t m, uint64_t seed,
                               int keep_host, char* err, size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    e->eng->gen_er_device(n, m, seed, keep_host != 0);
    bind_model(e);
  });
}

int scaph_engine_gen_zipf(scaph_engine* e, uint32_t n, double x_min, double alpha, uint32_t cap,
                          uint64_t seed, char* err, size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    e->eng->gen_zipf(n, x_min, alpha, cap, seed);
    bind_model(e);
  });
}

int scaph_engine_info(scaph_engine* e, uint32_t* n, uint64_t* m, uint32_t* num_pages, int* resident,
                      int* has_pull) {
  if (!e) return 1;
  if (n) *n = e->eng->num_vertices();
  if (m) *m = e->eng->num_edges();
  if (num_pages) *num_pages = (uint32_t)e->eng->pages().size();
  if (resident) *resident = e->eng->resident() ? 1 : 0;
  if (has_pull) *has_pull = e->eng->has_pull() ? 1 : 0;
  return 0;
}

int scaph_engine_set_page_levels(scaph_engine* e, const uint32_t* levels, uint32_t n, char* err,
                                 size_t errlen) {
  if (!e || !levels) return 1;
  return guarded(err, errlen, [&] {
    if (n != e->eng->num_vertices())
      throw scaph::Error(SCAPH_STATUS_INVALID, "one level per vertex expected");
    e->eng->set_page_levels(levels);
  });
}

int scaph_engine_pool(scaph_engine* e, int64_t* slots, uint64_t* device_bytes) {
  if (!e) return 1;
  if (slots) *slots = e->eng->pool_size();
  if (device_bytes) *device_bytes = e->eng->device_bytes();
  return 0;
}

int scaph_engine_bfs(scaph_engine* e, uint32_t src, scaph_bfs_stats* out, char* err,
                     size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    e->last = e->eng->bfs(src);
    if (out) {
      std::memset(out, 0, sizeof(*out));
      out->src = e->last.src;
      out->depth = e->last.depth;
      out->ms = e->last.ms;
      out->n_reached = e->last.n_reached;
      out->m_reached = e->last.m_reached;
      out->page_execs = e->last.page_execs;
      out->page_execs_cached = e->last.page_execs_cached;
      out->pages_full = e->last.pages_full;
      out->pages_compacted = e->last.pages_compacted;
      out->bytes_h2d = e->last.bytes_h2d;
      out->h2d_ms = e->last.h2d_ms;
      out->pages_prefetched = e->last.pages_prefetched;
      out->pages_direct = e->last.pages_direct;
    }
  });
}

namespace {
void fill_stats(const scaph::RunStats& r, scaph_bfs_stats* out) {
  std::memset(out, 0, sizeof(*out));
  out->src = r.src;
  out->depth = r.depth;
  out->ms = r.ms;
  out->n_reached = r.n_reached;
  out->m_reached = r.m_reached;
  out->page_execs = r.page_execs;
  out->page_execs_cached = r.page_execs_cached;
  out->pages_full = r.pages_full;
  out->pages_compacted = r.pages_compacted;
  out->bytes_h2d = r.bytes_h2d;
  out->h2d_ms = r.h2d_ms;
  out->pages_prefetched = r.pages_prefetched;
  out->pages_direct = r.pages_direct;
}
}  // namespace

int scaph_engine_set_weights(scaph_engine* e, const uint32_t* wgh, uint64_t count, char* err,
                             size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] { e->eng->set_weights(wgh, count); });
}

int scaph_engine_wcc(scaph_engine* e, scaph_bfs_stats* out, char* err, size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    e->last = e->eng->wcc();
    if (out) fill_stats(e->last, out);
  });
}

int scaph_engine_sssp(scaph_engine* e, uint32_t src, scaph_bfs_stats* out, char* err,
                      size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    e->last = e->eng->sssp(src);
    if (out) fill_stats(e->last, out);
  });
}

int scaph_engine_set_schedule(scaph_engine* e, const char* name, double alpha, double beta,
                              char* err, size_t errlen) {
  if (!e || !name) return 1;
  return guarded(err, errlen, [&] {
    const std::string s(name);
    if (s != "fixed" && s != "push" && s != "pull" && s != "learned")
      throw scaph::Error(1, "schedule must be fixed, push, pull or learned");
    if (s == "learned" && !e->learned.ready())
      throw scaph::Error(1, "learned schedule has no model (scaph_engine_set_model / -model)");
    if (alpha > 0) e->cfg.alpha = alpha;
    if (beta > 0) e->cfg.beta = beta;
    e->cfg.schedule = s;
    e->eng->set_config_schedule(s, e->cfg.alpha, e->cfg.beta);
    apply_schedule(e);
  });
}

int scaph_engine_set_model(scaph_engine* e, const uint16_t* params_bf16, const float* feat_mean,
                           const float* feat_scale, float prefetch_gain, char* err,
                           size_t errlen) {
  if (!e || !params_bf16 || !feat_mean || !feat_scale) return 1;
  return guarded(err, errlen, [&] {
    e->learned.set(params_bf16, feat_mean, feat_scale, prefetch_gain);
    e->eng->set_device_model(params_bf16, feat_mean, feat_scale);
  });
}

int scaph_engine_level_records(scaph_engine* e, scaph_level_record* out, uint32_t max,
                               uint32_t* count) {
  if (!e) return 1;
  const auto& lv = e->last.levels;
  const uint32_t k = (uint32_t)std::min<size_t>(lv.size(), max);
  for (uint32_t i = 0; i < k && out; ++i) {
    out[i].level = lv[i].level;
    out[i].dir = lv[i].dir;
    out[i].n_f = lv[i].n_f;
    out[i].m_f = lv[i].m_f;
    out[i].m_u = lv[i].m_u;
    out[i].n_next = lv[i].n_next;
    out[i].m_next = lv[i].m_next;
    out[i].scanned = lv[i].scanned;
    out[i].examined = lv[i].examined;
    out[i].ms = lv[i].ms;
    out[i].kernel_ms = lv[i].kernel_ms;
    for (int j = 0; j < 12; ++j) out[i].page_hist[j] = lv[i].page_hist[j];
    out[i].pages_active = lv[i].pages_active;
    out[i].pages_cached = lv[i].pages_cached;
    out[i].pages_full = lv[i].pages_full;
    out[i].pages_compacted = lv[i].pages_compacted;
    out[i].pages_prefetched = lv[i].pages_prefetched;
    out[i].pages_direct = lv[i].pages_direct;
    out[i].bytes_h2d = lv[i].bytes_h2d;
    out[i].h2d_ms = lv[i].h2d_ms;
    out[i].logit = lv[i].logit;
    out[i].sparse = lv[i].sparse ? 1u : 0u;
    out[i].pad = 0;
  }
  if (count) *count = (uint32_t)lv.size();
  return 0;
}

int scaph_engine_explore(scaph_engine* e, uint32_t src, int reps, scaph_explore_record* out,
                         uint32_t max, uint32_t* count, char* err, size_t errlen) {
  if (!e) return 1;
  return guarded(err, errlen, [&] {
    const auto recs = e->eng->explore(src, reps);
    const uint32_t k = (uint32_t)std::min<size_t>(recs.size(), max);
    for (uint32_t i = 0; i < k && out; ++i) {
      out[i].n_f = recs[i].n_f;
      out[i].m_f = recs[i].m_f;
      out[i].n_next = recs[i].n_next;
      out[i].level = recs[i].level;
      out[i].t_push_ms = recs[i].t_push_ms;
      out[i].t_pull_ms = recs[i].t_pull_ms;
      out[i].pages_active_next = recs[i].pages_active_next;
      for (int j = 0; j < 16; ++j) out[i].x[j] = recs[i].x[j];
    }
    if (count) *count = (uint32_t)recs.size();
  });
}

namespace {
scaph::MsTune from_c(const scaph_ms_tune& t) {
  scaph::MsTune o;
  o.push_div = t.push_div;
  o.push_mdiv = t.push_mdiv;
  o.pull_block = t.pull_block;
  o.use_graph = t.use_graph;
  o.segment = t.segment ? t.segment : 8;
  o.pull_grid = t.pull_grid;
  o.persistent = t.persistent;
  return o;
}
scaph_ms_tune to_c(const scaph::MsTune& t) {
  scaph_ms_tune o;
  o.push_div = t.push_div;
  o.push_mdiv = t.push_mdiv;
  o.pull_block = t.pull_block;
  o.use_graph = t.use_graph;
  o.segment = t.segment;
  o.pull_grid = t.pull_grid;
  o.persistent = t.persistent;
  return o;
}
}  // namespace

int scaph_engine_msbfs(scaph_engine* e, const uint32_t* sources, uint32_t count,
                       const scaph_ms_tune* tune, scaph_ms_stats* out, char* err, size_t errlen) {
  if (!e || !sources) return 1;
  return guarded(err, errlen, [&] {
    const scaph::MsTune t = tune ? from_c(*tune) : e->ms_tune;
    const scaph::MsStats s = e->multi().run(sources, count, t);
    if (out) {
      out->ms = s.ms;
      out->depth = s.depth;
      out->queries = s.queries;
      out->pairs = s.pairs;
    }
  });
}

int scaph_engine_msbfs_levels(scaph_engine* e, uint32_t query, uint32_t* host_out, char* err,
                              size_t errlen) {
  if (!e || !host_out) return 1;
  return guarded(err, errlen, [&] {
    if (!e->ms || e->ms->stale()) throw scaph::Error(SCAPH_STATUS_INVALID, "no multi-source run");
    e->ms->levels(query, host_out);
  });
}

int scaph_engine_msbfs_counters(scaph_engine* e, scaph_ms_counters* out, uint32_t max,
                                uint32_t* count) {
  if (!e) return 1;
  return guarded(nullptr, 0, [&] {
    if (!e->ms || e->ms->stale()) throw scaph::Error(SCAPH_STATUS_INVALID, "no multi-source run");
    const auto v = e->ms->counters();
    for (uint32_t i = 0; i < v.size() && i < max && out; ++i) out[i] = v[i];
    if (count) *count = (uint32_t)v.size();
  });
}

int scaph_engine_msbfs_autotune(scaph_engine* e, const uint32_t* sources, uint32_t count,
                                scaph_ms_tune* best, double* best_ms, scaph_ms_tune* log_tunes,
                                double* log_ms, uint32_t log_max, uint32_t* log_count,
                                char* err, size_t errlen) {
  if (!e || !sources) return 1;
  return guarded(err, errlen, [&] {
    std::vector<std::pair<scaph::MsTune, double>> log;
    const scaph::MsTune t = e->multi().autotune(sources, count, &log);
    e->ms_tune = t;
    double bm = 0;
    for (const auto& x : log)
      if (x.first == t) bm = x.second;
    if (best) *best = to_c(t);
    if (best_ms) *best_ms = bm;
    for (uint32_t i = 0; i < log.size() && i < log_max; ++i) {
      if (log_tunes) log_tunes[i] = to_c(log[i].first);
      if (log_ms) log_ms[i] = log[i].second;
    }
    if (log_count) *log_count = (uint32_t)log.size();
  });
}

int scaph_engine_levels(scaph_engine* e, uint32_t* host_out) {
  if (!e || !host_out) return 1;
  return guarded(nullptr, 0, [&] { e->eng->levels_to_host(host_out); });
}

int scaph_engine_validate(scaph_engine* e, uint32_t src, uint64_t* out2) {
  if (!e || !out2) return 1;
  return guarded(nullptr, 0, [&] {
    auto r = e->eng->validate(src);
    out2[0] = r[0];
    out2[1] = r[1];
  });
}

int scaph_engine_out_degrees(scaph_engine* e, uint32_t* host_out) {
  if (!e || !host_out) return 1;
  return guarded(nullptr, 0, [&] {
    const uint32_t n = e->eng->num_vertices();
    H_ERR(scaph_copy_d2h_async(host_out, e->eng->d_degree(), (size_t)n * 4, e->eng->stream()));
    H_ERR(scaph_stream_sync(e->eng->stream()));
  });
}

int scaph_engine_host_csr(scaph_engine* e, uint64_t* offsets, uint32_t* ngh) {
  if (!e) return 1;
  const scaph::HostCsr* g = e->eng->host_csr();
  if (!g) return 1;
  if (offsets) std::memcpy(offsets, g->offsets.data(), g->offsets.size() * 8);
  if (ngh) std::memcpy(ngh, g->ngh.data(), g->ngh.size() * 4);
  return 0;
}

int scaph_engine_check(scaph_engine* e, uint32_t src, uint64_t* errors, char* err, size_t errlen) {
  if (!e || !errors) return 1;
  return guarded(err, errlen, [&] {
    const scaph::HostCsr* g = e->eng->host_csr();
    if (!g) throw scaph::Error(1, "no host copy of the graph to check against");
    const uint32_t n = e->eng->num_vertices();
    std::vector<uint32_t> gpu(n), capped, exact;
    e->eng->levels_to_host(gpu.data());
    scaph::cpu_bfs(*g, src, 63u, &capped);
    scaph::cpu_bfs(*g, src, SCAPH_UNREACHED, &exact);
    uint64_t bad = 0;
    for (uint32_t v = 0; v < n; ++v) {
      if (scaph::compat63(gpu[v]) != capped[v]) ++bad;
      if (gpu[v] != exact[v]) ++bad;
    }
    *errors = bad;
  });
}

int scaph_engine_graph_desc(scaph_engine* e, int transposed, void* desc_out) {
  if (!e || !desc_out) return 1;
  const scaph_graph_desc& d = transposed ? e->eng->in_desc() : e->eng->out_desc();
  std::memcpy(desc_out, &d, sizeof(d));
  return 0;
}

int scaph_engine_device_degree(scaph_engine* e, const uint32_t** degree) {
  if (!e || !degree) return 1;
  *degree = e->eng->d_degree();
  return 0;
}

int scaph_host_plan_pool(int64_t max_memory_mib, uint32_t n, uint32_t num_pages,
                         uint64_t pagesize_words, int64_t* slots) {
  if (!slots || pagesize_words == 0) return 1;
  *slots = scaph::plan_pool_slots(max_memory_mib, n, num_pages, pagesize_words);
  return 0;
}

int scaph_host_page_cuts(const uint64_t* offsets, uint32_t n, uint64_t pagesize_words, int weighted,
                         void* pages_out, uint32_t max_pages, uint32_t* num_pages) {
  if (!offsets || !num_pages) return 1;
  std::vector<scaph_page> pages;
  std::string err;
  if (!scaph::page_cuts_raw(offsets, n, pagesize_words, weighted != 0, &pages, &err)) return 2;
  *num_pages = (uint32_t)pages.size();
  if (pages_out) {
    if (pages.size() > max_pages) return 3;
    std::memcpy(pages_out, pages.data(), pages.size() * sizeof(scaph_page));
  }
  return 0;
}

static int copy_csr_out(const scaph::HostCsr& g, uint64_t* offsets, uint32_t* ngh, uint64_t cap,
                        uint64_t* num_edges) {
  if (num_edges) *num_edges = g.num_edges();
  if (offsets) std::memcpy(offsets, g.offsets.data(), g.offsets.size() * 8);
  if (ngh) {
    if (g.num_edges() > cap) return 3;
    std::memcpy(ngh, g.ngh.data(), g.ngh.size() * 4);
  }
  return 0;
}

int scaph_host_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint64_t* offsets, uint32_t* ngh, uint64_t ngh_cap,
                        uint64_t* num_edges) {
  if (scale == 0 || scale > 31) return 1;
  scaph::RmatParams p;
  p.scale = scale;
  p.edge_factor = edge_factor;
  p.a = a;
  p.b = b;
  p.c = c;
  p.seed = seed;
  return copy_csr_out(scaph::generate_rmat(p), offsets, ngh, ngh_cap, num_edges);
}

int scaph_host_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t* offsets, uint32_t* ngh,
                      uint64_t ngh_cap, uint64_t* num_edges) {
  if (n == 0) return 1;
  return copy_csr_out(scaph::generate_er(n, m, seed), offsets, ngh, ngh_cap, num_edges);
}

int scaph_host_write_graph(const char* base, uint32_t n, const uint64_t* offsets,
                           const uint32_t* ngh, uint64_t pagesize_words, int idx64, int write_wgh,
                           char* err, size_t errlen) {
  if (!base || !offsets) return 1;
  scaph::HostCsr g;
  g.n = n;
  g.offsets.assign(offsets, offsets + (size_t)n + 1);
  g.ngh.assign(ngh, ngh + g.offsets[n]);
  std::string msg;
  if (!scaph::write_csr_files(base, g, idx64 != 0, write_wgh != 0, &msg) ||
      (pagesize_words && !scaph::write_page_files(base, g, pagesize_words, false, nullptr, &msg))) {
    put_err(err, errlen, msg);
    return 2;
  }
  put_err(err, errlen, "");
  return 0;
}

int scaph_host_cpu_bfs(uint32_t n, const uint64_t* offsets, const uint32_t* ngh, uint32_t src,
                       uint32_t inf, uint32_t* levels) {
  if (!offsets || !levels || (offsets[n] && !ngh)) return 1;
  scaph::HostCsr g;
  g.n = n;
  g.offsets.assign(offsets, offsets + (size_t)n + 1);
  g.ngh.assign(ngh, ngh + offsets[n]);
  std::vector<uint32_t> out;
  scaph::cpu_bfs(g, src, inf, &out);
  std::memcpy(levels, out.data(), (size_t)n * 4);
  return 0;
}

struct scaph_batch {
  std::unique_ptr<scaph::Batch> b;
  bool tuned = false;
};

int scaph_batch_create_rmat(uint32_t count, uint32_t min_scale, uint32_t max_scale,
                            uint32_t min_ef, uint32_t max_ef, uint64_t seed, int device,
                            scaph_batch** out, char* err, size_t errlen) {
  if (!out) return 1;
  *out = nullptr;
  return guarded(err, errlen, [&] {
    std::unique_ptr<scaph_batch> h(new scaph_batch());
    h->b.reset(new scaph::Batch(count, min_scale, max_scale, min_ef, max_ef, seed, device));
    *out = h.release();
  });
}

void scaph_batch_destroy(scaph_batch* b) { delete b; }

int scaph_batch_info(scaph_batch* b, uint32_t* count, unsigned long long* vertices,
                     unsigned long long* edges) {
  if (!b) return 1;
  if (count) *count = b->b->count();
  if (vertices) *vertices = b->b->vertices();
  if (edges) *edges = b->b->edges();
  return 0;
}

int scaph_batch_graph_info(scaph_batch* b, uint32_t g, scaph_batch_entry* out) {
  if (!b || !out || g >= b->b->count()) return 1;
  *out = b->b->graph(g);
  return 0;
}

int scaph_batch_run(scaph_batch* b, const scaph_batch_tune* tune, double* ms, char* err,
                    size_t errlen) {
  if (!b) return 1;
  return guarded(err, errlen, [&] {
    scaph::BatchTune t = b->b->tuned();
    if (tune) {
      t.block = tune->block;
      t.wgs_per_cu = tune->wgs_per_cu;
      t.use_graph = tune->use_graph;
    }
    const double v = b->b->run(t);
    if (ms) *ms = v;
  });
}

int scaph_batch_levels(scaph_batch* b, uint32_t g, uint32_t* host_out) {
  if (!b || !host_out || g >= b->b->count()) return 1;
  return guarded(nullptr, 0, [&] { b->b->levels(g, host_out); });
}

int scaph_batch_csr(scaph_batch* b, uint32_t g, uint64_t* offsets, uint32_t* ngh,
                    unsigned long long cap, unsigned long long* num_edges) {
  if (!b || g >= b->b->count()) return 1;
  return guarded(nullptr, 0, [&] {
    uint64_t m = 0;
    b->b->csr(g, offsets, ngh, cap, &m);
    if (num_edges) *num_edges = m;
  });
}

int scaph_batch_autotune(scaph_batch* b, scaph_batch_tune* best, double* best_ms,
                         scaph_batch_tune* log_tunes, double* log_ms, uint32_t log_max,
                         uint32_t* log_count, char* err, size_t errlen) {
  if (!b) return 1;
  return guarded(err, errlen, [&] {
    std::vector<std::pair<scaph::BatchTune, double>> log;
    const scaph::BatchTune t = b->b->autotune(&log);
    for (uint32_t i = 0; i < log.size() && i < log_max; ++i) {
      if (log_tunes) log_tunes[i] = {log[i].first.block, log[i].first.wgs_per_cu,
                                     log[i].first.use_graph};
      if (log_ms) log_ms[i] = log[i].second;
      if (best_ms && log[i].first == t) *best_ms = log[i].second;
    }
    if (best) *best = {t.block, t.wgs_per_cu, t.use_graph};
    if (log_count) *log_count = (uint32_t)log.size();
  });
}

}  // extern "C"

// Chunk masks of the page-slot cache (src/tools.h:10-85).
extern "C" int scaph_host_max_zero_run(uint32_t mask) { return scaph::max_zero_run(mask); }
extern "C" int scaph_host_fit_zero_run(uint32_t mask, int required) {
  return scaph::fit_zero_run(mask, required);
}
extern "C" uint32_t scaph_host_set_run(uint32_t mask, int off, int number) {
  return scaph::set_run(mask, off, number);
}
extern "C" uint32_t scaph_host_clear_run(uint32_t mask, int off, int number) {
  return scaph::clear_run(mask, off, number);
}
