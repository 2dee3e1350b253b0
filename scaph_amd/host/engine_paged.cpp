// engine_paged.cpp — out-of-core (paged) BFS.  A bounded pool of device
// page slots caches the pages held in pinned host memory; every level the
// host classifies the active pages (cached / full copy / compacted copy,
// src/pagehandle.h:326-379), compacts the cold ones on worker threads
// (src/compress.h:42-56) and streams copies on a side stream while the
// exec streams run the push kernel page by page (src/app.cuh:395-468).
// Levels are synchronous: the frontier of level L is a bitmap, each page
// runs once per level (more with -multi, which re-runs bound pages), and a
// vertex is claimed by its first edge exactly as in the resident mode.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <thread>

#include "engine_internal.h"

namespace scaph {

namespace {

size_t bitmap_words(uint32_t n) { return (((size_t)n + 63) / 64) * 2; }

// Out-edges below which splitting a level's cached pages into more grouped
// launches costs more in launches than the slots it frees early.
constexpr uint64_t kMinGroupEdges = 1ull << 22;

// Calls f(i) for each Node i of page pg (vertex pg.left + i) whose frontier
// bit is set, in vertex order, skipping zero bitmap words.
template <typename F>
void for_active(const scaph_page& pg, const uint32_t* front, F f) {
  if (pg.nodenum == 0) return;
  const uint64_t first = pg.left, last = (uint64_t)pg.left + pg.nodenum - 1;
  for (uint64_t w = first >> 5; w <= last >> 5; ++w) {
    uint32_t bits = front[w];
    if (w == first >> 5) bits &= ~0u << (first & 31);
    if (w == last >> 5 && (last & 31) != 31) bits &= (2u << (last & 31)) - 1u;
    while (bits) {
      const uint32_t b = (uint32_t)__builtin_ctz(bits);
      bits &= bits - 1;
      f((uint32_t)(w * 32 + b - first));
    }
  }
}

// Active Node records of one page plus their adjacency, re-indexed: the
// layout compress_stat writes (src/compress.h:18-40).  Returns words used.
uint64_t compact_page(const uint32_t* page, const scaph_page& pg, const uint32_t* front,
                      uint32_t active_nodes, uint32_t* out) {
  const uint32_t* edges = page + 3ull * pg.nodenum;
  uint32_t* eout = out + 3ull * active_nodes;
  uint64_t off = 0;
  uint32_t cnt = 0;
  for_active(pg, front, [&](uint32_t i) {
    const uint32_t v = pg.left + i;
    const uint32_t idx = page[3ull * i + 1], len = page[3ull * i + 2];
    if (cnt < active_nodes) {
      out[3ull * cnt] = v;
      out[3ull * cnt + 1] = (uint32_t)off;
      out[3ull * cnt + 2] = len;
      std::memcpy(eout + off, edges + idx, (size_t)len * 4);
    }
    off += len;
    ++cnt;
  });
  if (cnt != active_nodes) throw Error(SCAPH_STATUS_INVALID, "compaction count mismatch");
  return 3ull * active_nodes + off;
}

// The weights of compact_page's output at the same word offsets (the Node
// header words stay unused): wpage holds the page's weights in edge order.
void compact_weights(const uint32_t* page, const scaph_page& pg, const uint32_t* wpage,
                     const uint32_t* front, uint32_t active_nodes, uint32_t* out) {
  uint32_t* wout = out + 3ull * active_nodes;
  uint64_t off = 0;
  for_active(pg, front, [&](uint32_t i) {
    const uint32_t idx = page[3ull * i + 1], len = page[3ull * i + 2];
    std::memcpy(wout + off, wpage + idx, (size_t)len * 4);
    off += len;
  });
}

// The compaction workers of one level, joined on every exit from it: an
// error raised by the scheduling loop then reaches the caller instead of
// destroying joinable threads (std::terminate).
struct Workers {
  std::vector<std::thread> t;
  void join() {
    for (auto& x : t)
      if (x.joinable()) x.join();
  }
  ~Workers() { join(); }
};

}  // namespace

void Engine::setup_paged() {
  paged_.reset(new PagedState());
  PagedState& ps = *paged_;
  const size_t np = pages_out_.size();
  ps.chunk = P_ / 32;
  ps.thres = (uint64_t)cfg_.threshold * ps.chunk;
  // pinned host pages (src/graph.h:96-104); the compaction area is sized by
  // the threshold instead of the reference's fixed half page
  ps.h_page.resize(np);
  ps.h_tail.resize(np);
  for (size_t p = 0; p < np; ++p) {
    const uint64_t used = page_used_words(pages_out_[p], false);
    ps.h_page[p].alloc(std::max<uint64_t>(used, 1) * 4);
    fill_page(host_, pages_out_[p], false, ps.h_page[p].get<uint32_t>());
    if (ps.thres) ps.h_tail[p].alloc(ps.thres * 4);
  }
  ps.d_hpage.assign(np, 0);
  if (cfg_.page_direct > 0 && cfg_.page_groups > 0)
    for (size_t p = 0; p < np; ++p) {
      void* d = nullptr;
      H_ERR(scaph_host_device_pointer(ps.h_page[p].get(), &d));
      ps.d_hpage[p] = reinterpret_cast<unsigned long long>(d);
    }
  ps.pool.alloc((size_t)pool_size_ * P_ * 4);
  ps.copy.reset(new Stream((uint32_t)cfg_.tricknum));
  ps.copy_beg.reset(new Event());
  ps.copy_end.reset(new Event());
  const int ns = cfg_.num_stream;
  uint32_t max_nodes = 1;
  for (const auto& pg : pages_out_) max_nodes = std::max(max_nodes, pg.nodenum);
  ps.ws_bytes = scaph_push_workspace_bytes(max_nodes, P_);
  for (int s = 0; s < ns; ++s) {
    ps.exec.emplace_back(new Stream((uint32_t)cfg_.tricknum));
    ps.exec_done.emplace_back(new Event());
    ps.ws.emplace_back(ps.ws_bytes);
  }
  if (cfg_.page_groups > 0) {
    // one workspace for the group stream: every segment of the graph, up to
    // 16 M active Nodes per group, and tiles for all edges
    uint64_t segs = 0;
    for (const auto& pg : pages_out_) segs += (pg.nodenum + SCAPH_RUN_SEGMENT - 1) / SCAPH_RUN_SEGMENT;
    ps.group_nodes_cap = std::max<uint64_t>(std::min<uint64_t>(n_, 16ull << 20), max_nodes);
    ps.group_ws_bytes = scaph_push_runs_workspace_bytes((uint32_t)segs, ps.group_nodes_cap,
                                                        std::max<uint64_t>(m_, 1));
    ps.group_ws.alloc(ps.group_ws_bytes);
    ps.group.reset(new Stream((uint32_t)cfg_.tricknum));
    ps.group_done.reset(new Event());
    ps.h_runs.alloc(np * sizeof(scaph_page_run) + 64);
    ps.d_runs.alloc(np * sizeof(scaph_page_run) + 64);
  }
  ps.d_unv.alloc(np * 4 + 64);
  ps.h_unv.alloc(np * 4 + 64);
  if (cfg_.page_predict == "learned") {
    // vertices with in-edges: only they can become active after the source
    const size_t words = bitmap_words(n_);
    std::vector<std::atomic<uint32_t>> live(words);
    for (auto& w : live) w.store(0, std::memory_order_relaxed);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    const uint64_t m = host_.num_edges();
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        for (uint64_t k = m * t / nt; k < m * (t + 1) / nt; ++k) {
          const uint32_t v = host_.ngh[k];
          const uint32_t bit = 1u << (v & 31);
          if (!(live[v >> 5].load(std::memory_order_relaxed) & bit))
            live[v >> 5].fetch_or(bit, std::memory_order_relaxed);
        }
      });
    for (auto& x : th) x.join();
    std::vector<uint32_t> h(words);
    for (size_t w = 0; w < words; ++w) h[w] = live[w].load(std::memory_order_relaxed);
    ps.live.alloc(words * 4);
    H_ERR(scaph_copy_h2d_async(ps.live.get(), h.data(), words * 4, main_->get()));
    main_->sync();
  }
  ps.h_nodes.alloc(np * 4 + 64);
  ps.h_datas.alloc(np * 4 + 64);
  ps.h_totals.alloc(64);
  ps.h_front.alloc(bitmap_words(n_) * 4);
  // device page list for the count kernel
  dpages_out_.alloc(std::max<size_t>(np, 1) * sizeof(scaph_page));
  H_ERR(scaph_copy_h2d_async(dpages_out_.get(), pages_out_.data(), np * sizeof(scaph_page),
                             main_->get()));
  // initial pool fill: the first pool_size pages (src/app.cuh:330-339)
  for (int d = 0; d < pool_size_; ++d) {
    const uint64_t used = page_used_words(pages_out_[d], false);
    H_ERR(scaph_copy_h2d_async(ps.pool.get<uint32_t>() + (uint64_t)d * P_,
                               ps.h_page[d].get(), used * 4, ps.copy->get()));
  }
  ps.copy->sync();
  main_->sync();
  ps.handle.init((int)np, pool_size_, pages_out_, ps.thres, cfg_.multi);
}

RunStats Engine::bfs_paged(uint32_t src) {
  PagedState& ps = *paged_;
  void* st = main_->get();
  RunStats rs;
  rs.src = src;
  const size_t np = pages_out_.size();
  const size_t words = bitmap_words(n_);
  auto* cnt = counters_.get<scaph_level_counters>();
  auto* hc = hcount_.get<scaph_level_counters>();
  uint32_t* level = level_.get<uint32_t>();
  uint32_t* visited = visited_.get<uint32_t>();
  const uint32_t* deg = deg_out_.get<uint32_t>();
  const auto t0 = std::chrono::steady_clock::now();

  H_ERR(scaph_bfs_init(level, visited, nullptr, n_, src, nullptr, cnt, st));
  int fa = 0;
  // level-0 frontier bitmap: the source alone
  *hcount_.get<uint32_t>() = src;
  H_ERR(scaph_copy_h2d_async(queue_[0].get(), hcount_.get(), 4, st));
  H_ERR(scaph_frontier_queue_to_bitmap(queue_[0].get<uint32_t>(), 1, front_[0].get<uint32_t>(), n_,
                                       st));
  main_->sync();
  uint64_t n_f = 1, m_f = out_degree(src), m_u = m_ - m_f;
  rs.n_reached = 1;
  rs.m_reached = m_f;
  uint32_t cur = 0;
  while (n_f > 0) {
    const auto tl = std::chrono::steady_clock::now();
    // get_page_info / get_total_info + D2H (src/app.cuh:188-207)
    H_ERR(scaph_page_counts(dpages_out_.get<scaph_page>(), (uint32_t)np, front_[fa].get<uint32_t>(),
                            deg, 0, page_nodes_.get<uint32_t>(), page_datas_.get<uint32_t>(),
                            totals_.get<unsigned long long>(), st));
    H_ERR(scaph_copy_d2h_async(ps.h_nodes.get(), page_nodes_.get(), np * 4, st));
    H_ERR(scaph_copy_d2h_async(ps.h_datas.get(), page_datas_.get(), np * 4, st));
    H_ERR(scaph_copy_d2h_async(ps.h_totals.get(), totals_.get(), 16, st));
    H_ERR(scaph_fill_u32(front_[fa ^ 1].get<uint32_t>(), 0u, words, st));
    H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt), 0u, sizeof(scaph_level_counters) / 4, st));
    const bool learned_pages = predict_launch(cur, n_f, m_f, m_u);
    main_->sync();
    const uint32_t* nodes = ps.h_nodes.get<uint32_t>();
    if (ps.h_totals.get<unsigned long long>()[1] == 0) break;
    LevelRecord rec;
    rec.level = cur;
    rec.dir = 0;
    rec.n_f = n_f;
    rec.m_f = m_f;
    rec.m_u = m_u;
    page_histogram(nodes, &rec);

    predict_apply(cur, learned_pages);
    const uint64_t full0 = rs.pages_full, comp0 = rs.pages_compacted, pre0 = rs.pages_prefetched;
    const uint64_t dir0 = rs.pages_direct;
    const uint64_t h2d0 = rs.bytes_h2d;
    const double h2d_ms0 = rs.h2d_ms;
    for (size_t p = 0; p < np; ++p) {
      if (!nodes[p]) continue;
      ++rec.pages_active;
      if (ps.handle.vp[p].cached) ++rec.pages_cached;
    }
    schedule_pages(fa, cur, nullptr, &rs);
    rec.pages_full = (uint32_t)(rs.pages_full - full0);
    rec.pages_compacted = (uint32_t)(rs.pages_compacted - comp0);
    rec.pages_prefetched = (uint32_t)(rs.pages_prefetched - pre0);
    rec.pages_direct = (uint32_t)(rs.pages_direct - dir0);
    rec.bytes_h2d = rs.bytes_h2d - h2d0;
    rec.h2d_ms = (float)(rs.h2d_ms - h2d_ms0);
    H_ERR(scaph_copy_d2h_async(hc, cnt, sizeof(scaph_level_counters), st));
    main_->sync();
    rec.n_next = hc->n_next;
    rec.m_next = hc->m_next;
    rec.scanned = hc->scanned;
    rec.ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tl)
                 .count();
    rec.kernel_ms = rec.ms;
    rs.levels.push_back(rec);
    n_f = hc->n_next;
    m_f = hc->m_next;
    m_u = m_u > m_f ? m_u - m_f : 0;
    rs.n_reached += n_f;
    rs.m_reached += m_f;
    fa ^= 1;
    ++cur;
  }
  main_->sync();
  rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rs.depth = cur;
  return rs;
}

// One level (or relax round) over the active pages of front_[fa], whose
// counts are in ps.h_nodes / ps.h_datas: classification (src/pagehandle.h:
// 326-379), compaction of the cold pages on worker threads, copies on the
// copy stream and page launches on the exec (and group) streams
// (gpu_push_batch, src/app.cuh:395-468).  `relax` selects the WCC / SSSP
// relax kernel instead of the BFS push.
void Engine::schedule_pages(int fa, uint32_t cur, const RelaxRound* relax, RunStats* rs) {
  PagedState& ps = *paged_;
  const size_t np = pages_out_.size();
  const size_t words = bitmap_words(n_);
  auto* cnt = counters_.get<scaph_level_counters>();
  uint32_t* level = level_.get<uint32_t>();
  uint32_t* visited = visited_.get<uint32_t>();
  const uint32_t* deg = deg_out_.get<uint32_t>();
  void* st = main_->get();
  const int ns = (int)ps.exec.size();
  const uint32_t* nodes = ps.h_nodes.get<uint32_t>();
  const uint32_t* datas = ps.h_datas.get<uint32_t>();
  // cold pages need the frontier on the host for compaction
  const bool weighted = relax && relax->weights;
  if (relax) ps.handle.set_prediction({}, 0);
  const std::vector<char> whole = ps.handle.whole_copies(datas);
  // uncached pages with little active data (-page_direct) are pushed
  // straight from pinned host memory: the kernels read their active Node
  // records and out-lists over PCIe, with no host compaction, no slot and
  // no copy.  The page cache does not see them (their data counts as zero
  // in queue_reset); they go out in one grouped launch ahead of the cached
  // pages' groups.
  std::vector<int> direct;
  std::vector<uint32_t> sched_datas;
  const uint32_t* qdatas = datas;
  if (!relax && cfg_.page_direct > 0 && ps.group && !ps.d_hpage.empty() && ps.d_hpage[0]) {
    uint64_t dn = 0;
    for (size_t p = 0; p < np; ++p)
      if (!ps.handle.vp[p].cached && datas[p] > 0 && !whole[p] &&
          datas[p] <= (uint64_t)cfg_.page_direct && dn + nodes[p] <= ps.group_nodes_cap) {
        direct.push_back((int)p);
        dn += nodes[p];
      }
    if (!direct.empty()) {
      sched_datas.assign(datas, datas + np);
      for (int p : direct) sched_datas[p] = 0;
      qdatas = sched_datas.data();
    }
  }
  std::vector<int> cold;
  for (size_t p = 0; p < np; ++p)
    if (!ps.handle.vp[p].cached && qdatas[p] > 0 && !whole[p]) cold.push_back((int)p);
  if (!cold.empty()) {
    if (cold.size() * 4 > np) {
      H_ERR(scaph_copy_d2h_async(ps.h_front.get(), front_[fa].get(), words * 4, st));
    } else {
      // few cold pages: only the frontier words of their vertex ranges (a
      // small level otherwise pays for the whole n/8-byte bitmap)
      for (int p : cold) {
        const uint64_t w0 = pages_out_[p].left / 32;
        const uint64_t w1 = ((uint64_t)pages_out_[p].left + pages_out_[p].nodenum + 31) / 32;
        H_ERR(scaph_copy_d2h_async(ps.h_front.get<uint32_t>() + w0,
                                   front_[fa].get<uint32_t>() + w0, (w1 - w0) * 4, st));
      }
    }
    main_->sync();
  }
  ps.resident_at_reset.assign(np, 0);
  for (size_t p = 0; p < np; ++p) ps.resident_at_reset[p] = ps.handle.vp[p].cached ? 1 : 0;
  const bool grouping = cfg_.page_groups > 0 && !relax;
  ps.handle.queue_reset(qdatas, grouping);
  rs->pages_full += ps.handle.num_full;
  // grouped launches over the cached active pages: when copies wait for
  // slots, about page_groups of them (at least kMinGroupEdges out-edges
  // each, the first a quarter of that), so slots free as groups finish;
  // otherwise one launch sequence for the whole level
  uint64_t group_target = 0;
  bool first_group = true;
  std::vector<int> group_pages;
  const bool copies_pending = ps.handle.num_full > 0 || !cold.empty();
  if (grouping) {
    uint64_t e = 0;
    for (int p : ps.handle.group_queue()) e += datas[p] - 3ull * nodes[p];
    group_target = copies_pending
                       ? std::max<uint64_t>(e / (uint64_t)cfg_.page_groups, kMinGroupEdges)
                       : std::max<uint64_t>(e, 1);
  }

  if (!direct.empty()) {
    auto* tab = ps.h_runs.get<scaph_page_run>();
    uint64_t g_edges = 0, g_nodes = 0;
    uint32_t segs = 0;
    for (size_t k = 0; k < direct.size(); ++k) {
      const int p = direct[k];
      scaph_page_run& r = tab[k];
      r.buf = ps.d_hpage[p];
      r.first_vertex = pages_out_[p].left;
      r.nodenum = pages_out_[p].nodenum;
      r.seg0 = segs;
      r.reserved = 0;
      segs += (r.nodenum + SCAPH_RUN_SEGMENT - 1) / SCAPH_RUN_SEGMENT;
      g_edges += datas[p] - 3ull * nodes[p];
      g_nodes += nodes[p];
    }
    // on the group stream, whose first group waits for group_done: the run
    // table in h_runs is rewritten only after this launch has finished
    void* gs = ps.group->get();
    H_ERR(scaph_copy_h2d_async(ps.d_runs.get(), tab, direct.size() * sizeof(scaph_page_run), gs));
    H_ERR(scaph_bfs_push_runs(ps.d_runs.get<scaph_page_run>(), (uint32_t)direct.size(), segs,
                              g_nodes, g_edges, n_, deg, level, visited,
                              front_[fa].get<uint32_t>(), front_[fa ^ 1].get<uint32_t>(), cnt,
                              cur, ps.group_ws.get(), ps.group_ws_bytes, &push_tune_, gs));
    ps.group_done->record(gs);
    rs->page_execs += direct.size();
    rs->pages_direct += direct.size();
  }

  // compaction workers (parallel_compress_stat_async, src/compress.h:42-56)
  std::atomic<size_t> next{0}, done{0};
  std::atomic<bool> failed{false};
  Workers workers;
  const int nw = (int)std::min<size_t>(cold.size(), 8);
  for (int w = 0; w < nw; ++w) {
    workers.t.emplace_back([&] {
      for (size_t k = next++; k < cold.size(); k = next++) {
        const int p = cold[k];
        try {
          const uint64_t used = compact_page(ps.h_page[p].get<uint32_t>(), pages_out_[p],
                                             ps.h_front.get<uint32_t>(), nodes[p],
                                             ps.h_tail[p].get<uint32_t>());
          if (weighted)
            compact_weights(ps.h_page[p].get<uint32_t>(), pages_out_[p],
                            ps.h_wpage[p].get<uint32_t>(), ps.h_front.get<uint32_t>(), nodes[p],
                            ps.h_wtail[p].get<uint32_t>());
          VPstat& v = ps.handle.vp[p];
          v.nodenum = nodes[p];
          v.datanum = used;
          v.chunknum = (int)((used + ps.chunk - 1) / ps.chunk);
          v.shared = true;
          ps.handle.merge_write(p);
        } catch (...) {
          failed = true;
        }
        ++done;
      }
    });
  }
  rs->pages_compacted += cold.size();

  // scheduling loop (gpu_push_batch, src/app.cuh:414-459), event driven
  std::vector<int> last(ns, -1);
  int copy_last = -1;
  bool copy_ahead = false;  // copy_last is a look-ahead copy for the next level
  float copy_ms = 0;
  bool copy_timed = false;
  while (!ps.handle.finished()) {
    if (failed) break;
    bool progress = false;
    for (int s = 0; s < ns; ++s) {
      if (!ps.exec_done[s]->done()) continue;
      if (last[s] != -1) {
        ps.handle.release_exec_finished(last[s]);
        last[s] = -1;
        progress = true;
      }
      int e;
      if (ps.handle.get_execute(&e)) {
        const VPstat& v = ps.handle.vp[e];
        const uint32_t* buf = ps.pool.get<uint32_t>() + (uint64_t)v.dpid * P_ +
                              (uint64_t)v.dpoff * ps.chunk;
        const uint64_t edges_bound = v.shared ? v.datanum - 3ull * v.nodenum
                                              : (uint64_t)pages_out_[e].edgenum;
        if (relax) {
          // a compacted page holds only active Nodes, each naming its vertex
          H_ERR(scaph_relax_page(buf, v.shared ? v.nodenum : pages_out_[e].nodenum,
                                 v.shared ? SCAPH_PAGE_COMPACTED : pages_out_[e].left,
                                 front_[fa].get<uint32_t>(),
                                 n_, relax->weights, relax->weights ? ps.pool.get<uint32_t>()
                                                                    : nullptr,
                                 relax->value, front_[fa ^ 1].get<uint32_t>(), cnt,
                                 ps.relax_ws[s].get(), ps.relax_ws[s].bytes(),
                                 ps.exec[s]->get()));
        } else {
          // page counts of this frontier: the exact active Nodes and out-edges
          scaph_page_hint hint;
          hint.first_vertex = v.shared ? SCAPH_PAGE_COMPACTED : pages_out_[e].left;
          hint.active_nodes = nodes[e];
          hint.active_edges = datas[e] - 3ull * nodes[e];
          H_ERR(scaph_bfs_push_page(buf, v.nodenum, std::max<uint64_t>(edges_bound, 1), n_, deg,
                                    level, visited, front_[fa].get<uint32_t>(),
                                    front_[fa ^ 1].get<uint32_t>(), cnt, cur, ps.ws[s].get(),
                                    ps.ws_bytes, cfg_.page_hint ? &hint : nullptr, &push_tune_,
                                    ps.exec[s]->get()));
        }
        ps.exec_done[s]->record(ps.exec[s]->get());
        last[s] = e;
        ++rs->page_execs;
        if (ps.resident_at_reset[e]) ++rs->page_execs_cached;
        progress = true;
      }
    }
    if (grouping && ps.group_done->done()) {
      if (!group_pages.empty()) {
        for (int p : group_pages) ps.handle.release_exec_finished(p);
        group_pages.clear();
        progress = true;
      }
      auto& gq = ps.handle.group_queue();
      if (!gq.empty()) {
        uint64_t target = group_target;
        if (first_group && copies_pending) target = std::max<uint64_t>(target / 4, 1);
        first_group = false;
        auto* tab = ps.h_runs.get<scaph_page_run>();
        uint64_t g_edges = 0, g_nodes = 0;
        uint32_t segs = 0;
        while (!gq.empty()) {
          const int p = gq.front();
          if (!group_pages.empty() &&
              (g_edges >= target || g_nodes + nodes[p] > ps.group_nodes_cap))
            break;
          ps.handle.take_group();
          const VPstat& v = ps.handle.vp[p];
          scaph_page_run& r = tab[group_pages.size()];
          r.buf = reinterpret_cast<unsigned long long>(ps.pool.get<uint32_t>() +
                                                       (uint64_t)v.dpid * P_);
          r.first_vertex = pages_out_[p].left;
          r.nodenum = pages_out_[p].nodenum;
          r.seg0 = segs;
          r.reserved = 0;
          segs += (r.nodenum + SCAPH_RUN_SEGMENT - 1) / SCAPH_RUN_SEGMENT;
          g_edges += datas[p] - 3ull * nodes[p];
          g_nodes += nodes[p];
          group_pages.push_back(p);
          ++rs->page_execs;
          if (ps.resident_at_reset[p]) ++rs->page_execs_cached;
        }
        void* gs = ps.group->get();
        H_ERR(scaph_copy_h2d_async(ps.d_runs.get(), tab,
                                   group_pages.size() * sizeof(scaph_page_run), gs));
        H_ERR(scaph_bfs_push_runs(ps.d_runs.get<scaph_page_run>(), (uint32_t)group_pages.size(),
                                  segs, g_nodes, g_edges, n_, deg, level, visited,
                                  front_[fa].get<uint32_t>(), front_[fa ^ 1].get<uint32_t>(),
                                  cnt, cur, ps.group_ws.get(), ps.group_ws_bytes, &push_tune_,
                                  gs));
        ps.group_done->record(gs);
        progress = true;
      }
    }
    if (ps.copy_end->done()) {
      if (copy_last != -1) {
        if (copy_timed) copy_ms += Event::elapsed_ms(*ps.copy_beg, *ps.copy_end);
        if (copy_ahead) ps.handle.release_lookahead_finished(copy_last);
        else ps.handle.release_copy_finished(copy_last);
        copy_last = -1;
        progress = true;
      }
      int s;
      if (!relax && ps.handle.predicting() && done.load() == cold.size() &&
          ps.handle.merge_pending() == 0 && ps.handle.launches_queued() &&
          ps.handle.get_lookahead(datas, &s)) {
        // the link is idle for this level while page launches still wait:
        // copy a page the next level is expected to need into a free slot
        // no later level needs
        const uint64_t w = page_used_words(pages_out_[s], false);
        ps.copy_beg->record(ps.copy->get());
        H_ERR(scaph_copy_h2d_async(ps.pool.get<uint32_t>() + (uint64_t)ps.handle.vp[s].dpid * P_,
                                   ps.h_page[s].get(), w * 4, ps.copy->get()));
        ps.copy_end->record(ps.copy->get());
        copy_timed = true;
        rs->bytes_h2d += w * 4;
        ++rs->pages_prefetched;
        copy_last = s;
        copy_ahead = true;
        progress = true;
      } else if (ps.handle.get_empty(&s)) {
        copy_ahead = false;
        // PageCopy (src/app.cuh:371-393): a compacted run into its chunks,
        // or the whole page into its slot
        const VPstat& v = ps.handle.vp[s];
        uint32_t* dst = ps.pool.get<uint32_t>() + (uint64_t)v.dpid * P_ +
                        (uint64_t)v.dpoff * ps.chunk;
        const void* srcp = v.shared ? ps.h_tail[s].get() : ps.h_page[s].get();
        const uint64_t w = v.shared ? v.datanum : page_used_words(pages_out_[s], false);
        ps.copy_beg->record(ps.copy->get());
        H_ERR(scaph_copy_h2d_async(dst, srcp, w * 4, ps.copy->get()));
        if (weighted && v.shared) {
          // the compacted weights at the compacted page's offsets
          H_ERR(scaph_copy_h2d_async(ps.wpool.get<uint32_t>() + (dst - ps.pool.get<uint32_t>()),
                                     ps.h_wtail[s].get(), w * 4, ps.copy->get()));
          rs->bytes_h2d += w * 4;
        } else if (weighted && pages_out_[s].edgenum) {
          // the page's weights into the same slot of the weight pool
          const uint64_t off = (uint64_t)v.dpid * P_ + 3ull * pages_out_[s].nodenum;
          H_ERR(scaph_copy_h2d_async(ps.wpool.get<uint32_t>() + off, ps.h_wpage[s].get(),
                                     (uint64_t)pages_out_[s].edgenum * 4, ps.copy->get()));
          rs->bytes_h2d += (uint64_t)pages_out_[s].edgenum * 4;
        }
        ps.copy_end->record(ps.copy->get());
        copy_timed = true;
        rs->bytes_h2d += w * 4;
        copy_last = s;
        progress = true;
      }
    }
    if (!progress) std::this_thread::yield();
  }
  workers.join();
  if (failed) throw Error(SCAPH_STATUS_INVALID, "page compaction failed");
  for (int s = 0; s < ns; ++s) ps.exec[s]->sync();
  if (ps.group) ps.group->sync();
  ps.copy->sync();
  if (copy_last != -1) {
    copy_ms += Event::elapsed_ms(*ps.copy_beg, *ps.copy_end);
    if (copy_ahead) ps.handle.release_lookahead_finished(copy_last);
    else ps.handle.release_copy_finished(copy_last);
  }
  rs->h2d_ms += copy_ms;
}

// Page-activity prediction for level `cur` (whose frontier has n_f vertices
// and m_f out-edges, m_u edges still unexplored), handed to the page cache
// before its queue_reset:
//  * learned: the model's second output (the fraction of pages active at
//    the next level, scaph_mlp_decide on this frontier) picks that many
//    pages, those with the most vertices a later level can still reach;
//    pages with none are never needed again;
//  * exact: every page's next active level from set_page_levels (Belady's
//    order: the page needed furthest ahead is reused first) and its data at
//    the next level, so a page that level will copy whole is copied whole
//    now — what no predictor can improve on.
// predict_launch queues the learned prediction's kernels and copies on the
// main stream ahead of the level's one synchronisation (true when it did);
// predict_apply turns the results into the page cache's prediction.
bool Engine::predict_launch(uint32_t cur, uint64_t n_f, uint64_t m_f, uint64_t m_u) {
  if (cfg_.page_predict != "learned" || !device_model_ready()) return false;
  PagedState& ps = *paged_;
  const size_t np = pages_out_.size();
  ensure_model_buffers();
  DeviceModel& m = *model_;
  void* st = main_->get();
  H_ERR(scaph_page_unvisited(dpages_out_.get<scaph_page>(), (uint32_t)np, visited_.get<uint32_t>(),
                             ps.live.get<uint32_t>(), ps.d_unv.get<uint32_t>(), st));
  H_ERR(scaph_copy_d2h_async(ps.h_unv.get(), ps.d_unv.get(), np * 4, st));
  auto* hcc = m.hccnt.get<scaph_level_counters>();
  std::memset(hcc, 0, sizeof(*hcc));
  hcc->n_next = n_f;
  hcc->m_next = m_f;
  H_ERR(scaph_copy_h2d_async(m.ccnt.get(), hcc, sizeof(*hcc), st));
  // the page counts of this frontier are in page_nodes_ (already queued for
  // the host; the decision kernel reads and clears them)
  H_ERR(scaph_mlp_decide(m.params.get<uint16_t>(), m.norm.get<float>(),
                         page_nodes_.get<uint32_t>(), dpages_out_.get<scaph_page>(), (uint32_t)np,
                         m.ccnt.get<scaph_level_counters>(), m_u + m_f, cur, m.out.get<float>(),
                         st));
  H_ERR(scaph_copy_d2h_async(m.hout.get(), m.out.get(), 32 * 4, st));
  return true;
}

void Engine::predict_apply(uint32_t cur, bool learned) {
  PagedState& ps = *paged_;
  const size_t np = pages_out_.size();
  if (cfg_.page_predict == "exact" && !ps.level_mask.empty()) {
    std::vector<double> value(np, 0.0);
    std::vector<char> keep(np, 0);
    const bool more = cur + 1 < 64;
    for (size_t p = 0; p < np; ++p) {
      const uint64_t later = more ? ps.level_mask[p] & (~0ull << (cur + 1)) : 0;
      if (!later) continue;
      const int next = __builtin_ctzll(later);
      value[p] = 1000.0 - next;
      keep[p] = next == (int)cur + 1 && (uint64_t)next < ps.level_datas.size() &&
                ps.level_datas[next][p] > ps.thres;
    }
    ps.handle.set_prediction(std::move(value), 1000.0 - (cur + 1), std::move(keep));
    return;
  }
  if (!learned) {
    ps.handle.set_prediction({}, 0);
    return;
  }
  const uint32_t* unv = ps.h_unv.get<uint32_t>();
  const float frac = std::min(1.f, std::max(0.f, model_->hout.get<float>()[1]));
  const size_t want = (size_t)std::lround(frac * (double)np);
  std::vector<int> alive;
  uint32_t umax = 1;
  for (size_t p = 0; p < np; ++p)
    if (unv[p]) {
      alive.push_back((int)p);
      umax = std::max(umax, unv[p]);
    }
  std::stable_sort(alive.begin(), alive.end(), [&](int a, int b) { return unv[a] > unv[b]; });
  std::vector<double> value(np, 0.0);
  for (size_t i = 0; i < alive.size(); ++i) {
    const int p = alive[i];
    value[p] = (i < want ? 2.0 : 1.0) + (double)unv[p] / umax;
  }
  // no whole-page copies: the model predicts which pages the next level
  // touches, not how much of each
  ps.handle.set_prediction(std::move(value), 2.0);
}

void Engine::set_page_levels(const uint32_t* levels) {
  if (resident_ || !paged_) throw Error(SCAPH_STATUS_INVALID, "page levels need the paged mode");
  PagedState& ps = *paged_;
  const size_t np = pages_out_.size();
  ps.level_mask.assign(np, 0);
  ps.level_datas.clear();
  for (size_t p = 0; p < np; ++p) {
    const scaph_page& pg = pages_out_[p];
    for (uint64_t v = pg.left; v < (uint64_t)pg.left + pg.nodenum; ++v) {
      const uint32_t l = levels[v];
      if (l >= 64) continue;
      ps.level_mask[p] |= 1ull << l;
      if (ps.level_datas.size() <= l) ps.level_datas.resize(l + 1, std::vector<uint64_t>(np, 0));
      ps.level_datas[l][p] += h_deg_out_[v] + 3ull;
    }
  }
}

// 12-bin histogram of the per-page active-vertex share (fraction of pages
// per bin; bin 0 = no active vertex, bins 1..11 split (0, 1] evenly in
// log2 of the share from 2^-10 up).
void Engine::page_histogram(const uint32_t* nodes, LevelRecord* rec) {
  rec->page_hist.fill(0.f);
  const size_t np = pages_out_.size();
  if (np == 0) return;
  for (size_t p = 0; p < np; ++p) {
    const double share = (double)nodes[p] / std::max<uint32_t>(pages_out_[p].nodenum, 1);
    int bin = 0;
    if (share > 0) {
      const double l = std::log2(share);  // <= 0
      bin = 1 + (int)std::min(10.0, std::max(0.0, std::floor(l + 11.0)));
      if (bin > 11) bin = 11;
    }
    rec->page_hist[bin] += 1.f;
  }
  for (auto& x : rec->page_hist) x /= (float)np;
}

}  // namespace scaph
