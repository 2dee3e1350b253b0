// engine_relax.cpp — the reference's other App plugins on the same engine:
// WCC (app/wcc.cu) and SSSP (app/sssp.cu).  Resident, each round turns the
// active set (a bitmap) into a list and runs scaph_relax_round over it
// (with -sssp_delta, only its vertices below a moving bound); paged, the
// round's active pages go through the paged BFS scheduler and
// scaph_relax_page.  The vertices whose value a round lowered form the next
// active set, and the rounds stop when none is active, as the reference's
// run() stops (src/app.cuh:220-265).
#include <algorithm>
#include <chrono>

#include "engine_internal.h"

namespace scaph {

namespace {

size_t bitmap_words(uint32_t n) { return (((size_t)n + 63) / 64) * 2; }

}  // namespace

RunStats Engine::wcc() { return relax(SCAPH_RELAX_WCC, 0); }

RunStats Engine::sssp(uint32_t src) {
  if (!has_weights()) throw Error(SCAPH_STATUS_INVALID, "SSSP needs edge weights (_csr.wgh)");
  return relax(SCAPH_RELAX_SSSP, src);
}

Engine::DeviceState Engine::device_state() {
  if (n_ == 0) throw Error(SCAPH_STATUS_INVALID, "no graph loaded");
  DeviceState d;
  d.value = level_.get<uint32_t>();
  d.active = front_[0].get<uint32_t>();
  d.active_words = bitmap_words(n_);
  d.degree = deg_out_.get<uint32_t>();
  d.stream = main_->get();
  d.copy_stream = paged_ && paged_->copy ? paged_->copy->get() : nullptr;
  d.n = n_;
  return d;
}

RunStats Engine::relax_prepared(int mode) {
  if (mode != SCAPH_RELAX_WCC && mode != SCAPH_RELAX_SSSP)
    throw Error(SCAPH_STATUS_INVALID, "relax mode must be SCAPH_RELAX_WCC or SCAPH_RELAX_SSSP");
  if (mode == SCAPH_RELAX_SSSP && !has_weights())
    throw Error(SCAPH_STATUS_INVALID, "SSSP-mode rounds need edge weights (_csr.wgh)");
  return relax(mode, 0, true);
}

// Active vertices of front_[fa] (the list it converts to is not used).
uint64_t Engine::count_active(int fa) {
  void* st = main_->get();
  auto* cnt = counters_.get<scaph_level_counters>();
  auto* hc = hcount_.get<scaph_level_counters>();
  H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt), 0u, sizeof(scaph_level_counters) / 4, st));
  H_ERR(scaph_frontier_bitmap_to_queue(front_[fa].get<uint32_t>(), n_, queue_[1].get<uint32_t>(),
                                       cnt, st));
  H_ERR(scaph_copy_d2h_async(hc, cnt, sizeof(scaph_level_counters), st));
  main_->sync();
  return hc->q_len;
}

RunStats Engine::relax(int mode, uint32_t src, bool prepared) {
  if (n_ == 0) throw Error(SCAPH_STATUS_INVALID, "no graph loaded");
  if (mode == SCAPH_RELAX_SSSP && src >= n_)
    throw Error(SCAPH_STATUS_INVALID, "source vertex out of range");
  if (!resident_) return relax_paged(mode, src, prepared);
  const size_t ws_bytes = scaph_relax_workspace_bytes(m_);
  if (relax_ws_.bytes() < ws_bytes) relax_ws_.alloc(ws_bytes);
  void* st = main_->get();
  auto* cnt = counters_.get<scaph_level_counters>();
  auto* hc = hcount_.get<scaph_level_counters>();
  uint32_t* value = level_.get<uint32_t>();
  const size_t words = bitmap_words(n_);
  RunStats rs;
  rs.src = src;
  const auto t0 = std::chrono::steady_clock::now();
  int fa = 0;
  uint64_t n_act;
  if (prepared) {
    // the app's gpu_init wrote the values and the active set
    n_act = count_active(0);
  } else if (mode == SCAPH_RELAX_WCC) {
    H_ERR(scaph_relax_init(value, n_, mode, src, st));
    // every vertex starts active (app/wcc.cu:18-19); bits past n are ignored
    H_ERR(scaph_fill_u32(front_[0].get<uint32_t>(), 0xFFFFFFFFu, words, st));
    n_act = n_;
  } else {
    H_ERR(scaph_relax_init(value, n_, mode, src, st));
    *hcount_.get<uint32_t>() = src;
    H_ERR(scaph_copy_h2d_async(queue_[1].get(), hcount_.get(), 4, st));
    H_ERR(scaph_fill_u32(front_[0].get<uint32_t>(), 0u, words, st));
    H_ERR(scaph_frontier_queue_to_bitmap(queue_[1].get<uint32_t>(), 1, front_[0].get<uint32_t>(),
                                         n_, st));
    n_act = 1;
  }
  const uint32_t* weights = mode == SCAPH_RELAX_SSSP ? wpool_.get<uint32_t>() : nullptr;
  // near/far rounds (SSSP, -sssp_delta): only active vertices below `bound`
  // relax; the rest stay active, and the bound moves on when none is below
  // it, which saves re-relaxing vertices whose distance is still falling
  const uint64_t delta = mode == SCAPH_RELAX_SSSP ? (uint64_t)cfg_.sssp_delta : 0;
  uint64_t bound = delta;
  if (delta && !split_.get()) split_.alloc(64);
  auto* hsplit = reinterpret_cast<unsigned long long*>(hcount_.get<char>() + 256);
  uint32_t round = 0;
  while (n_act > 0) {
    H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt), 0u, sizeof(scaph_level_counters) / 4,
                         st));
    H_ERR(scaph_fill_u32(front_[fa ^ 1].get<uint32_t>(), 0u, words, st));
    H_ERR(scaph_frontier_bitmap_to_queue(front_[fa].get<uint32_t>(), n_, queue_[0].get<uint32_t>(),
                                         cnt, st));
    const uint32_t* q = queue_[0].get<uint32_t>();
    uint64_t n_q = n_act, n_far = 0;
    if (delta) {
      const uint32_t b = (uint32_t)std::min<uint64_t>(bound, 0xFFFFFFFFull);
      H_ERR(scaph_relax_split(q, (uint32_t)n_act, value, b, n_, queue_[1].get<uint32_t>(),
                              front_[fa ^ 1].get<uint32_t>(),
                              split_.get<unsigned long long>(), st));
      H_ERR(scaph_copy_d2h_async(hsplit, split_.get(), 24, st));
      main_->sync();
      n_q = hsplit[0];
      n_far = hsplit[1];
      if (n_q + n_far != n_act) throw Error(SCAPH_STATUS_INVALID, "active set size mismatch");
      if (n_q == 0) {
        // every active vertex is far: the bound moves past the least of them
        bound = std::max(bound + delta, (uint64_t)hsplit[2] + delta);
        fa ^= 1;
        continue;
      }
      q = queue_[1].get<uint32_t>();
    }
    H_ERR(scaph_relax_round(&g_out_, q, (uint32_t)n_q, weights, pool_out_.get<uint32_t>(), value,
                            front_[fa ^ 1].get<uint32_t>(), cnt, relax_ws_.get(),
                            relax_ws_.bytes(), st));
    H_ERR(scaph_copy_d2h_async(hc, cnt, sizeof(scaph_level_counters), st));
    main_->sync();
    if (hc->reserved) throw Error(SCAPH_STATUS_INVALID, "relax chunk list overflow");
    if (hc->q_len != n_act) throw Error(SCAPH_STATUS_INVALID, "active set size mismatch");
    LevelRecord rec;
    rec.level = round;
    rec.n_f = n_q;
    rec.n_next = hc->n_next;
    rec.scanned = hc->scanned;
    rs.levels.push_back(rec);
    rs.m_reached += hc->scanned;  // edges relaxed over all rounds
    // far vertices kept their bits; newly lowered ones were counted once
    n_act = hc->n_next + n_far;
    fa ^= 1;
    ++round;
  }
  rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rs.depth = round;
  rs.n_reached = n_;
  return rs;
}

// Paged rounds: the active pages of each round go through the same page
// scheduler as a paged BFS level (cached pages first, the others copied
// whole, or compacted to their active Nodes when at most -threshold of a
// page is active, with their weights for SSSP), each page relaxed by
// scaph_relax_page on an exec stream.
RunStats Engine::relax_paged(int mode, uint32_t src, bool prepared) {
  PagedState& ps = *paged_;
  void* st = main_->get();
  const size_t np = pages_out_.size();
  const size_t words = bitmap_words(n_);
  auto* cnt = counters_.get<scaph_level_counters>();
  auto* hc = hcount_.get<scaph_level_counters>();
  uint32_t* value = level_.get<uint32_t>();
  const bool weighted = mode == SCAPH_RELAX_SSSP;
  // chunk workspaces: every page's long lists fit
  uint64_t max_edges = 1;
  for (const auto& pg : pages_out_) max_edges = std::max<uint64_t>(max_edges, pg.edgenum);
  const size_t ws_bytes = scaph_relax_workspace_bytes(max_edges);
  if (ps.relax_ws.size() != ps.exec.size() ||
      (!ps.relax_ws.empty() && ps.relax_ws[0].bytes() < ws_bytes)) {
    ps.relax_ws.clear();
    for (size_t s = 0; s < ps.exec.size(); ++s) ps.relax_ws.emplace_back(ws_bytes);
  }
  if (weighted) {
    // weight slots for the pages cached now (BFS levels may have replaced
    // slots since the last SSSP run); later copies bring their own
    if (!ps.wpool.get()) ps.wpool.alloc(ps.pool.bytes());
    for (size_t p = 0; p < np; ++p) {
      const VPstat& v = ps.handle.vp[p];
      if (!v.cached || v.shared || pages_out_[p].edgenum == 0) continue;
      const uint64_t off = (uint64_t)v.dpid * P_ + 3ull * pages_out_[p].nodenum;
      H_ERR(scaph_copy_h2d_async(ps.wpool.get<uint32_t>() + off, ps.h_wpage[p].get(),
                                 (uint64_t)pages_out_[p].edgenum * 4, ps.copy->get()));
    }
    ps.copy->sync();
  }
  RunStats rs;
  rs.src = src;
  const auto t0 = std::chrono::steady_clock::now();
  int fa = 0;
  uint64_t n_act;
  if (prepared) {
    n_act = count_active(0);
  } else if (mode == SCAPH_RELAX_WCC) {
    H_ERR(scaph_relax_init(value, n_, mode, src, st));
    H_ERR(scaph_fill_u32(front_[0].get<uint32_t>(), 0xFFFFFFFFu, words, st));
    n_act = n_;
  } else {
    H_ERR(scaph_relax_init(value, n_, mode, src, st));
    *hcount_.get<uint32_t>() = src;
    H_ERR(scaph_copy_h2d_async(queue_[1].get(), hcount_.get(), 4, st));
    H_ERR(scaph_fill_u32(front_[0].get<uint32_t>(), 0u, words, st));
    H_ERR(scaph_frontier_queue_to_bitmap(queue_[1].get<uint32_t>(), 1, front_[0].get<uint32_t>(),
                                         n_, st));
    n_act = 1;
  }
  RelaxRound rr;
  rr.value = value;
  rr.weights = weighted ? ps.wpool.get<uint32_t>() : nullptr;
  uint32_t round = 0;
  while (n_act > 0) {
    const auto tl = std::chrono::steady_clock::now();
    H_ERR(scaph_page_counts(dpages_out_.get<scaph_page>(), (uint32_t)np, front_[fa].get<uint32_t>(),
                            deg_out_.get<uint32_t>(), 0, page_nodes_.get<uint32_t>(),
                            page_datas_.get<uint32_t>(), totals_.get<unsigned long long>(), st));
    H_ERR(scaph_copy_d2h_async(ps.h_nodes.get(), page_nodes_.get(), np * 4, st));
    H_ERR(scaph_copy_d2h_async(ps.h_datas.get(), page_datas_.get(), np * 4, st));
    H_ERR(scaph_copy_d2h_async(ps.h_totals.get(), totals_.get(), 16, st));
    H_ERR(scaph_fill_u32(front_[fa ^ 1].get<uint32_t>(), 0u, words, st));
    H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt), 0u, sizeof(scaph_level_counters) / 4,
                         st));
    main_->sync();
    if (ps.h_totals.get<unsigned long long>()[1] == 0) break;
    schedule_pages(fa, round, &rr, &rs);
    H_ERR(scaph_copy_d2h_async(hc, cnt, sizeof(scaph_level_counters), st));
    main_->sync();
    if (hc->reserved) throw Error(SCAPH_STATUS_INVALID, "relax chunk list overflow");
    LevelRecord rec;
    rec.level = round;
    rec.n_f = n_act;
    rec.n_next = hc->n_next;
    rec.scanned = hc->scanned;
    rec.ms = (float)std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tl)
                 .count();
    rs.levels.push_back(rec);
    rs.m_reached += hc->scanned;
    n_act = hc->n_next;
    fa ^= 1;
    ++round;
  }
  rs.ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  rs.depth = round;
  rs.n_reached = n_;
  return rs;
}

}  // namespace scaph
