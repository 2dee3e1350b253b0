// engine_levels.cpp — the resident level loop (App::run's traversal,
// src/app.cuh:220-265, level-synchronous with the push/pull switch), the
// learned schedule's device decision, and the training-trace explorer.
#include <algorithm>
#include <cstring>

#include <cmath>

#include "engine_internal.h"

namespace scaph {

namespace {

constexpr size_t kMaxTimedLevels = 1024;
constexpr size_t kCounterWords = sizeof(scaph_level_counters) / 4;
constexpr size_t kDecideFloats = 32;  // logit, threshold, pages active, pad, hist[12], feat[16]
// Word offset of the learned decision's output inside the counters buffer
// (beside the level counters, so one copy returns both): after the level's
// record (words 0..11) and the queue conversion's record (12..23), so that
// the whole hand-off is 64 words, one store per lane of the publishing wave.
constexpr size_t kDecideWord = 32;
static_assert(kDecideWord >= 2 * kCounterWords && kDecideWord + kDecideFloats <= 64,
              "hand-off layout");
// Word of the hand-off flag in the mapped hand-off buffer.
constexpr size_t kHandFlagWord = 256;
// Push levels whose frontier has more out-edges than this snapshot visited
// first, so that a pull next can take the frontier as visited & ~snapshot.
constexpr uint64_t kSnapEdges = 1u << 22;

size_t bitmap_words(uint32_t n) { return (((size_t)n + 63) / 64) * 2; }

}  // namespace

// Builds the frontier forms a step needs: the pull step reads a bitmap,
// the push step a vertex list.
void Engine::ensure_forms(Front& f, bool queue, bool bitmap, uint64_t n_f) {
  void* st = main_->get();
  auto* cnt = counters_.get<scaph_level_counters>();
  if (bitmap && !f.front_ok) {
    if (f.snap_ok)
      H_ERR(scaph_frontier_from_visited(visited_.get<uint32_t>(), vsnap_.get<uint32_t>(),
                                        front_[f.fa].get<uint32_t>(), n_, st));
    else
      H_ERR(scaph_frontier_queue_to_bitmap(queue_[f.qa].get<uint32_t>(), (uint32_t)n_f,
                                           front_[f.fa].get<uint32_t>(), n_, st));
    f.front_ok = true;
  }
  if (queue && !f.queue_ok) {
    H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt + 1), 0u, kCounterWords, st));
    H_ERR(scaph_frontier_bitmap_to_queue(front_[f.fa].get<uint32_t>(), n_,
                                         queue_[f.qa].get<uint32_t>(), cnt + 1, st));
    f.queue_ok = true;
  }
}

// Launches one level in direction d (counters must be cleared first).
void Engine::launch_step(Dir d, Front& f, uint32_t cur, uint64_t n_f, uint64_t m_f, bool sparse) {
  void* st = main_->get();
  auto* cnt = counters_.get<scaph_level_counters>();
  uint32_t* level = level_.get<uint32_t>();
  uint32_t* visited = visited_.get<uint32_t>();
  const uint32_t* deg = deg_out_.get<uint32_t>();
  if (d == Dir::kPush) {
    // with a pull step possible next, keep visited as it was before this
    // push: the next frontier's bitmap is then visited & ~snapshot
    // (small frontiers skip it: their list converts cheaply with atomics)
    const bool snap = vsnap_.get() != nullptr && m_f > kSnapEdges;
    if (snap) H_ERR(scaph_copy_d2d_async(vsnap_.get(), visited, vsnap_.bytes(), st));
    H_ERR(scaph_bfs_push(&g_out_, deg, level, visited, queue_[f.qa].get<uint32_t>(), (uint32_t)n_f,
                         m_f, queue_[f.qa ^ 1].get<uint32_t>(), cnt, cur, ws_.get(), ws_bytes_,
                         &push_tune_, st));
    f.qa ^= 1;
    f.queue_ok = true;
    f.front_ok = false;
    f.snap_ok = snap;
  } else {
    scaph_pull_heads heads{in_head_.get<uint32_t>(),
                           live_base_.get() ? static_visited_.get<uint32_t>() : nullptr,
                           live_base_.get<uint32_t>()};
    if (sparse) {
      H_ERR(scaph_bfs_pull_sparse(&g_in_, in_head_.get() ? &heads : nullptr,
                                  hub_list_.get<uint32_t>(), num_hubs_, deg, level, visited,
                                  front_[f.fa].get<uint32_t>(), front_[f.fa ^ 1].get<uint32_t>(),
                                  cnt, cur, m_, pws_.get(), pws_bytes_, &pull_tune_, st));
    } else {
      // a sparse frontier: most non-hub probes end at the summary's clear bits
      scaph_pull_tuning tune = pull_tune_;
      if (cfg_.pull_summary > 0 && cfg_.pull_summary_div >= 0) {
        uint32_t sh = (uint32_t)std::max(5, cfg_.pull_summary);
        while (sh < 31 && ((n_ + (1ull << sh) - 1) >> sh) > (1ull << 22)) ++sh;
        if (cfg_.pull_summary_div == 0 || n_f * (uint64_t)cfg_.pull_summary_div < (n_ >> sh))
          tune.summary_shift = sh;
      }
      H_ERR(scaph_bfs_pull(&g_in_, in_head_.get() ? &heads : nullptr, hub_list_.get<uint32_t>(),
                           num_hubs_, deg, level, visited,
                           front_[f.fa].get<uint32_t>(), front_[f.fa ^ 1].get<uint32_t>(), cnt,
                           cur, m_, pws_.get(), pws_bytes_, &tune, st));
    }
    f.fa ^= 1;
    f.front_ok = true;
    f.queue_ok = false;
    f.snap_ok = false;
  }
}

// Spins on the mapped flag the level's publish kernel writes last; checks
// the stream now and then so that a failed launch raises instead of hanging.
void Engine::wait_handoff(uint32_t seq) {
  const volatile uint32_t* flag = hand_.host<uint32_t>() + kHandFlagWord;
  for (uint64_t spins = 1;; ++spins) {
    if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return;
    if ((spins & 0x3FF) == 0) {
      const int st = scaph_stream_query(main_->get());
      if (st == 0) {
        // the stream drained: the flag is visible now or never will be
        if (__atomic_load_n(flag, __ATOMIC_ACQUIRE) == seq) return;
        throw Error(SCAPH_STATUS_INVALID, "level hand-off flag not written");
      }
      if (st != SCAPH_STATUS_NOT_READY) raise_status(st, __FILE__, __LINE__);
    }
    __builtin_ia32_pause();
  }
}

void Engine::ensure_model_buffers() {
  if (!model_) model_.reset(new DeviceModel());
  DeviceModel& m = *model_;
  if (!m.params.get()) {
    m.params.alloc(SCAPH_MLP_NPARAM * 2);
    m.norm.alloc(32 * 4);
    m.out.alloc(kDecideFloats * 4);
    m.ccnt.alloc(sizeof(scaph_level_counters));
    m.hout.alloc(kDecideFloats * 4);
    m.hccnt.alloc(sizeof(scaph_level_counters));
    // an all-zero model with unit scale: features are still computed
    std::vector<uint16_t> z(SCAPH_MLP_NPARAM, 0);
    std::vector<float> norm(32, 0.f);
    for (int i = 16; i < 32; ++i) norm[i] = 1.f;
    H_ERR(scaph_copy_h2d_async(m.params.get(), z.data(), z.size() * 2, main_->get()));
    H_ERR(scaph_copy_h2d_async(m.norm.get(), norm.data(), norm.size() * 4, main_->get()));
    main_->sync();
  }
  if (m.page_cnt.bytes() < std::max<size_t>(pages_out_.size(), 1) * 4) {
    // the decision kernels expect (and leave) page_cnt all zero
    m.page_cnt.alloc(std::max<size_t>(pages_out_.size(), 1) * 4);
    H_ERR(scaph_fill_u32(m.page_cnt.get<uint32_t>(), 0u, m.page_cnt.bytes() / 4, main_->get()));
    main_->sync();
  }
}

void Engine::set_device_model(const uint16_t* params, const float* mean, const float* scale) {
  ensure_model_buffers();
  std::vector<float> norm(32);
  for (int i = 0; i < 16; ++i) {
    norm[i] = mean[i];
    norm[16 + i] = scale[i] == 0.f ? 1.f : scale[i];
  }
  H_ERR(scaph_copy_h2d_async(model_->params.get(), params, SCAPH_MLP_NPARAM * 2, main_->get()));
  H_ERR(scaph_copy_h2d_async(model_->norm.get(), norm.data(), norm.size() * 4, main_->get()));
  main_->sync();
  model_->ready = true;
  // W1 is [64][16] row-major: unit j's weight for feature i at j * 16 + i
  bool scalar = true;
  for (int j = 0; j < 64 && scalar; ++j)
    for (int i = 4; i < 16; ++i)
      if ((params[j * 16 + i] & 0x7FFFu) != 0u) scalar = false;
  model_->scalar_only = scalar && cfg_.learned_fused;
}

bool Engine::device_model_ready() const { return model_ && model_->ready; }

RunStats Engine::bfs_resident(uint32_t src) {
  void* st = main_->get();
  RunStats rs;
  rs.src = src;
  while (ev_.size() < 2) ev_.emplace_back(new Event());
  const uint64_t deg_src = out_degree(src);
  auto* cnt = counters_.get<scaph_level_counters>();
  const bool learned = cfg_.schedule == "learned" && device_model_ready() && has_in_;
  if (learned) ensure_model_buffers();
  const float* hout = learned ? hand_.host<float>() + kDecideWord : nullptr;
  auto* hc = hand_.host<scaph_level_counters>();

  ev_[0]->record(st);
  H_ERR(scaph_bfs_init(level_.get<uint32_t>(), visited_.get<uint32_t>(),
                       has_in_ ? static_visited_.get<uint32_t>() : nullptr, n_, src,
                       queue_[0].get<uint32_t>(), cnt, st));
  Dir dir = Dir::kPush;
  uint64_t n_f = 1, m_f = deg_src, m_u = m_ - deg_src, n_f_prev = 0;
  Front f;
  uint32_t cur = 0;
  rs.n_reached = 1;
  rs.m_reached = deg_src;
  std::vector<size_t> lev_ev;
  bool have_decision = false;
  float logit = 0.f;
  while (n_f > 0) {
    SchedState ss{dir, cur, n_, m_, n_f, m_f, m_u, n_f_prev,
                  rs.levels.empty() ? nullptr : &rs.levels.back()};
    Dir d;
    if (learned && have_decision && std::fabs(logit) >= cfg_.learned_margin)
      d = logit > 0.f ? Dir::kPull : Dir::kPush;
    else if (learned)
      d = beamer(ss);  // the fixed rule: the first level, and calls the model is unsure of
    else
      d = sched_ ? sched_(ss) : beamer(ss);
    if (d == Dir::kPull && !has_in_) d = Dir::kPush;
    ensure_forms(f, d == Dir::kPush, d == Dir::kPull, n_f);
    // the counters start cleared: by scaph_bfs_init for level 0, by the
    // previous level's hand-off after that
    const bool timed = rs.levels.size() < kMaxTimedLevels;
    const size_t eb = 2 + 3 * rs.levels.size();
    if (timed) {
      while (ev_.size() < eb + 3) ev_.emplace_back(new Event());
      ev_[eb]->record(st);
    }
    // after a pull level the unvisited vertices left are known exactly: it
    // examined every one and found n_next of them
    const LevelRecord* prev = rs.levels.empty() ? nullptr : &rs.levels.back();
    const bool sparse = d == Dir::kPull && prev && prev->dir == (int)Dir::kPull &&
                        cfg_.pull_sparse > 0 &&
                        (prev->examined - prev->n_next) * (uint64_t)cfg_.pull_sparse < n_;
    launch_step(d, f, cur, n_f, m_f, sparse);
    if (timed) {
      ev_[eb + 1]->record(st);
      lev_ev.push_back(eb);
    }
    const bool fused = learned && model_->scalar_only;
    if (learned && !fused) {
      // features of the new frontier and the model's choice, on the device;
      // the decision lands beside the counters so one copy brings both back
      DeviceModel& m = *model_;
      const bool bm = d == Dir::kPull;
      H_ERR(scaph_mlp_decide_frontier(
          m.params.get<uint16_t>(), m.norm.get<float>(),
          bm ? front_[f.fa].get<uint32_t>() : nullptr, bm ? nullptr : queue_[f.qa].get<uint32_t>(),
          g_out_.pages, (uint32_t)pages_out_.size(), n_, cnt, m_u, cur + 1,
          m.page_cnt.get<uint32_t>(), counters_.get<float>() + kDecideWord, st));
    }
    // the counters (and the decision) land in mapped host memory, the
    // counters are cleared for the next level, and a flag says so
    const uint32_t seq = ++hand_seq_;
    if (fused) {
      // scalar-feature model: the decision is formed inside the hand-off
      H_ERR(scaph_mlp_decide_publish(model_->params.get<uint16_t>(), model_->norm.get<float>(),
                                     counters_.get<uint32_t>(), (uint32_t)kDecideWord, m_u,
                                     cur + 1, (uint32_t)(kDecideWord + kDecideFloats),
                                     (uint32_t)kCounterWords, hand_.dev<uint32_t>(),
                                     hand_.dev<uint32_t>() + kHandFlagWord, seq, st));
    } else {
      H_ERR(scaph_publish_words(counters_.get<uint32_t>(),
                                (uint32_t)(learned ? kDecideWord + kDecideFloats : kCounterWords),
                                (uint32_t)kCounterWords, hand_.dev<uint32_t>(),
                                hand_.dev<uint32_t>() + kHandFlagWord, seq, st));
    }
    if (timed) ev_[eb + 2]->record(st);
    wait_handoff(seq);
    LevelRecord rec;
    rec.level = cur;
    rec.dir = (int)d;
    rec.logit = learned && have_decision ? logit : 0.f;
    rec.n_f = n_f;
    rec.m_f = m_f;
    rec.m_u = m_u;
    rec.n_next = hc->n_next;
    rec.m_next = hc->m_next;
    rec.scanned = hc->scanned;
    rec.examined = hc->examined;
    rec.sparse = sparse;
    if (hc->reserved) throw Error(SCAPH_STATUS_INVALID, "pull work list overflow");
    if (learned) {
      logit = hout[0];
      have_decision = true;
      for (int i = 0; i < 12; ++i) rec.page_hist[i] = hout[4 + i];
    }
    rs.levels.push_back(rec);
    n_f_prev = n_f;
    n_f = hc->n_next;
    m_f = hc->m_next;
    m_u = m_u > m_f ? m_u - m_f : 0;
    rs.n_reached += n_f;
    rs.m_reached += m_f;
    dir = d;
    ++cur;
  }
  ev_[1]->record(st);
  ev_[1]->sync();
  rs.ms = Event::elapsed_ms(*ev_[0], *ev_[1]);
  for (size_t i = 0; i < lev_ev.size(); ++i) {
    const size_t e = lev_ev[i];
    const Event& prev_end = i == 0 ? *ev_[0] : *ev_[lev_ev[i - 1] + 2];
    rs.levels[i].ms = Event::elapsed_ms(prev_end, *ev_[e + 2]);
    rs.levels[i].kernel_ms = Event::elapsed_ms(*ev_[e], *ev_[e + 1]);
  }
  rs.depth = cur;
  return rs;
}

std::vector<Engine::ExploreRecord> Engine::explore(uint32_t src, int reps) {
  if (!resident_ || !has_in_) throw Error(SCAPH_STATUS_INVALID, "explore needs resident pull pages");
  if (src >= n_) throw Error(SCAPH_STATUS_INVALID, "source vertex out of range");
  reps = std::max(1, reps);
  ensure_model_buffers();
  DeviceModel& m = *model_;
  void* st = main_->get();
  const size_t words = bitmap_words(n_);
  // snapshot buffers follow the loaded graph's size
  if (snap_level_.bytes() < std::max<size_t>(n_, 1) * 4) {
    snap_level_.alloc(std::max<size_t>(n_, 1) * 4);
    snap_visited_.alloc(words * 4);
    snap_queue_.alloc(std::max<size_t>(n_, 1) * 4);
    snap_front_.alloc(words * 4);
  }
  auto* cnt = counters_.get<scaph_level_counters>();
  auto* hc = hcount_.get<scaph_level_counters>();
  Event e0, e1;
  const uint64_t deg_src = out_degree(src);
  H_ERR(scaph_bfs_init(level_.get<uint32_t>(), visited_.get<uint32_t>(),
                       static_visited_.get<uint32_t>(), n_, src, queue_[0].get<uint32_t>(), cnt,
                       st));
  uint64_t n_f = 1, m_f = deg_src, m_u = m_ - deg_src;
  Front f;
  uint32_t cur = 0;
  std::vector<ExploreRecord> recs;
  const float* hout = m.hout.get<float>();
  auto snapshot = [&](bool save) {
    auto cp = [&](void* dst, const void* srcp, size_t bytes) {
      H_ERR(scaph_copy_d2d_async(dst, srcp, bytes, st));
    };
    if (save) {
      cp(snap_level_.get(), level_.get(), (size_t)n_ * 4);
      cp(snap_visited_.get(), visited_.get(), words * 4);
      cp(snap_queue_.get(), queue_[f.qa].get(), n_f * 4);
      cp(snap_front_.get(), front_[f.fa].get(), words * 4);
    } else {
      cp(level_.get(), snap_level_.get(), (size_t)n_ * 4);
      cp(visited_.get(), snap_visited_.get(), words * 4);
      cp(queue_[f.qa].get(), snap_queue_.get(), n_f * 4);
      cp(front_[f.fa].get(), snap_front_.get(), words * 4);
    }
  };
  while (n_f > 0) {
    ensure_forms(f, true, true, n_f);
    // features of the frontier entering this level (same kernel as run time)
    auto* hcc = m.hccnt.get<scaph_level_counters>();
    std::memset(hcc, 0, sizeof(*hcc));
    hcc->n_next = n_f;
    hcc->m_next = m_f;
    H_ERR(scaph_copy_h2d_async(m.ccnt.get(), hcc, sizeof(*hcc), st));
    H_ERR(scaph_frontier_page_counts(front_[f.fa].get<uint32_t>(), nullptr, nullptr, g_out_.pages,
                                     (uint32_t)pages_out_.size(), n_, m.page_cnt.get<uint32_t>(),
                                     st));
    H_ERR(scaph_mlp_decide(m.params.get<uint16_t>(), m.norm.get<float>(), m.page_cnt.get<uint32_t>(),
                           g_out_.pages, (uint32_t)pages_out_.size(),
                           m.ccnt.get<scaph_level_counters>(), m_u + m_f, cur, m.out.get<float>(),
                           st));
    H_ERR(scaph_copy_d2h_async(m.hout.get(), m.out.get(), kDecideFloats * 4, st));
    snapshot(true);
    main_->sync();
    ExploreRecord r;
    for (int i = 0; i < 16; ++i) r.x[i] = hout[16 + i];
    if (!recs.empty()) recs.back().pages_active_next = hout[2] / (float)pages_out_.size();
    r.level = cur;
    r.n_f = n_f;
    r.m_f = m_f;
    float t[2] = {1e30f, 1e30f};
    uint64_t nn[2] = {0, 0}, mn[2] = {0, 0};
    for (int dd = 0; dd < 2; ++dd) {
      for (int k = 0; k < reps; ++k) {
        snapshot(false);
        Front g = f;
        H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt), 0u, kCounterWords, st));
        e0.record(st);
        launch_step(dd == 0 ? Dir::kPush : Dir::kPull, g, cur, n_f, m_f);
        e1.record(st);
        H_ERR(scaph_copy_d2h_async(hc, cnt, sizeof(*hc), st));
        main_->sync();
        t[dd] = std::min(t[dd], Event::elapsed_ms(e0, e1));
        nn[dd] = hc->n_next;
        mn[dd] = hc->m_next;
      }
    }
    if (nn[0] != nn[1] || mn[0] != mn[1])
      throw Error(SCAPH_STATUS_INVALID, "push and pull disagree on the next frontier");
    r.t_push_ms = t[0];
    r.t_pull_ms = t[1];
    r.n_next = nn[0];
    r.pages_active_next = 0.f;
    recs.push_back(r);
    // continue with the faster direction
    const Dir win = t[1] < t[0] ? Dir::kPull : Dir::kPush;
    snapshot(false);
    H_ERR(scaph_fill_u32(reinterpret_cast<uint32_t*>(cnt), 0u, kCounterWords, st));
    launch_step(win, f, cur, n_f, m_f);
    main_->sync();
    n_f = nn[0];
    m_f = mn[0];
    m_u = m_u > m_f ? m_u - m_f : 0;
    ++cur;
  }
  return recs;
}

}  // namespace scaph
