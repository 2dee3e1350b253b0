// engine.cpp — resident BFS engine: graph upload, the level loop with the
// push/pull direction switch, validation and the CPU check.
#include "engine.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <deque>
#include <queue>
#include <thread>

#include "engine_internal.h"
#include "flags.h"

namespace scaph {

double Total_throughput = 0;

namespace {

constexpr uint32_t kMaxHubs = 262144;  // frontier bits the pull expand can keep in LDS
constexpr uint64_t kAlignWords = 64;  // page buffers start on 256-byte boundaries
constexpr size_t kMaxTimedLevels = 1024;
constexpr size_t kCounterWords = sizeof(scaph_level_counters) / 4;

uint64_t align_words(uint64_t w) { return (w + kAlignWords - 1) / kAlignWords * kAlignWords; }

size_t bitmap_words(uint32_t n) { return (((size_t)n + 63) / 64) * 2; }

}  // namespace

// ------------------------------------------------------------------ flags

Config& global_flags() {
  static Config c;
  static bool registered = false;
  if (!registered) {
    registered = true;
    FlagRegistry& r = FlagRegistry::global();
    r.add("pagesize", FlagType::kInt32, &c.pagesize, "subgraph pagesize");
    r.add("max_memory", FlagType::kInt32, &c.max_memory, "gpu maximum memory size");
    r.add("multi", FlagType::kInt32, &c.multi, "max entry times");
    r.add("num_stream", FlagType::kInt32, &c.num_stream, "gpu stream number");
    r.add("num_blks", FlagType::kInt32, &c.num_blks, "kernel mapping blocks");
    r.add("num_thds", FlagType::kInt32, &c.num_thds, "kernel mapping threads");
    r.add("threshold", FlagType::kInt32, &c.threshold, "threshold of compress point");
    r.add("check", FlagType::kBool, &c.check, "check result");
    r.add("tricknum", FlagType::kInt32, &c.tricknum, "CUs withheld from the traversal kernels");
    r.add("src", FlagType::kInt64, &c.src, "traversal start point");
    r.add("graphfile", FlagType::kString, &c.graphfile, "graphfile path");
    r.add("gen", FlagType::kString, &c.gen, "generator spec used when graphfile is empty");
    r.add("alpha", FlagType::kDouble, &c.alpha, "push->pull when m_f > m_u/alpha");
    r.add("beta", FlagType::kDouble, &c.beta, "pull->push when n_f shrinks below n/beta");
    r.add("schedule", FlagType::kString, &c.schedule, "fixed|push|pull|learned");
    r.add("model", FlagType::kString, &c.model, "learned schedule weights");
    r.add("rounds", FlagType::kInt32, &c.rounds, "timed rounds");
    r.add("device", FlagType::kInt32, &c.device, "HIP device ordinal");
    r.add("levels_out", FlagType::kString, &c.levels_out, "write u32 levels to this file");
    r.add("compat63", FlagType::kBool, &c.compat63, "report levels capped at 63 like the reference");
    r.add("trace", FlagType::kString, &c.trace, "per-iteration JSONL trace file");
    r.add("build_pull", FlagType::kBool, &c.build_pull, "keep transposed pages for pull");
    r.add("pull_grid", FlagType::kInt32, &c.pull_grid, "cap on pull expand blocks (0 = default)");
    r.add("pull_waves", FlagType::kInt32, &c.pull_waves,
          "waves per pull block, 4 or 8 (0 = default); the hub bits are staged once per block");
    r.add("pull_hubs", FlagType::kInt32, &c.pull_hubs,
          "highest-out-degree vertices whose frontier bits the pull keeps in LDS (<= 262144)");
    r.add("pull_sparse", FlagType::kInt32, &c.pull_sparse,
          "pull over a list of the unvisited vertices when fewer than n / this are left (0 = never)");
    r.add("pull_summary", FlagType::kInt32, &c.pull_summary,
          "pull: test a frontier summary of one bit per 2^this vertices before each non-hub "
          "in-neighbour's frontier word (0 = never, else >= 5; raised until n / 2^this <= 2^22)");
    r.add("pull_summary_div", FlagType::kInt32, &c.pull_summary_div,
          "pull: use the summary on levels whose frontier has fewer than n / 2^pull_summary / "
          "this vertices (0 = on every pull level)");
    r.add("pull_phases", FlagType::kInt32, &c.pull_phases,
          "2: a pull window's vertices are probed by their head records first and the undecided "
          "ones by their lists after (1: both per vertex; 3: as 2, the head probes pipelined "
          "one batch ahead)");
    r.add("pull_simd", FlagType::kInt32, &c.pull_simd,
          "8: pull kernel registers capped so 8 waves share a SIMD (0 = compiler's choice)");
    r.add("learned_fused", FlagType::kBool, &c.learned_fused,
          "a model reading only scalar features decides inside the level hand-off");
    r.add("learned_margin", FlagType::kDouble, &c.learned_margin,
          "learned schedule: the model decides when |pull logit| >= this, else the fixed rule");
    r.add("sssp_delta", FlagType::kInt32, &c.sssp_delta,
          "SSSP: relax only vertices below a bound that grows by this much (0 = all)");
    r.add("pull_head_index", FlagType::kBool, &c.pull_head_index,
          "pull head records carry the in-list's offset in its page, so a list past its head is "
          "read without its Node record");
    r.add("pull_prefetch", FlagType::kBool, &c.pull_prefetch,
          "request the next batch's head records before probing the current one");
    r.add("pull_head_live", FlagType::kBool, &c.pull_head_live,
          "head records only for vertices with in-edges, numbered by live rank");
    r.add("pull_head_outdeg", FlagType::kBool, &c.pull_head_outdeg,
          "head records carry the out-degree instead of a third in-list entry");
    r.add("push_grid", FlagType::kInt32, &c.push_grid, "cap on push expand blocks (0 = default)");
    r.add("push_rows", FlagType::kBool, &c.push_rows, "push expand reads coalesced edge rows");
    r.add("push_small", FlagType::kInt32, &c.push_small,
          "out-edges below which a push level runs as one launch (0 = 8192, 1 = never)");
    r.add("push_coop", FlagType::kInt32, &c.push_coop,
          "frontiers of at most 256 vertices with at most this many out-edges run in one launch, "
          "their edges spread over the grid (0 = 2^20, 1 = never)");
    r.add("push_scan", FlagType::kInt32, &c.push_scan,
          "prep blocks scanned by one block before the scan takes three launches (0 = 1024)");
    r.add("push_xcd", FlagType::kInt32, &c.push_xcd,
          "out-edges / 1024 from which a push level expands XCD-sliced (default 65536; 0 = off)");
    r.add("page_hint", FlagType::kBool, &c.page_hint,
          "paged: size page pushes by their page counts and read full pages through the frontier bits");
    r.add("page_groups", FlagType::kInt32, &c.page_groups,
          "paged: grouped launches per level for the cached active pages while copies wait for "
          "slots, else one (0 = one per page)");
    r.add("page_direct", FlagType::kInt64, &c.page_direct,
          "paged: uncached pages with at most this many active words are pushed straight from "
          "pinned host memory over PCIe, without compaction or a copy (0 = off)");
    r.add("pull_head", FlagType::kBool, &c.pull_head,
          "keep each in-list's first entries beside its in-degree for the pull probes");
    r.add("page_words", FlagType::kInt64, &c.page_words, "exact page size in u32 words (0: -pagesize)");
    r.add("page_predict", FlagType::kString, &c.page_predict,
          "paged: next-level page-activity prediction: off|learned|exact");
  }
  return c;
}

bool parse_config(const std::vector<std::string>& args, Config* out,
                  std::vector<std::string>* positional, std::string* err) {
  Config& g = global_flags();
  g = Config();
  std::vector<std::string> pos;
  bool help = false;
  if (!parse_flag_vector(args, &pos, err, &help)) return false;
  if (help) {
    if (err) *err = FlagRegistry::global().usage();
    return false;
  }
  if (g.pagesize <= 0 || g.pagesize > 1023) {
    if (err) *err = "ERROR: -pagesize must be in 1..1023 (Mi words)";
    return false;
  }
  if (g.threshold < 0 || g.threshold > 16) {
    // the compacted copy must fit the half-page tail of each pinned page
    // buffer (src/compress.h:47 with src/graph.h:100)
    if (err) *err = "ERROR: -threshold must be in 0..16 (1/32 pages)";
    return false;
  }
  if (g.num_stream < 1 || g.num_stream > 64 || g.multi < 1 || g.multi > 64) {
    if (err) *err = "ERROR: -num_stream and -multi must be in 1..64";
    return false;
  }
  if (g.tricknum < 0 || g.tricknum > 4095) {
    // the upper bound is the device's CU count, checked when streams open
    if (err) *err = "ERROR: -tricknum must be in 0..(CUs - 1)";
    return false;
  }
  if (g.sssp_delta < 0) {
    if (err) *err = "ERROR: -sssp_delta must be >= 0";
    return false;
  }
  if (g.page_groups < 0 || g.page_groups > 4096) {
    if (err) *err = "ERROR: -page_groups must be in 0..4096";
    return false;
  }
  if (g.page_direct < 0) {
    if (err) *err = "ERROR: -page_direct must be >= 0";
    return false;
  }
  if (g.page_words < 0 || (g.page_words > 0 && g.page_words < 64)) {
    if (err) *err = "ERROR: -page_words must be 0 or at least 64";
    return false;
  }
  if (g.alpha <= 0 || g.beta <= 0) {
    if (err) *err = "ERROR: -alpha and -beta must be positive";
    return false;
  }
  if (g.page_predict != "off" && g.page_predict != "learned" && g.page_predict != "exact") {
    if (err) *err = "ERROR: -page_predict must be off, learned or exact";
    return false;
  }
  if (g.schedule != "fixed" && g.schedule != "push" && g.schedule != "pull" &&
      g.schedule != "learned") {
    if (err) *err = "ERROR: -schedule must be fixed, push, pull or learned";
    return false;
  }
  *out = g;
  if (positional) *positional = pos;
  return true;
}

// ----------------------------------------------------------------- engine

Engine::Engine(const Config& cfg)
    : cfg_(cfg), P_(cfg.page_words > 0 ? (uint64_t)cfg.page_words : (uint64_t)cfg.pagesize << 20) {
  H_ERR(scaph_init(cfg.device));
  main_.reset(new Stream((uint32_t)cfg_.tricknum));
  hcount_.alloc(4096);
  hand_.alloc(4096);
  std::memset(hand_.host(), 0, hand_.bytes());
  pull_tune_.grid_blocks = (uint32_t)std::max(0, cfg.pull_grid);
  pull_tune_.head_outdeg = cfg.pull_head_outdeg ? 1u : 0u;
  pull_tune_.no_prefetch = cfg.pull_prefetch ? 0u : 1u;
  pull_tune_.block_waves = (uint32_t)std::max(0, cfg.pull_waves);
  pull_tune_.waves_per_simd = (uint32_t)std::max(0, cfg.pull_simd);
  pull_tune_.two_phase = cfg.pull_phases == 3 ? 2u : cfg.pull_phases == 2 ? 1u : 0u;
  push_tune_.grid_blocks = (uint32_t)std::max(0, cfg.push_grid);
  push_tune_.edge_rows = cfg.push_rows ? 1u : 0u;
  push_tune_.small_edges = (uint32_t)std::max(0, cfg.push_small);
  push_tune_.coop_edges = (uint32_t)std::max(0, cfg.push_coop);
  push_tune_.scan_single = (uint32_t)std::max(0, cfg.push_scan);
  push_tune_.xcd_slices = (uint32_t)std::max(0, cfg.push_xcd);
}

Engine::~Engine() = default;

void Engine::reset_graph() {
  ++graph_gen_;
  in_head_.reset();
  live_base_.reset();
  n_ = 0;
  m_ = 0;
  host_ = HostCsr();
  pages_out_.clear();
  pages_in_.clear();
  pool_out_.reset();
  pool_in_.reset();
  dpages_out_.reset();
  dptr_out_.reset();
  dpages_in_.reset();
  dptr_in_.reset();
  wpool_.reset();
  weights_loaded_ = false;
  deg_out_.reset();
  deg_in_.reset();
  static_visited_.reset();
  hub_list_.reset();
  num_hubs_ = 0;
  h_deg_out_.clear();
  g_out_ = scaph_graph_desc{};
  g_in_ = scaph_graph_desc{};
  has_in_ = false;
  resident_ = true;
  paged_.reset();
}

uint64_t Engine::device_bytes() const {
  uint64_t b = pool_out_.bytes() + pool_in_.bytes() + deg_out_.bytes() + deg_in_.bytes() +
               static_visited_.bytes() + level_.bytes() + visited_.bytes() + front_[0].bytes() +
               front_[1].bytes() + queue_[0].bytes() + queue_[1].bytes() + ws_.bytes() +
               pws_.bytes() + vsnap_.bytes() + in_head_.bytes() + live_base_.bytes() + hub_list_.bytes() +
               counters_.bytes() + dpages_out_.bytes() + dptr_out_.bytes() + dpages_in_.bytes() +
               dptr_in_.bytes() + page_nodes_.bytes() + page_datas_.bytes() + totals_.bytes() +
               wpool_.bytes();
  if (paged_) {
    const PagedState& ps = *paged_;
    b += ps.pool.bytes() + ps.group_ws.bytes() + ps.d_runs.bytes() + ps.wpool.bytes();
    for (const auto& w : ps.ws) b += w.bytes();
    for (const auto& w : ps.relax_ws) b += w.bytes();
  }
  return b;
}

// Budget of src/app.cuh:281-286, 323-324: max_memory MiB minus the vertex
// arrays (value, stat, degree: 9 bytes per vertex) and the per-page lists
// (24 bytes per page), divided into P-word page slots; -1 when that leaves
// no more than one slot ("vertex data out of gpu memory").
int64_t plan_pool_slots(int64_t max_memory_mib, uint32_t n, uint32_t list_size,
                        uint64_t page_words) {
  int64_t budget = max_memory_mib << 20;
  budget -= (int64_t)n * (4 + 1 + 4);
  budget -= (int64_t)list_size * (4 + 4 + 16);
  const int64_t slot = (int64_t)page_words * 4;
  if (budget <= slot) return -1;
  return std::min<int64_t>(budget / slot, list_size);
}

// The pull expand's transposed pages are budgeted after the out-pages.
void Engine::plan_residency() {
  const int64_t list = (int64_t)pages_out_.size();
  const int64_t pool = plan_pool_slots(cfg_.max_memory, n_, (uint32_t)list, P_);
  if (pool < 0) throw Error(SCAPH_STATUS_INVALID, "vertex data out of gpu memory. exit!");
  pool_size_ = (int)pool;
  resident_ = pool_size_ >= list;
  // slots left over once every out-page has one
  const int64_t spare = ((((int64_t)cfg_.max_memory << 20) - (int64_t)n_ * 9 - list * 24) /
                         ((int64_t)P_ * 4)) - list;
  const int64_t need_in = (int64_t)pages_in_.size();
  has_in_ = cfg_.build_pull && resident_ && need_in > 0 && spare >= need_in;
}

// Weights in the pages' own layout: page p's out-edge k sits at word
// offset(p) + 3 * nodenum + k of the packed pool, and its weight at the same
// offset of wpool_, so a kernel holding an out-edge's address finds its
// weight at a fixed distance.  Pages are consecutive vertex ranges, so page
// p's edges are CSR edges [sum of earlier edgenum, + edgenum).
void Engine::set_weights(const uint32_t* wgh, uint64_t count) {
  if (count != m_ || (m_ && !wgh)) throw Error(SCAPH_STATUS_INVALID, "one weight per edge expected");
  if (!resident_) {
    // paged: each page's weights in pinned memory, copied with the page
    PagedState& ps = *paged_;
    ps.h_wpage.clear();
    ps.h_wpage.resize(pages_out_.size());
    ps.h_wtail.clear();
    ps.h_wtail.resize(pages_out_.size());
    uint64_t e0 = 0;
    for (size_t p = 0; p < pages_out_.size(); ++p) {
      const uint64_t ne = pages_out_[p].edgenum;
      ps.h_wpage[p].alloc(std::max<uint64_t>(ne, 1) * 4);
      if (ps.thres) ps.h_wtail[p].alloc(ps.thres * 4);
      std::copy(wgh + e0, wgh + e0 + ne, ps.h_wpage[p].get<uint32_t>());
      e0 += ne;
    }
    ps.wpool.reset();  // slots are filled on the next SSSP run
    weights_loaded_ = true;
    return;
  }
  uint64_t total = 0;
  std::vector<uint64_t> offs(pages_out_.size());
  for (size_t p = 0; p < pages_out_.size(); ++p) {
    offs[p] = total;
    total += align_words(page_used_words(pages_out_[p], false));
  }
  if (total * 4 != pool_out_.bytes() && !(total == 0 && pool_out_.bytes() == 4))
    throw Error(SCAPH_STATUS_INVALID, "page pool layout mismatch");
  wpool_.alloc(std::max<uint64_t>(total, 1) * 4);
  PinnedBuf stage;
  const uint64_t stage_words = std::max<uint64_t>(P_, 1ull << 22);
  stage.alloc(stage_words * 4);
  uint32_t* host = stage.get<uint32_t>();
  void* st = main_->get();
  uint64_t e0 = 0;
  for (size_t p = 0; p < pages_out_.size(); ++p) {
    const scaph_page& pg = pages_out_[p];
    const uint64_t words = page_used_words(pg, false);
    std::fill(host, host + 3ull * pg.nodenum, 0u);
    std::copy(wgh + e0, wgh + e0 + pg.edgenum, host + 3ull * pg.nodenum);
    H_ERR(scaph_copy_h2d_async(wpool_.get<uint32_t>() + offs[p], host, words * 4, st));
    main_->sync();
    e0 += pg.edgenum;
  }
  weights_loaded_ = true;
}

void Engine::upload_pages(const HostCsr& g, bool weighted, std::vector<scaph_page>* pages,
                          DevBuf* pool, DevBuf* d_pages, DevBuf* d_ptr, scaph_graph_desc* desc) {
  std::vector<uint64_t> offs(pages->size());
  uint64_t total = 0;
  for (size_t p = 0; p < pages->size(); ++p) {
    offs[p] = total;
    total += align_words(page_used_words((*pages)[p], weighted));
  }
  pool->alloc(std::max<uint64_t>(total, 1) * 4);
  // stage through a pinned buffer, a few pages at a time
  const uint64_t stage_words = std::max<uint64_t>(P_, 1ull << 22);
  PinnedBuf stage;
  stage.alloc(stage_words * 4);
  uint32_t* host = stage.get<uint32_t>();
  void* st = main_->get();
  for (size_t p = 0; p < pages->size(); ++p) {
    const uint64_t used = page_used_words((*pages)[p], weighted);
    fill_page(g, (*pages)[p], weighted, host);
    H_ERR(scaph_copy_h2d_async(pool->get<uint32_t>() + offs[p], host, used * 4, st));
    main_->sync();
  }
  std::vector<uint64_t> ptr(pages->size());
  for (size_t p = 0; p < pages->size(); ++p)
    ptr[p] = reinterpret_cast<uint64_t>(pool->get<uint32_t>() + offs[p]);
  d_pages->alloc(std::max<size_t>(pages->size(), 1) * sizeof(scaph_page));
  d_ptr->alloc(std::max<size_t>(pages->size(), 1) * 8);
  H_ERR(scaph_copy_h2d_async(d_pages->get(), pages->data(), pages->size() * sizeof(scaph_page), st));
  H_ERR(scaph_copy_h2d_async(d_ptr->get(), ptr.data(), ptr.size() * 8, st));
  main_->sync();
  desc->page_ptr = d_ptr->get<const uint64_t>();
  desc->pages = d_pages->get<const scaph_page>();
  desc->num_pages = (uint32_t)pages->size();
  desc->num_vertices = g.n;
}

void Engine::finish_setup(const std::vector<uint32_t>* in_deg_host) {
  void* st = main_->get();
  if (has_in_) {
    static_visited_.alloc(bitmap_words(n_) * 4);
    if (!deg_in_.get()) {
      deg_in_.alloc((size_t)n_ * 4);
      H_ERR(scaph_copy_h2d_async(deg_in_.get(), in_deg_host->data(), (size_t)n_ * 4, st));
    }
    H_ERR(scaph_bitmap_from_zero_degree(static_visited_.get<uint32_t>(), deg_in_.get<uint32_t>(),
                                        n_, st));
  }
  alloc_state();
  main_->sync();
}

void Engine::alloc_state() {
  const size_t words = bitmap_words(n_);
  level_.alloc((size_t)n_ * 4);
  visited_.alloc(words * 4);
  if (has_in_) vsnap_.alloc(words * 4);
  front_[0].alloc(words * 4);
  front_[1].alloc(words * 4);
  queue_[0].alloc(std::max<size_t>(n_, 1) * 4);
  queue_[1].alloc(std::max<size_t>(n_, 1) * 4);
  // level counters (four records), then the learned decision's output
  // (word offset 64), all zero to start
  counters_.alloc(1024);
  H_ERR(scaph_fill_u32(counters_.get<uint32_t>(), 0u, 1024 / 4, main_->get()));
  // the resident push expands whole frontiers; the paged mode keeps one
  // page-sized workspace per exec stream instead
  ws_bytes_ = resident_ ? scaph_push_workspace_bytes(std::max<uint32_t>(n_, 1),
                                                      std::max<uint64_t>(m_, 1))
                        : 0;
  if (ws_bytes_) ws_.alloc(ws_bytes_);
  pws_bytes_ = has_in_ ? scaph_pull_workspace_bytes(std::max<uint64_t>(m_, 1)) : 0;
  if (pws_bytes_) pws_.alloc(pws_bytes_);
  const size_t np = std::max<size_t>(pages_out_.size(), 1);
  page_nodes_.alloc(np * 4);
  page_datas_.alloc(np * 4);
  totals_.alloc(64);
}

void Engine::load_csr(HostCsr g, bool keep_host, HostCsr* in_edges) {
  reset_graph();
  if (g.offsets.size() != (size_t)g.n + 1) throw Error(SCAPH_STATUS_INVALID, "malformed CSR");
  n_ = g.n;
  m_ = g.num_edges();
  std::string err;
  if (!page_cuts(g, P_, false, &pages_out_, &err)) throw Error(SCAPH_STATUS_INVALID, err);
  HostCsr t;
  std::vector<uint32_t> in_deg;
  h_deg_out_.resize(n_);
  for (uint32_t v = 0; v < n_; ++v) h_deg_out_[v] = g.degree(v);
  std::vector<uint32_t> hubs;
  if (cfg_.build_pull) {
    if (n_ > 0x7FFFFFFFu) throw Error(SCAPH_STATUS_INVALID, "pull needs fewer than 2^31 vertices");
    if (in_edges) t = std::move(*in_edges);
    else t = transpose(g);
    hubs = select_hubs(h_deg_out_, hub_count());
    encode_hub_in_lists(&t, hubs);
    if (!page_cuts(t, P_, false, &pages_in_, &err)) throw Error(SCAPH_STATUS_INVALID, err);
  }
  plan_residency();
  deg_out_.alloc(std::max<size_t>(n_, 1) * 4);
  H_ERR(scaph_copy_h2d_async(deg_out_.get(), h_deg_out_.data(), (size_t)n_ * 4, main_->get()));
  if (resident_) {
    upload_pages(g, false, &pages_out_, &pool_out_, &dpages_out_, &dptr_out_, &g_out_);
    if (has_in_) {
      num_hubs_ = (uint32_t)hubs.size();
      hub_list_.alloc(std::max<size_t>(hubs.size(), 1) * 4);
      H_ERR(scaph_copy_h2d_async(hub_list_.get(), hubs.data(), hubs.size() * 4, main_->get()));
      upload_pages(t, false, &pages_in_, &pool_in_, &dpages_in_, &dptr_in_, &g_in_);
      in_deg.resize(n_);
      for (uint32_t v = 0; v < n_; ++v) in_deg[v] = t.degree(v);
      if (cfg_.pull_head) {
        // records of the vertices with in-edges only, numbered per 1024-id
        // window (live_base_), as scaph_list_heads_live writes them
        const size_t windows = (bitmap_words(n_) + 31) / 32;
        std::vector<uint32_t> base(windows, 0);
        std::vector<uint32_t> head;
        head.reserve((size_t)n_ * 4);
        uint32_t live = 0;
        const bool live_only = cfg_.pull_head_live;
        const bool indexed = heads_indexable(pages_in_);
        size_t pg = 0;
        for (uint32_t v = 0; v < n_; ++v) {
          if ((v & 1023u) == 0) base[v >> 10] = live;
          while (pg + 1 < pages_in_.size() && pages_in_[pg + 1].left <= v) ++pg;
          const uint32_t d = t.degree(v);
          if (d == 0 && live_only) continue;
          ++live;
          const uint64_t o = t.offsets[v];
          uint32_t r[4];
          for (uint32_t j = 0; j < 3; ++j) r[j] = j < d ? t.ngh[o + j] : 0xFFFFFFFFu;
          r[3] = d;
          if (cfg_.pull_head_outdeg) {
            r[2] = d;
            r[3] = h_deg_out_[v];
          }
          if (indexed) {
            // the list's word offset in its page's edge array (its Node idx)
            const uint64_t idx = o - t.offsets[pages_in_[pg].left];
            r[2] = (uint32_t)(idx << 8) | std::min<uint32_t>(d, 255u);
          }
          head.insert(head.end(), r, r + 4);
        }
        pull_tune_.head_indexed = indexed ? 1u : 0u;
        for (size_t k = (n_ + 1023u) >> 10; k < windows; ++k) base[k] = live;
        // one spare record: nothing reads past the last live vertex's
        head.insert(head.end(), 4, 0xFFFFFFFFu);
        in_head_.alloc(head.size() * 4);
        H_ERR(scaph_copy_h2d_async(in_head_.get(), head.data(), head.size() * 4, main_->get()));
        if (live_only) {
          live_base_.alloc(windows * 4);
          H_ERR(scaph_copy_h2d_async(live_base_.get(), base.data(), windows * 4, main_->get()));
        }
        main_->sync();
      }
    }
    finish_setup(&in_deg);
    // edge weights (the _csr.wgh file) for the SSSP app
    if (g.wgh.size() == m_ && m_ > 0) set_weights(g.wgh.data(), g.wgh.size());
    if (keep_host) host_ = std::move(g);
  } else {
    host_ = std::move(g);
    finish_setup(nullptr);
    setup_paged();
    if (host_.wgh.size() == m_ && m_ > 0) set_weights(host_.wgh.data(), host_.wgh.size());
  }
}

void Engine::load_files(const std::string& base) {
  // fetch_graph (src/graph.h:24-108): the flat CSR files plus graph.info and
  // page.info; page buffers are re-derived from the flat CSR and checked
  // against page.info so that a stale page set is refused.
  HostCsr g;
  std::string err;
  if (!read_csr_files(base, &g, &err)) throw Error(SCAPH_STATUS_INVALID, err);
  GraphInfo info;
  std::vector<scaph_page> listed;
  if (read_graph_info(base, &info, &err)) {
    if (info.nodenum != g.n || (uint32_t)info.edgenum != (uint32_t)g.num_edges())
      throw Error(SCAPH_STATUS_INVALID, "graph.info does not match the CSR files");
    P_ = info.pagesize;  // src/graph.h:86 takes the page size from graph.info
    if (!read_page_info(base, &listed, &err)) throw Error(SCAPH_STATUS_INVALID, err);
    std::vector<scaph_page> cuts;
    if (!page_cuts(g, P_, false, &cuts, &err)) throw Error(SCAPH_STATUS_INVALID, err);
    if (cuts.size() != listed.size() ||
        std::memcmp(cuts.data(), listed.data(), cuts.size() * sizeof(scaph_page)) != 0)
      throw Error(SCAPH_STATUS_INVALID, "page.info does not match the page cut of the CSR");
  }
  // the transposed-page generator's output (pagegencsr ... t), used instead
  // of transposing here once checked to hold exactly the reversed edges
  HostCsr t;
  bool have_t = false;
  if (cfg_.build_pull) {
    if (std::FILE* f = std::fopen((base + "_in_csr.idx").c_str(), "rb")) {
      std::fclose(f);
      if (!read_csr_files(base + "_in", &t, &err)) throw Error(SCAPH_STATUS_INVALID, err);
      if (t.n != g.n || t.num_edges() != g.num_edges() ||
          edge_multiset_digest(t, true) != edge_multiset_digest(g, false))
        throw Error(SCAPH_STATUS_INVALID, base + "_in_csr.* is not the transpose of the graph");
      have_t = true;
    }
  }
  load_csr(std::move(g), true, have_t ? &t : nullptr);
}

void Engine::load_from_config() {
  if (!cfg_.graphfile.empty()) {
    load_files(cfg_.graphfile);
    return;
  }
  if (cfg_.gen.empty()) throw Error(SCAPH_STATUS_INVALID, "no -graphfile and no -gen given");
  if (cfg_.gen.rfind("rmat-device:", 0) == 0) {
    HostCsr dummy;
    std::string err;
    // parse the same keys as the host generator
    RmatParams p;
    std::string rest = cfg_.gen.substr(12);
    size_t pos = 0;
    while (pos < rest.size()) {
      size_t comma = rest.find(',', pos);
      std::string item = rest.substr(pos, comma == std::string::npos ? std::string::npos : comma - pos);
      size_t eq = item.find('=');
      if (eq != std::string::npos) {
        std::string k = item.substr(0, eq), v = item.substr(eq + 1);
        if (k == "scale") p.scale = (uint32_t)std::stoul(v);
        else if (k == "ef") p.edge_factor = (uint32_t)std::stoul(v);
        else if (k == "a") p.a = std::stod(v);
        else if (k == "b") p.b = std::stod(v);
        else if (k == "c") p.c = std::stod(v);
        else if (k == "seed") p.seed = std::stoull(v);
      }
      if (comma == std::string::npos) break;
      pos = comma + 1;
    }
    gen_rmat_device(p, cfg_.check);
    return;
  }
  HostCsr g;
  std::string err;
  if (!generate_from_spec(cfg_.gen, &g, &err)) throw Error(SCAPH_STATUS_INVALID, err);
  load_csr(std::move(g), true);
}

void Engine::gen_rmat_device(const RmatParams& p, bool keep_host_csr) {
  if (p.scale == 0 || p.scale > 31) throw Error(SCAPH_STATUS_INVALID, "rmat scale must be 1..31");
  uint32_t ta, tab, tabc;
  rmat_thresholds(p, &ta, &tab, &tabc);
  const uint64_t raw = (uint64_t)p.edge_factor << p.scale;
  build_device((uint32_t)(1ull << p.scale), raw, keep_host_csr, [&](uint64_t* keys, void* st) {
    H_ERR(scaph_gen_rmat_keys(keys, 0, raw, p.scale, ta, tab, tabc, p.seed, st));
  });
}

void Engine::gen_er_device(uint32_t n, uint64_t m, uint64_t seed, bool keep_host_csr) {
  if (n == 0) throw Error(SCAPH_STATUS_INVALID, "er needs n > 0");
  build_device(n, m, keep_host_csr, [&](uint64_t* keys, void* st) {
    H_ERR(scaph_gen_er_keys(keys, 0, m, n, seed, st));
  });
}

void Engine::gen_zipf(uint32_t n, double x_min, double alpha, uint32_t cap, uint64_t seed) {
  if (n == 0 || !(alpha > 1.0) || !(x_min > 0.0) || cap == 0)
    throw Error(SCAPH_STATUS_INVALID, "zipf needs n > 0, alpha > 1, x_min > 0, cap > 0");
  std::vector<uint32_t> deg(n);
  {
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
      th.emplace_back([&, t] {
        const uint64_t lo = (uint64_t)n * t / nt, hi = (uint64_t)n * (t + 1) / nt;
        for (uint64_t v = lo; v < hi; ++v)
          deg[v] = zipf_degree(seed, (uint32_t)v, x_min, alpha, cap);
      });
    for (auto& x : th) x.join();
  }
  uint64_t raw = 0;
  for (uint32_t d : deg) raw += d;
  void* st = main_->get();
  HostCsr g;
  g.n = n;
  g.offsets.resize((size_t)n + 1);
  {
    DevBuf d_deg((size_t)n * 4), offs(((size_t)n + 1) * 8), keys(std::max<uint64_t>(raw, 1) * 8),
        tmp(std::max<uint64_t>(raw, 1) * 8), ngh(std::max<uint64_t>(raw, 1) * 4),
        deg2((size_t)n * 4);
    H_ERR(scaph_copy_h2d_async(d_deg.get(), deg.data(), (size_t)n * 4, st));
    H_ERR(scaph_degrees_to_offsets(d_deg.get<uint32_t>(), offs.get<uint64_t>(), n, st));
    H_ERR(scaph_gen_zipf_keys(keys.get<uint64_t>(), offs.get<uint64_t>(), n, 0, raw, seed, st));
    uint64_t E = 0;
    H_ERR(scaph_csr_from_keys(keys.get<uint64_t>(), tmp.get<uint64_t>(), raw, n, 0, nullptr,
                              offs.get<uint64_t>(), ngh.get<uint32_t>(), deg2.get<uint32_t>(), &E,
                              st));
    g.ngh.resize(E);
    H_ERR(scaph_copy_d2h_async(g.offsets.data(), offs.get(), g.offsets.size() * 8, st));
    if (E) H_ERR(scaph_copy_d2h_async(g.ngh.data(), ngh.get(), E * 4, st));
    main_->sync();
  }
  load_csr(std::move(g), true);
}

void Engine::build_device(uint32_t n, uint64_t raw, bool keep_host_csr,
                          const std::function<void(uint64_t*, void*)>& keygen) {
  reset_graph();
  void* st = main_->get();
  n_ = n;
  DevBuf keys(raw * 8), tmp(raw * 8), offs(((size_t)n + 1) * 8), ngh(std::max<uint64_t>(raw, 1) * 4);
  std::vector<uint64_t> h_off((size_t)n + 1);
  std::string err;
  for (int dir = 0; dir < 2; ++dir) {
    if (dir == 1 && !cfg_.build_pull) break;
    DevBuf deg((size_t)n * 4);
    DevBuf hub_idx;
    if (dir == 1) {
      // hubs from the out-degrees (h_deg_out_ holds them after the first pass)
      const std::vector<uint32_t> hubs = select_hubs(h_deg_out_, hub_count());
      std::vector<uint32_t> idx(n, 0xFFFFFFFFu);
      for (uint32_t i = 0; i < hubs.size(); ++i) idx[hubs[i]] = i;
      hub_idx.alloc((size_t)n * 4);
      H_ERR(scaph_copy_h2d_async(hub_idx.get(), idx.data(), (size_t)n * 4, st));
      num_hubs_ = (uint32_t)hubs.size();
      hub_list_.alloc(std::max<size_t>(hubs.size(), 1) * 4);
      H_ERR(scaph_copy_h2d_async(hub_list_.get(), hubs.data(), hubs.size() * 4, st));
      main_->sync();
    }
    keygen(keys.get<uint64_t>(), st);
    uint64_t E = 0;
    H_ERR(scaph_csr_from_keys(keys.get<uint64_t>(), tmp.get<uint64_t>(), raw, n, dir,
                              hub_idx.get<uint32_t>(), offs.get<uint64_t>(), ngh.get<uint32_t>(),
                              deg.get<uint32_t>(), &E, st));
    H_ERR(scaph_copy_d2h_async(h_off.data(), offs.get(), h_off.size() * 8, st));
    main_->sync();
    if (dir == 0) m_ = E;
    else if (E != m_) throw Error(SCAPH_STATUS_INVALID, "transposed edge count differs");
    std::vector<scaph_page>& pages = dir == 0 ? pages_out_ : pages_in_;
    if (!page_cuts_raw(h_off.data(), n, P_, false, &pages, &err))
      throw Error(SCAPH_STATUS_INVALID, err);
    if (dir == 0) {
      if (keep_host_csr) {
        host_.n = n;
        host_.offsets = h_off;
        host_.ngh.resize(E);
        H_ERR(scaph_copy_d2h_async(host_.ngh.data(), ngh.get(), E * 4, st));
        main_->sync();
      }
      // page budget is decided once the transposed cut is known (below)
    }
    if (dir == 1 || !cfg_.build_pull) {
      plan_residency();
      if (!resident_)
        throw Error(SCAPH_STATUS_INVALID,
                    "device-generated graphs must fit the -max_memory budget (raise it or load "
                    "from files for out-of-core runs)");
    }
    // page buffers packed in one allocation
    DevBuf& pool = dir == 0 ? pool_out_ : pool_in_;
    DevBuf& dpages = dir == 0 ? dpages_out_ : dpages_in_;
    DevBuf& dptr = dir == 0 ? dptr_out_ : dptr_in_;
    scaph_graph_desc& desc = dir == 0 ? g_out_ : g_in_;
    uint64_t total = 0;
    std::vector<uint64_t> ptr(pages.size());
    for (size_t i = 0; i < pages.size(); ++i) {
      ptr[i] = total;
      total += align_words(page_used_words(pages[i], false));
    }
    pool.alloc(std::max<uint64_t>(total, 1) * 4);
    for (auto& x : ptr) x = reinterpret_cast<uint64_t>(pool.get<uint32_t>() + x);
    dpages.alloc(std::max<size_t>(pages.size(), 1) * sizeof(scaph_page));
    dptr.alloc(std::max<size_t>(pages.size(), 1) * 8);
    H_ERR(scaph_copy_h2d_async(dpages.get(), pages.data(), pages.size() * sizeof(scaph_page), st));
    H_ERR(scaph_copy_h2d_async(dptr.get(), ptr.data(), ptr.size() * 8, st));
    H_ERR(scaph_pages_from_csr(offs.get<uint64_t>(), ngh.get<uint32_t>(),
                               dpages.get<scaph_page>(), dptr.get<uint64_t>(),
                               (uint32_t)pages.size(), n, st));
    const bool indexed = dir == 1 && cfg_.pull_head && heads_indexable(pages);
    if (dir == 1) pull_tune_.head_indexed = indexed ? 1u : 0u;
    if (dir == 1 && cfg_.pull_head && !cfg_.pull_head_live) {
      in_head_.alloc((size_t)n * 16);
      if (indexed)
        H_ERR(scaph_list_heads_indexed(offs.get<uint64_t>(), ngh.get<uint32_t>(), n,
                                       deg_out_.get<uint32_t>(), dpages.get<scaph_page>(),
                                       (uint32_t)pages.size(), nullptr, nullptr,
                                       in_head_.get<uint32_t>(), st));
      else
        H_ERR(scaph_list_heads_from_csr(offs.get<uint64_t>(), ngh.get<uint32_t>(), n,
                                        cfg_.pull_head_outdeg ? deg_out_.get<uint32_t>() : nullptr,
                                        in_head_.get<uint32_t>(), st));
    } else if (dir == 1 && cfg_.pull_head) {
      // records of the vertices with in-edges only (deg holds the in-degrees)
      static_visited_.alloc(bitmap_words(n) * 4);
      H_ERR(scaph_bitmap_from_zero_degree(static_visited_.get<uint32_t>(), deg.get<uint32_t>(), n,
                                          st));
      const size_t windows = (bitmap_words(n) + 31) / 32;
      live_base_.alloc(windows * 4);
      H_ERR(scaph_live_window_base(static_visited_.get<uint32_t>(), n,
                                   live_base_.get<uint32_t>(), st));
      uint32_t live = 0, last = 0;
      H_ERR(scaph_copy_d2h_async(&last, live_base_.get<uint32_t>() + windows - 1, 4, st));
      main_->sync();
      {
        // live vertices = base of the last window + those in it
        std::vector<uint32_t> tail(32, ~0u);
        const size_t w0 = (windows - 1) * 32;
        const size_t nw = std::min<size_t>(32, bitmap_words(n) - w0);
        H_ERR(scaph_copy_d2h_async(tail.data(), static_visited_.get<uint32_t>() + w0, nw * 4, st));
        main_->sync();
        live = last;
        for (size_t k = 0; k < nw; ++k) live += (uint32_t)__builtin_popcount(~tail[k]);
      }
      in_head_.alloc(((size_t)live + 1) * 16);
      if (indexed)
        H_ERR(scaph_list_heads_indexed(offs.get<uint64_t>(), ngh.get<uint32_t>(), n,
                                       deg_out_.get<uint32_t>(), dpages.get<scaph_page>(),
                                       (uint32_t)pages.size(), static_visited_.get<uint32_t>(),
                                       live_base_.get<uint32_t>(), in_head_.get<uint32_t>(), st));
      else
        H_ERR(scaph_list_heads_live(offs.get<uint64_t>(), ngh.get<uint32_t>(), n,
                                    cfg_.pull_head_outdeg ? deg_out_.get<uint32_t>() : nullptr,
                                    static_visited_.get<uint32_t>(), live_base_.get<uint32_t>(),
                                    in_head_.get<uint32_t>(), st));
    }
    main_->sync();
    desc.page_ptr = dptr.get<const uint64_t>();
    desc.pages = dpages.get<const scaph_page>();
    desc.num_pages = (uint32_t)pages.size();
    desc.num_vertices = n;
    if (dir == 0) {
      deg_out_ = std::move(deg);
      h_deg_out_.resize(n);
      H_ERR(scaph_copy_d2h_async(h_deg_out_.data(), deg_out_.get(), (size_t)n * 4, st));
      main_->sync();
    } else {
      deg_in_ = std::move(deg);
    }
  }
  if (!has_in_) {
    in_head_.reset();
    live_base_.reset();
    pool_in_.reset();
    dpages_in_.reset();
    dptr_in_.reset();
    deg_in_.reset();
    hub_list_.reset();
    num_hubs_ = 0;
    g_in_ = scaph_graph_desc{};
  }
  keys.reset();
  tmp.reset();
  ngh.reset();
  offs.reset();
  finish_setup(nullptr);
}

uint32_t Engine::out_degree(uint32_t v) {
  if (v >= n_) throw Error(SCAPH_STATUS_INVALID, "vertex out of range");
  if (!h_deg_out_.empty()) return h_deg_out_[v];
  uint32_t d = 0;
  H_ERR(scaph_copy_d2h_async(&d, deg_out_.get<uint32_t>() + v, 4, main_->get()));
  main_->sync();
  return d;
}

uint32_t Engine::in_degree(uint32_t v) {
  if (v >= n_ || !deg_in_.get()) return 0;
  uint32_t d = 0;
  H_ERR(scaph_copy_d2h_async(&d, deg_in_.get<uint32_t>() + v, 4, main_->get()));
  main_->sync();
  return d;
}

bool Engine::heads_indexable(const std::vector<scaph_page>& in_pages) const {
  if (!cfg_.pull_head_index || !cfg_.pull_head_outdeg) return false;
  for (const auto& p : in_pages)
    if (p.edgenum >= (1u << 24)) return false;
  return true;
}

// Beamer's direction rule (GAP's form): push -> pull when the frontier's
// edges exceed the unexplored edges / alpha; pull -> push once the
// frontier shrinks and holds at most n / beta vertices.
Dir Engine::beamer(const SchedState& s) const {
  if (!has_in_) return Dir::kPush;
  if (cfg_.schedule == "push") return Dir::kPush;
  if (cfg_.schedule == "pull") return Dir::kPull;
  if (s.cur == Dir::kPush)
    return (double)s.m_f > (double)s.m_u / cfg_.alpha ? Dir::kPull : Dir::kPush;
  const bool shrinking = s.n_f < s.n_f_prev;
  return (shrinking && (double)s.n_f <= (double)s.n / cfg_.beta) ? Dir::kPush : Dir::kPull;
}

RunStats Engine::bfs(uint32_t src) {
  if (n_ == 0) throw Error(SCAPH_STATUS_INVALID, "no graph loaded");
  if (src >= n_) throw Error(SCAPH_STATUS_INVALID, "source vertex out of range");
  return resident_ ? bfs_resident(src) : bfs_paged(src);
}

void Engine::levels_to_host(uint32_t* out) const {
  H_ERR(scaph_copy_d2h_async(out, level_.get(), (size_t)n_ * 4, main_->get()));
  main_->sync();
}

std::array<uint64_t, 2> Engine::validate(uint32_t src) {
  if (!resident_) throw Error(SCAPH_STATUS_INVALID, "validation needs resident pages");
  DevBuf out(64);
  H_ERR(scaph_bfs_validate(&g_out_, has_in_ ? &g_in_ : nullptr, hub_list_.get<uint32_t>(),
                           level_.get<uint32_t>(), src,
                           out.get<unsigned long long>(), main_->get()));
  std::array<uint64_t, 2> r{};
  H_ERR(scaph_copy_d2h_async(r.data(), out.get(), 16, main_->get()));
  main_->sync();
  return r;
}

void cpu_bfs(const HostCsr& g, uint32_t src, uint32_t inf, std::vector<uint32_t>* out) {
  out->assign(g.n, inf);
  if (src >= g.n) return;
  std::vector<uint32_t> cur, next;
  (*out)[src] = 0;
  cur.push_back(src);
  while (!cur.empty()) {
    for (uint32_t u : cur) {
      const uint32_t lu = (*out)[u];
      for (uint64_t k = g.offsets[u]; k < g.offsets[(size_t)u + 1]; ++k) {
        const uint32_t v = g.ngh[k];
        if (lu + 1 < (*out)[v]) {
          (*out)[v] = lu + 1;
          next.push_back(v);
        }
      }
    }
    cur.swap(next);
    next.clear();
  }
}

void cpu_wcc(const HostCsr& g, std::vector<uint32_t>* out) {
  std::vector<uint32_t>& val = *out;
  val.resize(g.n);
  for (uint32_t v = 0; v < g.n; ++v) val[v] = v;
  for (bool changed = true; changed;) {
    changed = false;
    for (uint32_t u = 0; u < g.n; ++u)
      for (uint64_t k = g.offsets[u]; k < g.offsets[(size_t)u + 1]; ++k) {
        const uint32_t v = g.ngh[k];
        if (val[u] < val[v]) {
          val[v] = val[u];
          changed = true;
        }
      }
  }
}

void cpu_sssp(const HostCsr& g, uint32_t src, std::vector<uint32_t>* out) {
  std::vector<uint64_t> dist(g.n, UINT64_MAX);
  using Item = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  if (src < g.n && g.wgh.size() == g.ngh.size()) {
    dist[src] = 0;
    heap.push({0, src});
  }
  while (!heap.empty()) {
    const Item it = heap.top();
    heap.pop();
    if (it.first != dist[it.second]) continue;
    for (uint64_t k = g.offsets[it.second]; k < g.offsets[(size_t)it.second + 1]; ++k) {
      const uint64_t nd = it.first + g.wgh[k];
      if (nd < dist[g.ngh[k]]) {
        dist[g.ngh[k]] = nd;
        heap.push({nd, g.ngh[k]});
      }
    }
  }
  out->resize(g.n);
  for (uint32_t v = 0; v < g.n; ++v)
    (*out)[v] = dist[v] == UINT64_MAX ? SCAPH_RELAX_INF : (uint32_t)dist[v];
}

}  // namespace scaph
