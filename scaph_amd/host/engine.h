// engine.h — the host BFS engine: graph residency, the per-level loop and
// the direction schedule.  It plays the role of the reference's App<>
// (src/app.cuh:72-135): App<> in app.h is a thin template facade over it
// that keeps the reference's method names.
#ifndef SCAPH_HOST_ENGINE_H
#define SCAPH_HOST_ENGINE_H

#include <algorithm>
#include <array>
#include <cstdint>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "gen.h"
#include "graph_io.h"
#include "hip_abi.h"

namespace scaph {

// Flags of the reference (src/app.cuh:38-52 values, 57-70 names) plus the
// additions this engine needs.  FLAGS_threshold defaults to 3, the value
// app/bfs.cu:129 forces before constructing the app.
struct Config {
  int32_t pagesize = 10;      // page size in Mi u32 words
  int32_t max_memory = 4096;  // device budget in MiB
  int32_t multi = 1;          // re-execution ranks of cached pages
  int32_t num_stream = 5;     // exec streams for paged mode
  int32_t num_blks = 256;     // accepted; grids are sized from the work
  int32_t num_thds = 256;     // accepted; blocks are 256 threads (4 wave64)
  int32_t threshold = 3;      // cold-page compaction threshold, in 1/32 pages
  bool check = false;
  int32_t tricknum = 0;       // CUs withheld from every engine stream (a CU mask)
  int64_t src = 0;
  std::string graphfile;
  // additions
  std::string gen;               // generator spec used when graphfile is empty
  double alpha = 14.0;           // push -> pull when m_f > m_u / alpha
  double beta = 24.0;            // pull -> push when n_f shrinks below n / beta
  std::string schedule = "fixed";  // fixed | push | pull | learned
  std::string model;             // learned-schedule weights file
  int32_t rounds = 10;           // timed rounds of the CLI (app/bfs.cu:137)
  int32_t device = 0;
  std::string levels_out;        // write final levels (u32) here
  bool compat63 = false;         // print/write levels with the reference's cap of 63
  std::string trace;             // per-iteration JSONL trace file
  bool build_pull = true;        // keep transposed pages for the pull expand
  bool pull_head = true;         // 16-byte in-list head records in front of the in-pages
  int32_t pull_grid = 0;         // cap on pull blocks (0: as many as stay resident)
  int32_t pull_waves = 0;        // waves per pull block, 4 or 8 (0: the kernel default, 8)
  int32_t pull_simd = 0;         // 8: pull registers capped for 8 waves per SIMD (0: compiler)
  int32_t pull_phases = 3;       // 2: a window's heads are probed before its lists (1: together;
                                 // 3: as 2, the head probes pipelined one batch ahead)
  int32_t pull_hubs = 131072;    // in-list sources tagged as hubs, frontier bits in LDS
  int32_t pull_sparse = 128;     // pull over a list of the unvisited vertices when fewer than
                                 // n / this are left (known after a pull level; 0 = never)
  int32_t pull_summary = 5;      // pull: frontier summary of one bit per 2^this vertices tested
                                 // before each non-hub probe (0 = never, else >= 5) ...
  int32_t pull_summary_div = 4;  // ... on levels whose frontier has fewer than n / 2^shift / this
                                 // vertices (0: on every pull level)
  bool pull_head_outdeg = true;  // head records {e0, e1, in-degree, out-degree}
  bool pull_head_live = true;    // head records only for vertices with in-edges (live rank)
  bool pull_prefetch = true;     // request the next batch's head records early
  bool pull_head_index = true;   // head records carry the list's page offset (no Node read past
                                 // the head; needs -pull_head_outdeg and pages below 2^24 edges)
  bool learned_fused = true;     // a scalar-feature model decides inside the level hand-off
  double learned_margin = 0.5;   // the model decides only when |pull logit| >= this; nearer
                                 // the boundary the fixed rule does
  int32_t sssp_delta = 0;        // resident SSSP: bucket width of near/far rounds (0 = off)
  int32_t push_grid = 0;         // cap on push expand blocks (0: one per tile, at most 8192)
  bool push_rows = true;         // push expand reads coalesced 256-edge rows
  int32_t push_small = 0;        // out-edges below which a push level is one launch (0: 8192)
  int32_t push_coop = 0;         // frontiers of <= 256 vertices with at most this many out-edges:
                                 // one launch, edges spread over the grid (0: 2^20; 1: never)
  int32_t push_scan = 0;         // prep blocks scanned by one block; more: 3 launches (0: 1024)
  int32_t push_xcd = 65536;      // out-edges/1024 from which a push level expands XCD-sliced
  bool page_hint = true;         // paged: size page pushes by their page counts, read full
                                 // pages through the frontier bits
  int32_t page_groups = 8;       // paged: cached active pages of a level go out in about this
                                 // many grouped launches while copies wait for slots, else in
                                 // one (0: one launch per page)
  int64_t page_direct = 65536;   // paged: an uncached page with at most this many active words
                                 // is pushed from pinned host memory, no compaction or copy
                                 // (0 = off; needs -page_groups > 0)
  int64_t page_words = 0;        // exact page size in words (overrides -pagesize; tests)
  std::string page_predict = "off";  // paged: next-level page-activity prediction for slot reuse,
                                     // look-ahead copies and whole-page copies (off | learned |
                                     // exact: from set_page_levels, the bound for any predictor)
};

// Registers the flags above (once) against a process-wide Config.
Config& global_flags();
// Resets the global flags to defaults, parses args, returns a copy.
bool parse_config(const std::vector<std::string>& args, Config* out,
                  std::vector<std::string>* positional, std::string* err);

enum class Dir : int { kPush = 0, kPull = 1 };

// The per-iteration record — also the learned schedule's feature source.
struct LevelRecord {
  uint32_t level = 0;
  int dir = 0;
  uint64_t n_f = 0, m_f = 0, m_u = 0;   // frontier before the step
  uint64_t n_next = 0, m_next = 0;      // reached by the step
  uint64_t scanned = 0;                 // edges examined
  uint64_t examined = 0;                // pull: vertices whose in-list was read
  float ms = 0;                         // step time incl. host turnaround
  float kernel_ms = 0;                  // expand kernels alone
  std::array<float, 12> page_hist{};    // histogram of per-page active counts
  // paged mode: pages active / already cached at the level's start, whole
  // and compacted copies, copies made ahead for the next level, H2D bytes
  // and copy-stream busy time
  uint32_t pages_active = 0, pages_cached = 0, pages_full = 0, pages_compacted = 0;
  uint32_t pages_prefetched = 0;
  uint32_t pages_direct = 0;  // cold pages pushed from pinned host memory (-page_direct)
  uint64_t bytes_h2d = 0;
  float h2d_ms = 0;
  float logit = 0;  // learned schedule: the model's pull logit for this level (0 otherwise)
  bool sparse = false;  // a pull over the list of the unvisited vertices (-pull_sparse)
};

struct RunStats {
  uint32_t src = 0;
  double ms = 0;               // t_BFS: init .. levels final on device
  uint64_t n_reached = 0;      // n_c
  uint64_t m_reached = 0;      // m_c = sum of out-degrees of reached vertices
  uint32_t depth = 0;          // number of levels
  std::vector<LevelRecord> levels;
  // paged (out-of-core) mode
  uint64_t page_execs = 0, page_execs_cached = 0, pages_full = 0, pages_compacted = 0;
  uint64_t pages_prefetched = 0;  // copies made ahead for the next level
  uint64_t pages_direct = 0;      // cold pages pushed from pinned host memory
  uint64_t bytes_h2d = 0;
  double h2d_ms = 0;
};

// Direction policy.  The default is Beamer's rule with the flags' alpha
// and beta; the learned schedule supplies its own.
struct SchedState {
  Dir cur;
  uint32_t level;
  uint64_t n, m;                 // graph totals
  uint64_t n_f, m_f, m_u, n_f_prev;
  const LevelRecord* last;       // record of the previous step (or null)
};
using SchedFn = std::function<Dir(const SchedState&)>;

class Engine {
 public:
  explicit Engine(const Config& cfg);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Graph loading (replaces fetch_graph + load_graph, src/graph.h:24-108,
  // src/app.cuh:267-345).
  void load_files(const std::string& base);
  // in_edges: the graph's transpose when already at hand (read from the
  // transposed-page generator's files); otherwise it is computed here.
  void load_csr(HostCsr g, bool keep_host = true, HostCsr* in_edges = nullptr);
  void gen_rmat_device(const RmatParams& p, bool keep_host_csr);
  void gen_er_device(uint32_t n, uint64_t m, uint64_t seed, bool keep_host_csr);
  // Power-law digraph (zipf_degree out-degrees, uniform destinations):
  // edges built and sorted on the device, then loaded like a host CSR, so
  // a graph beyond the -max_memory budget runs out of core.
  void gen_zipf(uint32_t n, double x_min, double alpha, uint32_t cap, uint64_t seed);
  void load_from_config();  // graphfile, else gen spec

  // One BFS from src.  Levels stay on the device (levels_to_host to read).
  RunStats bfs(uint32_t src);
  // The reference's other App plugins on the resident graph
  // (app/wcc.cu, app/sssp.cu): label-correcting push rounds until no value
  // changes; the values land where the levels do (levels_to_host).  WCC:
  // value[v] = the least vertex id that reaches v.  SSSP: shortest path
  // lengths over the edge weights, 0x7FFFFFFF for unreached.
  RunStats wcc();
  RunStats sssp(uint32_t src);
  // The device state a plugin app initialises in its gpu_init (the
  // reference's public d_value / d_stat and streams, src/app.cuh:72-135):
  // the per-vertex values (n u32), the active-vertex bitmap (active_words
  // u32, bit v of word v/32; the reference's bool stat[]), the out-degrees,
  // the engine's exec stream and, in paged mode, its page copy stream.
  // Valid once a graph is loaded; buffers belong to the engine.
  struct DeviceState {
    uint32_t* value;
    uint32_t* active;
    size_t active_words;
    const uint32_t* degree;
    void* stream;
    void* copy_stream;
    uint32_t n;
  };
  DeviceState device_state();
  // Label-correcting rounds from the values and active set the app wrote
  // into device_state(): every out-edge (u, v) of an active u offers
  // value[u] (mode SCAPH_RELAX_WCC) or value[u] + w (SCAPH_RELAX_SSSP, edge
  // weights required) to v, a lowered v is active in the next round, and
  // the rounds stop when none is (the cond/update/active functors of
  // app/wcc_kernel.cuh and app/sssp_kernel.cuh run by src/app.cuh:220-265).
  RunStats relax_prepared(int mode);
  // Edge weights in CSR edge order (the _csr.wgh file), for SSSP; set
  // automatically when a loaded CSR carries them.
  void set_weights(const uint32_t* wgh, uint64_t count);
  bool has_weights() const { return weights_loaded_; }
  void levels_to_host(uint32_t* out) const;
  // Device-side BFS invariant check; returns {edge violations, orphans}.
  std::array<uint64_t, 2> validate(uint32_t src);

  void set_schedule(SchedFn fn) { sched_ = std::move(fn); }
  // Paged mode, -page_predict=exact: the levels of a previous BFS from the
  // source about to run (u32, unreached = 0xFFFFFFFF).  The scheduler then
  // knows every page's future activity: the bound no predictor can beat.
  void set_page_levels(const uint32_t* levels);
  // Learned schedule on the device: bf16 MLP parameters and the feature
  // normalisation; used by the level loop when -schedule=learned.
  void set_device_model(const uint16_t* params, const float* mean, const float* scale);
  bool device_model_ready() const;
  // indexed pull heads (-pull_head_index): with outdeg records and every
  // in-page holding fewer than 2^24 edges (the offset field's width)
  bool heads_indexable(const std::vector<scaph_page>& in_pages) const;
  // Training traces: at every level both directions run from the same
  // state (min of `reps` timings each), the faster one continues.
  struct ExploreRecord {
    float x[16];             // raw features of the frontier entering the level
    float t_push_ms, t_pull_ms;
    float pages_active_next;  // fraction of pages active at the next level
    uint32_t level;
    uint64_t n_f, m_f, n_next;
  };
  std::vector<ExploreRecord> explore(uint32_t src, int reps);
  void set_config_schedule(const std::string& s, double alpha, double beta) {
    cfg_.schedule = s;
    cfg_.alpha = alpha;
    cfg_.beta = beta;
  }
  Dir beamer(const SchedState& s) const;

  uint32_t num_vertices() const { return n_; }
  uint64_t num_edges() const { return m_; }
  uint32_t out_degree(uint32_t v);
  uint32_t in_degree(uint32_t v);
  bool resident() const { return resident_; }
  bool has_pull() const { return has_in_; }
  const HostCsr* host_csr() const { return host_.n ? &host_ : nullptr; }
  const std::vector<scaph_page>& pages() const { return pages_out_; }
  const Config& config() const { return cfg_; }
  uint64_t page_words() const { return P_; }
  int pool_size() const { return pool_size_; }
  uint64_t device_bytes() const;
  // Device pointers (for the C-ABI tests and the batched/learned parts).
  const scaph_graph_desc& out_desc() const { return g_out_; }
  const scaph_graph_desc& in_desc() const { return g_in_; }
  const uint32_t* d_level() const { return level_.get<uint32_t>(); }
  const uint32_t* d_degree() const { return deg_out_.get<uint32_t>(); }
  const uint32_t* d_hub_list() const { return hub_list_.get<uint32_t>(); }
  uint32_t num_hubs() const { return num_hubs_; }
  void* stream() const { return main_->get(); }
  // Changes whenever a graph is (re)loaded; lets companions drop state.
  uint64_t graph_generation() const { return graph_gen_; }

 private:
  struct PagedState;
  void reset_graph();
  void upload_pages(const HostCsr& g, bool weighted, std::vector<scaph_page>* pages, DevBuf* pool,
                    DevBuf* d_pages, DevBuf* d_ptr, scaph_graph_desc* desc);
  void finish_setup(const std::vector<uint32_t>* in_deg_host);
  void alloc_state();
  void plan_residency();
  void setup_paged();
  void build_device(uint32_t n, uint64_t raw_edges, bool keep_host_csr,
                    const std::function<void(uint64_t*, void*)>& keygen);
  RunStats bfs_resident(uint32_t src);
  struct Front {  // which buffers hold the current frontier, in which forms
    int qa = 0, fa = 0;
    bool queue_ok = true, front_ok = false;
    bool snap_ok = false;  // vsnap_ holds visited from before the push that made the queue
  };
  void ensure_forms(Front& f, bool queue, bool bitmap, uint64_t n_f);
  // sparse: a pull level over a list of the few unvisited vertices left
  void launch_step(Dir d, Front& f, uint32_t cur, uint64_t n_f, uint64_t m_f,
                   bool sparse = false);
  struct DeviceModel;
  std::unique_ptr<DeviceModel> model_;
  void ensure_model_buffers();
  RunStats bfs_paged(uint32_t src);
  bool predict_launch(uint32_t cur, uint64_t n_f, uint64_t m_f, uint64_t m_u);
  void predict_apply(uint32_t cur, bool learned);
  void page_histogram(const uint32_t* nodes, LevelRecord* rec);
  // WCC / SSSP rounds in paged mode: the value array, and the weight pool
  // (same slots as the page pool) for SSSP
  struct RelaxRound {
    uint32_t* value;
    const uint32_t* weights;
  };
  void schedule_pages(int fa, uint32_t cur, const RelaxRound* relax, RunStats* rs);
  RunStats relax_paged(int mode, uint32_t src, bool prepared);

  Config cfg_;
  uint64_t P_;  // page size in words
  uint32_t n_ = 0;
  uint64_t m_ = 0;
  bool resident_ = true;
  bool has_in_ = false;
  int pool_size_ = 0;
  HostCsr host_;                  // kept when available (check(), paged mode)
  std::vector<scaph_page> pages_out_, pages_in_;
  DevBuf pool_out_, pool_in_;     // packed resident pages
  DevBuf wpool_;                  // edge weights laid out like pool_out_ (SSSP)
  bool weights_loaded_ = false;   // resident: wpool_; paged: PagedState::h_wpage
  DevBuf relax_ws_;               // chunk list of the WCC/SSSP rounds
  DevBuf split_;                  // near/far split counters (3 u64)
  RunStats relax(int mode, uint32_t src, bool prepared = false);
  uint64_t count_active(int fa);
  DevBuf dpages_out_, dptr_out_, dpages_in_, dptr_in_;
  scaph_graph_desc g_out_{}, g_in_{};
  DevBuf deg_out_, deg_in_, static_visited_;
  DevBuf hub_list_;  // hub vertex ids (pull in-lists refer to them by index)
  DevBuf in_head_;   // per vertex with in-edges: first in-list entries, in-degree, out-degree
  DevBuf live_base_; // per 1024-id window: vertices with in-edges before it (in_head_ numbering)
  uint32_t num_hubs_ = 0;
  // -pull_hubs clamped to what the pull kernel's LDS holds
  uint32_t hub_count() const {
    return (uint32_t)std::min<int64_t>(std::max<int64_t>(cfg_.pull_hubs, 0), 262144);
  }
  // BFS state
  DevBuf level_, visited_, front_[2], queue_[2], counters_, ws_, page_nodes_, page_datas_, totals_;
  size_t ws_bytes_ = 0;
  DevBuf pws_;  // pull workspace (deferred heavy in-lists)
  DevBuf vsnap_;  // visited before the latest push level (its frontier = visited & ~vsnap_)
  size_t pws_bytes_ = 0;
  std::vector<uint32_t> h_deg_out_;
  std::unique_ptr<Stream> main_;
  PinnedBuf hcount_;  // pinned landing zone for per-level counters
  MappedBuf hand_;    // resident level hand-off: counters (+ decision) and a flag, kernel-written
  uint32_t hand_seq_ = 0;
  scaph_pull_tuning pull_tune_{};
  scaph_push_tuning push_tune_{};
  void wait_handoff(uint32_t seq);
  std::vector<std::unique_ptr<Event>> ev_;
  SchedFn sched_;
  std::unique_ptr<PagedState> paged_;
  DevBuf snap_level_, snap_visited_, snap_queue_, snap_front_;
  uint64_t graph_gen_ = 0;
};

// Page slots of the device pool for a -max_memory budget of max_memory_mib
// (src/app.cuh:281-286, 323-324): the budget less 9 bytes per vertex and 24
// per page, in whole P-word slots, at most one per page; -1 when the vertex
// arrays leave no more than one slot (the reference's "vertex data out of
// gpu memory" exit).
int64_t plan_pool_slots(int64_t max_memory_mib, uint32_t n, uint32_t list_size,
                        uint64_t page_words);

// CPU BFS used by bfs::check() (app/bfs.cu:67-117): FIFO level BFS over the
// out-edges with levels starting at `inf`; a vertex is relaxed when
// level(u) + 1 < level(v).  With inf = 63 it reproduces the reference's
// capped levels; with inf = SCAPH_UNREACHED it gives exact BFS depths.
void cpu_bfs(const HostCsr& g, uint32_t src, uint32_t inf, std::vector<uint32_t>* out);
// The reference apps' CPU checks: WCC label sweeps (app/wcc.cu:60-85) and
// Dijkstra over g.wgh (0x7FFFFFFF unreached) for SSSP, whose reference check
// is a stub (app/sssp.cu:57-61).
void cpu_wcc(const HostCsr& g, std::vector<uint32_t>* out);
void cpu_sssp(const HostCsr& g, uint32_t src, std::vector<uint32_t>* out);

// Reference cap: depth >= 63 (or unreached) reads as 63 (app/bfs.cu:18).
inline uint32_t compat63(uint32_t lvl) { return lvl <= 62 ? lvl : 63u; }

}  // namespace scaph

#endif
