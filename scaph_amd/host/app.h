// app.h — the per-app interface of the reference (src/app.cuh:72-135,
// app/bfs.cu:39-117) over the MI355X engine.
//
// App<vertex_t,index_t,value_t> keeps the reference's public surface:
// App(argc, argv) parses the gflags-style command line, load_graph() loads
// -graphfile (or a -gen spec), run() performs one BFS from -src and prints
// the reference's per-run banner, gpu_init() is the app's pure-virtual
// device initialisation and check() its CPU verification.  The work is
// done by scaph::Engine; only 32-bit vertex/index/value types exist, as in
// the reference's only instantiation (app/bfs.cu:121-123).
#ifndef SCAPH_HOST_APP_H
#define SCAPH_HOST_APP_H

#include <cstdio>
#include <iostream>
#include <string>
#include <type_traits>
#include <vector>

#include "engine.h"
#include "learned.h"

namespace scaph {

extern double Total_throughput;  // src/app.cuh:49

template <typename vertex_t, typename index_t, typename value_t>
class App {
  static_assert(sizeof(vertex_t) == 4 && sizeof(index_t) == 4 && sizeof(value_t) == 4,
                "the paged layout is 32-bit (src/page.h Node<u32,u32>)");

 public:
  App(int argc, char** argv) {
    std::vector<std::string> args;
    for (int i = 1; i < argc; ++i) args.push_back(argv[i]);
    std::string err;
    if (!parse_config(args, &cfg, &positional, &err)) throw Error(SCAPH_STATUS_INVALID, err);
    engine = new Engine(cfg);
    num_stream = cfg.num_stream;
    num_blks = cfg.num_blks;
    num_thds = cfg.num_thds;
    pagesize = (vertex_t)((uint64_t)cfg.pagesize << 20);
    chunksize = pagesize / 32;
    threshold = cfg.threshold * chunksize;
    gpu_memsize = (int64_t)cfg.max_memory << 20;
  }
  explicit App(const Config& c) : cfg(c) {
    engine = new Engine(cfg);
    pagesize = (vertex_t)((uint64_t)cfg.pagesize << 20);
    chunksize = pagesize / 32;
    threshold = cfg.threshold * chunksize;
  }
  virtual ~App() { delete engine; }
  App(const App&) = delete;
  App& operator=(const App&) = delete;

  // src/app.cuh:267-345
  void load_graph() {
    engine->load_from_config();
    if (cfg.schedule == "learned") {
      // -model names the weights file (scaph_amd/schedule_mlp.bin is the
      // trained one); the decision itself runs on the device
      if (cfg.model.empty())
        throw Error(SCAPH_STATUS_INVALID, "-schedule=learned needs -model=<file>");
      LearnedSchedule ls;
      ls.load_file(cfg.model);
      const auto mean = ls.mean_for(engine->num_vertices(), engine->num_edges());
      engine->set_device_model(ls.params(), mean.data(), ls.scale());
    }
    bind_graph();
    std::printf("|========================Dataset Information========================|\n");
    std::printf("|                       graph name      : %-26s|\n",
                cfg.graphfile.empty() ? cfg.gen.c_str() : cfg.graphfile.c_str());
    std::printf("|                       graph nodes     : %-26u|\n", (unsigned)vert_count);
    std::printf("|                       graph edges     : %-26llu|\n",
                (unsigned long long)engine->num_edges());
  }

  // src/app.cuh:220-265: one run of the app, timed, with the reference banner.
  void run() {
    gpu_init();
    last = execute();
    const double secs = last.ms / 1e3;
    const double gteps_ref = secs > 0 ? (double)engine->num_edges() / secs / 1e9 : 0.0;
    const double gteps = secs > 0 ? (double)last.m_reached / secs / 1e9 : 0.0;
    std::printf("|                        %4s costs     : %.8f seconds        |\n", name(), secs);
    std::printf("|                         Throughput    : %.8f GTEPS          |\n", gteps_ref);
    std::printf("|                 Traversed-edge rate   : %.8f GTEPS          |\n", gteps);
    Total_throughput += gteps_ref;
    if (cfg.check) {
      h_value.resize(vert_count);
      engine->levels_to_host(h_value.data());
    }
  }

  // Takes the loaded graph's sizes and device state (load_graph() calls it;
  // call it after loading through engine-> directly).
  void bind_graph() {
    vert_count = engine->num_vertices();
    edge_count = (index_t)engine->num_edges();
    list_size = (int)engine->pages().size();
    pool_size = engine->pool_size();
    pagesize = (vertex_t)engine->page_words();
    const Engine::DeviceState d = engine->device_state();
    d_value = reinterpret_cast<value_t*>(d.value);
    d_active = d.active;
    active_words = d.active_words;
    d_degree = d.degree;
    stream = d.stream;
    copystream = d.copy_stream;
  }

  void iter_init() {}  // per-level bookkeeping lives inside Engine::bfs
  virtual void gpu_init() = 0;
  virtual bool check() { return true; }
  // The app's traversal on the engine (the reference's cond/update functors
  // select it; here each app names its engine entry point).
  virtual RunStats execute() { return engine->bfs(source); }
  virtual const char* name() const { return "BFS"; }

  Config cfg;
  std::vector<std::string> positional;
  Engine* engine = nullptr;
  RunStats last;
  uint32_t source = 0;
  int64_t gpu_memsize = 0;
  vertex_t vert_count = 0;
  index_t edge_count = 0;
  vertex_t pagesize = 0, chunksize = 0, threshold = 0;
  int list_size = 0, pool_size = 0;
  int num_blks = 256, num_thds = 256, num_stream = 5;
  std::vector<value_t> h_value;
  // Device state for the app's gpu_init (the reference's public d_value,
  // d_stat and streams, src/app.cuh:72-135), set by load_graph(): the
  // per-vertex values, the active bitmap (bit v of word v/32 for the
  // reference's bool d_stat[v]; active_words u32), the out-degrees, and the
  // engine's exec and page-copy streams (opaque hipStream_t; the copy
  // stream is null when every page is resident).  An app initialises them
  // through the kernel C-ABI (scaph_hip.h) or kernels of its own, on
  // `stream`, before execute() runs.
  value_t* d_value = nullptr;
  uint32_t* d_active = nullptr;
  size_t active_words = 0;
  const uint32_t* d_degree = nullptr;
  void* stream = nullptr;
  void* copystream = nullptr;
};

// A plugin app on the engine's label-correcting rounds, for apps beyond the
// three built in: gpu_init() writes the start values and the active set
// into d_value / d_active; execute() runs rounds in which every out-edge of
// an active vertex offers its value (relax_mode() SCAPH_RELAX_WCC) or its
// value plus the edge weight (SCAPH_RELAX_SSSP) to the target, until no
// value is lowered — the role of the reference's cond/update/active
// functors (app/wcc_kernel.cuh, app/sssp_kernel.cuh) for min-relaxing apps.
template <typename vertex_t, typename index_t, typename value_t>
class relax_app : public App<vertex_t, index_t, value_t> {
 public:
  using Base = App<vertex_t, index_t, value_t>;
  relax_app(int argc, char** argv) : Base(argc, argv) {}
  explicit relax_app(const Config& c) : Base(c) {}
  virtual int relax_mode() const = 0;
  RunStats execute() override { return this->engine->relax_prepared(relax_mode()); }
};

// The BFS plugin (app/bfs.cu:39-117).
template <typename vertex_t, typename index_t, typename value_t>
class bfs : public App<vertex_t, index_t, value_t> {
 public:
  using Base = App<vertex_t, index_t, value_t>;
  bfs(int argc, char** argv) : Base(argc, argv) {}
  explicit bfs(const Config& c) : Base(c) {}

  // bfs_init + setsrc (app/bfs.cu:20-37, 58-65) run inside Engine::bfs;
  // here the source is chosen and checked.
  void gpu_init() override {
    if (this->cfg.src < 0 || (uint64_t)this->cfg.src >= this->engine->num_vertices())
      throw Error(SCAPH_STATUS_INVALID, "-src is out of range");
    this->source = (uint32_t)this->cfg.src;
  }

  // app/bfs.cu:67-117: CPU BFS and an element-wise comparison.  Compared in
  // the reference's capped form (63 = unreached) and in exact form.
  bool check() override {
    const HostCsr* g = this->engine->host_csr();
    if (!g) {
      std::cout << "check skipped: no host copy of the graph" << std::endl;
      return false;
    }
    if (this->h_value.size() != this->vert_count) {
      this->h_value.resize(this->vert_count);
      this->engine->levels_to_host(this->h_value.data());
    }
    std::vector<uint32_t> cpu;
    cpu_bfs(*g, this->source, 63u, &cpu);
    uint64_t errors = 0;
    for (vertex_t vid = 0; vid < this->vert_count; ++vid) {
      const uint32_t gpu = compat63(this->h_value[vid]);
      if (gpu != cpu[vid]) {
        if (errors < 10)
          std::cout << "vid: " << vid << " gpu: " << gpu << " cpu: " << cpu[vid] << std::endl;
        ++errors;
      }
    }
    std::vector<uint32_t> exact;
    cpu_bfs(*g, this->source, SCAPH_UNREACHED, &exact);
    for (vertex_t vid = 0; vid < this->vert_count; ++vid)
      if (exact[vid] != this->h_value[vid]) ++errors;
    std::cout << "total errors: " << errors << std::endl;
    return errors == 0;
  }
};

// Shared by the WCC and SSSP checks: GPU values against a CPU result.
inline bool compare_values(const std::vector<uint32_t>& gpu, const std::vector<uint32_t>& cpu) {
  uint64_t errors = 0;
  for (size_t vid = 0; vid < cpu.size(); ++vid) {
    if (gpu[vid] != cpu[vid]) {
      if (errors < 10)
        std::cout << "vid: " << vid << " gpu: " << gpu[vid] << " cpu: " << cpu[vid] << std::endl;
      ++errors;
    }
  }
  std::cout << "total error count: " << errors << std::endl;
  return errors == 0;
}

// The WCC plugin (app/wcc.cu:25-97): value[v] = least vertex id reaching v,
// every vertex active at the start.
template <typename vertex_t, typename index_t, typename value_t>
class wcc : public App<vertex_t, index_t, value_t> {
 public:
  using Base = App<vertex_t, index_t, value_t>;
  wcc(int argc, char** argv) : Base(argc, argv) {}
  explicit wcc(const Config& c) : Base(c) {}
  void gpu_init() override {}  // value[v] = v and all-active are set by Engine::wcc
  RunStats execute() override { return this->engine->wcc(); }
  const char* name() const override { return "WCC"; }
  bool check() override {
    const HostCsr* g = this->engine->host_csr();
    if (!g) {
      std::cout << "check skipped: no host copy of the graph" << std::endl;
      return false;
    }
    std::vector<uint32_t> gpu(this->vert_count), cpu;
    this->engine->levels_to_host(gpu.data());
    cpu_wcc(*g, &cpu);
    return compare_values(gpu, cpu);
  }
};

// The SSSP plugin (app/sssp.cu:33-61): shortest paths from -src over the
// _csr.wgh weights (or a generator's wmax=), 0x7FFFFFFF for unreached.
template <typename vertex_t, typename index_t, typename value_t>
class sssp : public App<vertex_t, index_t, value_t> {
 public:
  using Base = App<vertex_t, index_t, value_t>;
  sssp(int argc, char** argv) : Base(argc, argv) {}
  explicit sssp(const Config& c) : Base(c) {}
  void gpu_init() override {
    if (this->cfg.src < 0 || (uint64_t)this->cfg.src >= this->engine->num_vertices())
      throw Error(SCAPH_STATUS_INVALID, "-src is out of range");
    this->source = (uint32_t)this->cfg.src;
  }
  RunStats execute() override { return this->engine->sssp(this->source); }
  const char* name() const override { return "SSSP"; }
  // The reference's check is a stub returning true; this one runs Dijkstra.
  bool check() override {
    const HostCsr* g = this->engine->host_csr();
    if (!g || g->wgh.size() != g->ngh.size()) {
      std::cout << "check skipped: no host copy of the weighted graph" << std::endl;
      return false;
    }
    std::vector<uint32_t> gpu(this->vert_count), cpu;
    this->engine->levels_to_host(gpu.data());
    cpu_sssp(*g, this->source, &cpu);
    return compare_values(gpu, cpu);
  }
};

}  // namespace scaph

#endif
