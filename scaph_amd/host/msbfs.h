// msbfs.h — multi-source bitset BFS over an engine's resident graph: up to
// 256 queries per traversal (scaph_ms_* kernels), a host level loop or
// hipGraph-captured level segments, and the autotuner that picks the
// push/pull thresholds, the pull block size and the launch mode.  No
// reference counterpart.
#ifndef SCAPH_HOST_MSBFS_H
#define SCAPH_HOST_MSBFS_H

#include <cstdint>
#include <utility>
#include <vector>

#include "engine.h"
#include "hip_abi.h"
#include "scaph_hip.h"

namespace scaph {

struct MsTune {
  uint32_t push_div = 64;     // push while n_active * push_div < n
  uint32_t push_mdiv = 16;    // and m_active * push_mdiv < m
  uint32_t pull_block = 256;  // pull threads per block
  uint32_t use_graph = 0;     // replay captured level segments
  uint32_t segment = 8;       // levels per captured segment
  uint32_t pull_grid = 0;     // cap on pull blocks (0: one per block of vertices)
  uint32_t persistent = 0;    // all levels in one cooperative launch (scaph_ms_run_persistent)
  bool operator==(const MsTune& o) const {
    return push_div == o.push_div && push_mdiv == o.push_mdiv && pull_block == o.pull_block &&
           use_graph == o.use_graph && segment == o.segment && pull_grid == o.pull_grid &&
           persistent == o.persistent;
  }
};

struct MsStats {
  double ms = 0;           // init + all levels, device time
  uint32_t depth = 0;      // levels with a non-empty frontier
  uint32_t queries = 0;
  uint64_t pairs = 0;      // (vertex, query) pairs reached, sources included
};

class MultiSource {
 public:
  MultiSource(Engine& e, uint32_t max_levels);
  ~MultiSource();
  MultiSource(const MultiSource&) = delete;
  MultiSource& operator=(const MultiSource&) = delete;

  MsStats run(const uint32_t* sources, uint32_t count, const MsTune& t);
  // uint32 levels of one query of the last run.
  void levels(uint32_t query, uint32_t* host_out);
  std::vector<scaph_ms_counters> counters();
  // Times every candidate (after one untimed run each) and returns the
  // fastest; `log` receives every candidate's time.
  MsTune autotune(const uint32_t* sources, uint32_t count,
                  std::vector<std::pair<MsTune, double>>* log);
  bool stale() const { return e_.graph_generation() != gen_; }

 private:
  struct Segments {
    uint32_t count = 0;
    MsTune tune;
    std::vector<void*> graphs;
  };
  void launch_levels(uint32_t first, uint32_t last, const scaph_ms_params& p);
  Segments& segments(uint32_t count, const MsTune& t, const scaph_ms_params& p);
  void drop_graphs();

  Engine& e_;
  uint64_t gen_;
  uint32_t max_levels_;
  DevBuf state_, src_, lev_;
  scaph_ms_state st_{};
  PinnedBuf hctr_, hsrc_;
  uint32_t depth_ = 0, count_ = 0;
  std::vector<Segments> segs_;
};

}  // namespace scaph

#endif
