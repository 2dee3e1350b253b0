// batch.h — a batch of small independent R-MAT graphs on the device (one
// disjoint-union CSR), their BFS queries in one persistent launch
// (scaph_batch_bfs), optionally replayed from a captured hipGraph, and the
// autotuner that picks block size and workgroups per CU.  No reference
// counterpart.
#ifndef SCAPH_HOST_BATCH_H
#define SCAPH_HOST_BATCH_H

#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

#include "hip_abi.h"
#include "scaph_hip.h"

namespace scaph {

struct BatchTune {
  uint32_t block = 256;
  uint32_t wgs_per_cu = 4;
  uint32_t use_graph = 0;
  bool operator==(const BatchTune& o) const {
    return block == o.block && wgs_per_cu == o.wgs_per_cu && use_graph == o.use_graph;
  }
};

class Batch {
 public:
  // count graphs; graph g: scale uniform in [min_scale, max_scale], edge
  // factor uniform in [min_ef, max_ef], its own seed and a random source,
  // all drawn from SplitMix64 words of `seed` (word(seed, 4g + k)).
  Batch(uint32_t count, uint32_t min_scale, uint32_t max_scale, uint32_t min_ef, uint32_t max_ef,
        uint64_t seed, int device);
  ~Batch();
  Batch(const Batch&) = delete;
  Batch& operator=(const Batch&) = delete;

  uint32_t count() const { return (uint32_t)graphs_.size(); }
  uint64_t vertices() const { return total_v_; }
  uint64_t edges() const { return total_e_; }
  const scaph_batch_entry& graph(uint32_t g) const { return graphs_.at(g); }
  uint64_t graph_edges(uint32_t g);  // edges after dedup

  double run(const BatchTune& t);  // device ms of one batch
  void levels(uint32_t g, uint32_t* host_out);
  void csr(uint32_t g, uint64_t* offsets, uint32_t* ngh, uint64_t cap, uint64_t* count);
  BatchTune autotune(std::vector<std::pair<BatchTune, double>>* log);
  BatchTune tuned() const { return tuned_; }

 private:
  void* graph_for(const BatchTune& t);

  std::vector<scaph_batch_entry> graphs_;
  std::vector<uint32_t> order_;
  uint64_t total_v_ = 0, total_e_ = 0;
  uint32_t max_n_ = 0, cus_ = 0;
  std::unique_ptr<Stream> st_;
  DevBuf d_graphs_, d_order_, offsets_, ngh_, levels_, work_, qbuf_;
  std::vector<uint64_t> h_off_;  // offsets copied back lazily
  BatchTune tuned_;
  std::vector<std::pair<BatchTune, void*>> captured_;
};

}  // namespace scaph

#endif
