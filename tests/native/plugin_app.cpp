// plugin_app.cpp — TEST ONLY (built into tests/native/libplugin_app.so, never
// into the product libraries): a third-party app written against the App<>
// plugin surface (scaph_amd/host/app.h), the way a reference app subclasses
// App and initialises d_value / d_stat in its gpu_init (app/bfs.cu:39-65).
//
// "seed labels": label[v] = the least index i of the seeds s_0..s_{k-1} from
// which v is reachable (0xFFFFFFFF when none).  gpu_init writes label[s_i] =
// i (the least i for a repeated seed) and marks every seed active through
// the kernel C-ABI; WCC-mode rounds then carry the minimum along out-edges.
#include <algorithm>
#include <cstring>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "app.h"
#include "scaph_hip.h"

namespace {

class seed_labels : public scaph::relax_app<uint32_t, uint32_t, uint32_t> {
 public:
  seed_labels(const scaph::Config& c, std::vector<uint32_t> seeds)
      : relax_app(c), seeds_(std::move(seeds)) {}
  ~seed_labels() override {
    if (d_seeds_) scaph_dev_free(d_seeds_);
  }
  int relax_mode() const override { return SCAPH_RELAX_WCC; }
  const char* name() const override { return "SEED"; }

  void gpu_init() override {
    std::map<uint32_t, uint32_t> first;  // seed vertex -> least index
    for (uint32_t i = 0; i < seeds_.size(); ++i) {
      if (seeds_[i] >= vert_count) throw scaph::Error(SCAPH_STATUS_INVALID, "seed out of range");
      first.emplace(seeds_[i], i);
    }
    H_ERR(scaph_fill_u32(d_value, 0xFFFFFFFFu, vert_count, stream));
    H_ERR(scaph_fill_u32(d_active, 0u, active_words, stream));
    label_.assign(first.begin(), first.end());
    for (const auto& kv : label_)
      H_ERR(scaph_copy_h2d_async(d_value + kv.first, &kv.second, 4, stream));
    if (!d_seeds_) H_ERR(scaph_dev_alloc(&d_seeds_, std::max<size_t>(seeds_.size(), 1) * 4));
    H_ERR(scaph_copy_h2d_async(d_seeds_, seeds_.data(), seeds_.size() * 4, stream));
    H_ERR(scaph_frontier_queue_to_bitmap(static_cast<uint32_t*>(d_seeds_),
                                         (uint32_t)seeds_.size(), d_active, vert_count, stream));
    H_ERR(scaph_stream_sync(stream));  // the host staging above must outlive the copies
  }

 private:
  std::vector<uint32_t> seeds_;
  std::vector<std::pair<uint32_t, uint32_t>> label_;
  void* d_seeds_ = nullptr;
};

}  // namespace

// flags: space-separated engine flags (gflags syntax).  Returns 0, or the
// error status with its message in err.
extern "C" int plugin_seed_labels(uint32_t n, const uint64_t* offsets, const uint32_t* ngh,
                                  const uint32_t* seeds, uint32_t k, const char* flags,
                                  uint32_t* labels, uint64_t* rounds, char* err, size_t errlen) {
  try {
    std::vector<std::string> args;
    std::istringstream in(flags ? flags : "");
    for (std::string a; in >> a;) args.push_back(a);
    scaph::Config cfg;
    std::string msg;
    if (!scaph::parse_config(args, &cfg, nullptr, &msg)) throw scaph::Error(SCAPH_STATUS_INVALID, msg);
    seed_labels app(cfg, std::vector<uint32_t>(seeds, seeds + k));
    scaph::HostCsr g;
    g.n = n;
    g.offsets.assign(offsets, offsets + n + 1);
    g.ngh.assign(ngh, ngh + offsets[n]);
    app.engine->load_csr(std::move(g), false);
    app.bind_graph();
    app.run();
    app.engine->levels_to_host(labels);
    if (rounds) *rounds = app.last.depth;
    return 0;
  } catch (const scaph::Error& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return e.status ? e.status : 1;
  } catch (const std::exception& e) {
    if (err && errlen) std::snprintf(err, errlen, "%s", e.what());
    return 1;
  }
}
