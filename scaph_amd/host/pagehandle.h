// pagehandle.h — device page-slot cache for the out-of-core (paged) mode.
//
// Same policy as the reference's PageHandle (src/pagehandle.h:252-474):
//   * a slot holds one whole page ("bound") or up to 32 chunks of P/32
//     words, each run of chunks holding one compacted cold page ("shared");
//   * per iteration (queue_reset, :326-379): active pages already bound to
//     a slot run first (rank 0); inactive bound slots and shared slots are
//     free for reuse; uncached pages with more active data than the
//     threshold are copied whole, the rest are compacted and copied as
//     chunk runs (compaction workers append them as they finish);
//   * a page copy takes, in order: room in a shared slot (best fit), an
//     empty slot, or the slot of a page waiting for re-execution (:423-473);
//   * an executed page re-enters the queue at rank+1 while rank < multi-1
//     (:381-404), otherwise its slot returns to the empty list.
// The queues are plain C++ (the reference's MFinFout used tbb::mutex).
#ifndef SCAPH_HOST_PAGEHANDLE_H
#define SCAPH_HOST_PAGEHANDLE_H

#include <cstdint>
#include <deque>
#include <map>
#include <mutex>
#include <set>
#include <vector>

#include "scaph_hip.h"

namespace scaph {

// 32-chunk occupancy masks (src/tools.h:10-85).
int max_zero_run(uint32_t mask);
// Start of the best-fitting zero run of length >= required: the first run of
// exactly `required`, else the last of the shortest longer runs; -1 if none.
int fit_zero_run(uint32_t mask, int required);
uint32_t set_run(uint32_t mask, int off, int number);
uint32_t clear_run(uint32_t mask, int off, int number);

struct VPstat {
  bool cached = false;   // resident in a slot
  bool shared = false;   // this iteration's copy is a compacted chunk run
  int vpid = -1;
  int dpid = -1;
  int dpoff = 0;         // first chunk inside the slot
  uint32_t nodenum = 0;  // Node records to execute
  uint64_t datanum = 0;  // words to copy (compacted) / page words (full)
  int chunknum = 32;
};

struct DPstat {
  bool shared = false;
  uint32_t mark = 0;  // occupied chunks
  int vpid = -1;      // bound page (whole-page slots)
  int dpid = -1;
  int maxm = 32;      // longest free chunk run
};

class PageHandle {
 public:
  void init(int list_size, int pool_size, const std::vector<scaph_page>& pages, uint64_t thres,
            int maxentry);
  bool finished() const;
  // Start of an iteration: datas[p] = active data words of page p.  With
  // group_cached, the cached active pages wait in group_queue() for grouped
  // launches instead of joining the rank-0 queue.
  void queue_reset(const uint32_t* datas, bool group_cached = false);
  std::deque<int>& group_queue() { return groupq_; }
  // Moves the front of group_queue() into execution at rank 0.
  int take_group();
  void release_exec_finished(int vpid);
  void release_copy_finished(int vpid);
  bool get_execute(int* vpid);
  bool get_empty(int* srcv);
  // compaction workers (any thread): a cold page is ready to be copied
  void merge_write(int vpid);
  size_t merge_pending() const;

  // Page-activity prediction for the level about to start (call before
  // queue_reset; an empty vector restores the plain policy).  value[p]
  // ranks how soon page p is expected to be active after this level:
  // higher = sooner, 0 = never again.  With values, free slots are reused least valuable
  // first (a page expected never again goes before one expected next), and
  // pages with value >= lookahead_min are worth copying ahead of the next
  // level.  keep_whole pages (expected active next level with more data
  // than the threshold) are copied whole instead of compacted, so they stay
  // resident, while the pool has room for them.
  void set_prediction(std::vector<double> value, double lookahead_min,
                      std::vector<char> keep_whole = {});
  bool predicting() const { return !value_.empty(); }
  // A whole-page copy ahead of the next level: the most valuable page that
  // is not resident, not active in this level and valued >= lookahead_min,
  // into the least valuable free slot if no later level needs that slot's
  // page (value 0, or an empty or compacted-chunk slot).
  // Call only once every copy of this level has been issued.
  bool get_lookahead(const uint32_t* datas, int* vpid);
  // Which active pages of a level with per-page data `datas` queue_reset
  // copies whole (the others not resident are compacted).
  std::vector<char> whole_copies(const uint32_t* datas) const;
  void release_lookahead_finished(int vpid);
  // Pages of this level still waiting for their launch.
  bool launches_queued() const;

  std::vector<VPstat> vp;
  std::vector<DPstat> dp;
  // statistics of the current iteration
  int num_cached = 0, num_full = 0;

 private:
  bool merge_query(int* v);
  bool merge_read(int* v);
  bool shared_find(int required, int* dpid) const;
  bool exec_read_kick(int* vpid);
  double slot_value(int dpid) const;
  void push_empty(int dpid);
  void unbound(int dpid);
  void bound(int dpid, int vpid);
  void make_shared(int dpid);
  bool insert_shared(int dpid, int vpid);
  void remove_shared(int dpid, int vpid);

  uint64_t thres_ = 0;
  int maxentry_ = 1;
  std::vector<std::deque<int>> execq_;  // by rank
  std::deque<int> groupq_;              // cached active pages for grouped launches
  std::deque<int> emptyq_;
  std::set<int> sharedq_;
  std::set<int> expectq_;
  std::map<int, int> hold_;  // executing page -> rank
  mutable std::mutex merge_mu_;
  std::vector<int> mergeq_;
  size_t merge_reader_ = 0;
  std::vector<uint32_t> page_nodes_;
  uint64_t page_words_full_ = 0;
  std::vector<double> value_;
  std::vector<char> keep_whole_;
  double lookahead_min_ = 0;
};

}  // namespace scaph

#endif
