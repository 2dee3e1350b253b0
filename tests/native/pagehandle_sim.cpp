// pagehandle_sim.cpp — TEST ONLY (built into tests/native/libpagehandle_sim.so,
// never into the product libraries): drives the engine's PageHandle with a
// simulated executor and copy engine (random completion order) and checks
// the scheduling invariants on every step.
#include <algorithm>
#include <random>
#include <string>
#include <vector>

#include "pagehandle.h"

namespace {

struct SimError {
  int code;
};

void check_slots(const scaph::PageHandle& h) {
  // chunk masks equal the union of the runs of the pages placed there
  std::vector<uint32_t> want(h.dp.size(), 0);
  for (size_t p = 0; p < h.vp.size(); ++p) {
    const scaph::VPstat& v = h.vp[p];
    if (v.dpid < 0) continue;
    if (v.shared) {
      if (v.dpoff < 0 || v.dpoff + v.chunknum > 32) throw SimError{10};
      const uint32_t run =
          (v.chunknum >= 32 ? 0xFFFFFFFFu : ((1u << v.chunknum) - 1u)) << v.dpoff;
      if (want[v.dpid] & run) throw SimError{11};  // overlapping chunk runs
      want[v.dpid] |= run;
    } else if (h.dp[v.dpid].vpid != (int)p) {
      throw SimError{12};  // slot and page disagree on the binding
    }
  }
  for (size_t d = 0; d < h.dp.size(); ++d)
    if (h.dp[d].shared && h.dp[d].mark != want[d]) throw SimError{13};
}

}  // namespace

extern "C" int scaph_host_pagehandle_simulate(uint32_t list_size, uint32_t pool_size,
                                              uint64_t thres_words, int maxentry,
                                              uint32_t iterations, const uint32_t* datas,
                                              uint64_t seed, uint64_t* out_stats,
                                              int predict) {
  if (list_size == 0 || pool_size == 0 || pool_size > list_size || !datas) return 1;
  try {
    std::vector<scaph_page> pages(list_size);
    for (uint32_t i = 0; i < list_size; ++i)
      pages[i] = scaph_page{i * 100, i * 100 + 99, 100, 1000};
    scaph::PageHandle h;
    h.init((int)list_size, (int)pool_size, pages, thres_words, maxentry);
    std::mt19937_64 rng(seed);
    uint64_t execs = 0, copies = 0, compacted = 0, cached_execs = 0, ahead_copies = 0;
    for (uint32_t it = 0; it < iterations; ++it) {
      const uint32_t* d = datas + (uint64_t)it * list_size;
      if (predict) {
        // random next-level values and keep-whole marks
        std::vector<double> value(list_size);
        std::vector<char> keep(list_size);
        for (uint32_t p = 0; p < list_size; ++p) {
          value[p] = (double)(rng() % 4);
          keep[p] = (char)(rng() % 3 == 0);
        }
        h.set_prediction(value, 2.0, keep);
      }
      const std::vector<char> whole = h.whole_copies(d);
      std::vector<int> cold;
      for (uint32_t p = 0; p < list_size; ++p) {
        if (whole[p] && (h.vp[p].cached || d[p] == 0)) throw SimError{25};
        if (!whole[p] && d[p] > thres_words && !h.vp[p].cached) throw SimError{26};
        if (!h.vp[p].cached && d[p] > 0 && !whole[p]) cold.push_back((int)p);
      }
      std::vector<char> resident(list_size);
      for (uint32_t p = 0; p < list_size; ++p) resident[p] = h.vp[p].cached;
      h.queue_reset(d);
      std::vector<int> executed(list_size, 0), copied(list_size, 0);
      size_t cold_next = 0;
      std::vector<int> running;
      int copying = -1;
      bool ahead = false;
      uint64_t guard = 0;
      while (!h.finished()) {
        if (++guard > 1000000) throw SimError{20};  // no progress
        const int action = (int)(rng() % 4);
        if (action == 0 && cold_next < cold.size()) {
          // a compaction worker finishes (datas are in chunk units here)
          const int p = cold[cold_next++];
          h.vp[p].nodenum = 1;
          h.vp[p].datanum = d[p];
          h.vp[p].chunknum = (int)std::min<uint64_t>(32, d[p]);
          h.vp[p].shared = true;
          h.merge_write(p);
          ++compacted;
        } else if (action == 1) {
          if (!running.empty()) {
            const size_t k = rng() % running.size();
            h.release_exec_finished(running[k]);
            running.erase(running.begin() + (long)k);
          }
          int e;
          if (running.size() < 5 && h.get_execute(&e)) {
            if (!h.vp[e].cached && !h.vp[e].shared) throw SimError{21};  // not resident
            if (h.vp[e].dpid < 0) throw SimError{22};
            if (h.vp[e].shared && !copied[e]) throw SimError{23};
            running.push_back(e);
            ++executed[e];
            ++execs;
            if (resident[e]) ++cached_execs;
          }
        } else {
          if (copying != -1) {
            if (ahead) h.release_lookahead_finished(copying);
            else h.release_copy_finished(copying);
            copying = -1;
          }
          int s;
          if (predict && cold_next == cold.size() && h.merge_pending() == 0 &&
              h.get_lookahead(d, &s)) {
            // a copy ahead of the next level: never a page this level runs
            if (d[s] > 0 || h.vp[s].dpid < 0 || h.vp[s].shared) throw SimError{27};
            copying = s;
            ahead = true;
            ++copies;
            ++ahead_copies;
          } else if (h.get_empty(&s)) {
            if (h.vp[s].dpid < 0) throw SimError{24};
            copying = s;
            ahead = false;
            copied[s] = 1;
            ++copies;
          }
        }
        check_slots(h);
      }
      for (uint32_t p = 0; p < list_size; ++p)
        if (d[p] > 0 && executed[p] < 1) throw SimError{30};  // active page skipped
      if (copying != -1 && !ahead) throw SimError{31};
      if (copying != -1) h.release_lookahead_finished(copying);
      // every resident whole page is bound to its slot
      for (uint32_t p = 0; p < list_size; ++p)
        if (h.vp[p].cached && !h.vp[p].shared &&
            (h.vp[p].dpid < 0 || h.dp[h.vp[p].dpid].vpid != (int)p))
          throw SimError{32};
    }
    if (out_stats) {
      out_stats[0] = execs;
      out_stats[1] = copies;
      out_stats[2] = compacted;
      out_stats[3] = cached_execs;
      out_stats[4] = ahead_copies;
    }
    return 0;
  } catch (const SimError& e) {
    return e.code;
  }
}
