// pagehandle.cpp — slot cache policy (see pagehandle.h).
#include "pagehandle.h"

#include <algorithm>

namespace scaph {

int max_zero_run(uint32_t mask) {
  int best = 0, run = 0;
  for (int i = 0; i < 32; ++i) {
    if ((mask >> i) & 1u) {
      best = std::max(best, run);
      run = 0;
    } else {
      ++run;
    }
  }
  return std::max(best, run);
}

int fit_zero_run(uint32_t mask, int required) {
  if (required <= 0 || max_zero_run(mask) < required) return -1;
  int exact = -1, best_len = 33, best_off = -1;
  int i = 0;
  while (i < 32) {
    if ((mask >> i) & 1u) {
      ++i;
      continue;
    }
    int j = i;
    while (j < 32 && !((mask >> j) & 1u)) ++j;
    const int len = j - i;
    if (len == required) {
      exact = i;
      break;
    }
    if (len > required && len <= best_len) {  // ties: the later run wins
      best_len = len;
      best_off = i;
    }
    i = j;
  }
  return exact >= 0 ? exact : best_off;
}

uint32_t set_run(uint32_t mask, int off, int number) {
  for (int i = 0; i < number && off + i < 32; ++i) mask |= 1u << (off + i);
  return mask;
}

uint32_t clear_run(uint32_t mask, int off, int number) {
  for (int i = 0; i < number && off + i < 32; ++i) mask &= ~(1u << (off + i));
  return mask;
}

void PageHandle::init(int list_size, int pool_size, const std::vector<scaph_page>& pages,
                      uint64_t thres, int maxentry) {
  thres_ = thres;
  maxentry_ = std::max(1, maxentry);
  execq_.assign(maxentry_, {});
  vp.assign(list_size, VPstat());
  dp.assign(pool_size, DPstat());
  page_nodes_.resize(list_size);
  for (int i = 0; i < list_size; ++i) {
    VPstat& v = vp[i];
    v.cached = i < pool_size;
    v.shared = false;
    v.vpid = i;
    v.dpid = i < pool_size ? i : -1;
    v.dpoff = 0;
    v.nodenum = pages[i].nodenum;
    v.chunknum = 32;
    page_nodes_[i] = pages[i].nodenum;
  }
  for (int d = 0; d < pool_size; ++d) {
    dp[d].shared = false;
    dp[d].dpid = d;
    dp[d].vpid = d < list_size ? d : -1;
    dp[d].mark = dp[d].vpid >= 0 ? 0xFFFFFFFFu : 0u;
    dp[d].maxm = dp[d].vpid >= 0 ? 0 : 32;
  }
  emptyq_.clear();
  sharedq_.clear();
  expectq_.clear();
  hold_.clear();
  groupq_.clear();
  mergeq_.clear();
  merge_reader_ = 0;
}

bool PageHandle::finished() const {
  if (!expectq_.empty() || !hold_.empty() || !groupq_.empty()) return false;
  for (const auto& q : execq_)
    if (!q.empty()) return false;
  return true;
}

int PageHandle::take_group() {
  const int p = groupq_.front();
  groupq_.pop_front();
  hold_[p] = 0;
  return p;
}

void PageHandle::queue_reset(const uint32_t* datas, bool group_cached) {
  emptyq_.clear();
  sharedq_.clear();
  expectq_.clear();
  {
    std::lock_guard<std::mutex> g(merge_mu_);
    mergeq_.clear();
    merge_reader_ = 0;
  }
  for (auto& q : execq_) q.clear();
  groupq_.clear();
  hold_.clear();
  num_cached = 0;
  for (size_t d = 0; d < dp.size(); ++d) {
    if (dp[d].shared) {
      emptyq_.push_back((int)d);
    } else {
      const int pid = dp[d].vpid;
      if (pid != -1 && datas[pid] > 0) {
        (group_cached ? groupq_ : execq_[0]).push_back(pid);
        ++num_cached;
      } else {
        emptyq_.push_back((int)d);
      }
    }
  }
  if (!value_.empty())
    std::stable_sort(emptyq_.begin(), emptyq_.end(),
                     [&](int a, int b) { return slot_value(a) < slot_value(b); });
  num_full = 0;
  const std::vector<char> whole = whole_copies(datas);
  for (size_t i = 0; i < vp.size(); ++i) {
    if (whole[i]) {
      vp[i].shared = false;
      vp[i].nodenum = page_nodes_[i];
      vp[i].chunknum = 32;
      merge_write((int)i);
      ++num_full;
    }
  }
  for (size_t i = 0; i < vp.size(); ++i)
    if (datas[i] > 0) expectq_.insert((int)i);
}

void PageHandle::merge_write(int v) {
  std::lock_guard<std::mutex> g(merge_mu_);
  mergeq_.push_back(v);
}

size_t PageHandle::merge_pending() const {
  std::lock_guard<std::mutex> g(merge_mu_);
  return mergeq_.size() - merge_reader_;
}

bool PageHandle::merge_query(int* v) {
  std::lock_guard<std::mutex> g(merge_mu_);
  if (merge_reader_ >= mergeq_.size()) return false;
  *v = mergeq_[merge_reader_];
  return true;
}

bool PageHandle::merge_read(int* v) {
  std::lock_guard<std::mutex> g(merge_mu_);
  if (merge_reader_ >= mergeq_.size()) return false;
  *v = mergeq_[merge_reader_++];
  return true;
}

void PageHandle::release_exec_finished(int vpid) {
  VPstat& v = vp[vpid];
  if (!v.shared) {
    auto it = hold_.find(vpid);
    const int rank = it == hold_.end() ? maxentry_ - 1 : it->second;
    if (rank < maxentry_ - 1) execq_[rank + 1].push_back(vpid);
    else push_empty(v.dpid);
  } else {
    const int d = v.dpid;
    remove_shared(d, vpid);
    if (!dp[d].shared) {
      sharedq_.erase(d);
      push_empty(d);
    }
  }
  hold_.erase(vpid);
  expectq_.erase(vpid);
}

void PageHandle::release_copy_finished(int vpid) {
  vp[vpid].cached = true;
  execq_[0].push_back(vpid);
}

// Pages of this level copied whole: every uncached active page with more
// data than the threshold (src/pagehandle.h:326-379), plus, with a
// prediction, the keep_whole pages (expected active next level) while the
// pool has slots beyond the cached active pages and those copies; the rest
// of the uncached active pages are compacted.
std::vector<char> PageHandle::whole_copies(const uint32_t* datas) const {
  std::vector<char> w(vp.size(), 0);
  int room = (int)dp.size();
  for (size_t i = 0; i < vp.size(); ++i) {
    if (datas[i] == 0) continue;
    if (vp[i].cached && !vp[i].shared) --room;
    else if (datas[i] > thres_) {
      w[i] = 1;
      --room;
    }
  }
  if (keep_whole_.empty() || room <= 0) return w;
  std::vector<int> cand;
  for (size_t i = 0; i < vp.size(); ++i)
    if (keep_whole_[i] && datas[i] > 0 && !w[i] && !(vp[i].cached && !vp[i].shared))
      cand.push_back((int)i);
  std::stable_sort(cand.begin(), cand.end(), [&](int a, int b) { return datas[a] > datas[b]; });
  for (int i : cand) {
    if (room <= 0) break;
    w[i] = 1;
    --room;
  }
  return w;
}

void PageHandle::set_prediction(std::vector<double> value, double lookahead_min,
                                std::vector<char> keep_whole) {
  value_ = std::move(value);
  lookahead_min_ = lookahead_min;
  keep_whole_ = std::move(keep_whole);
}

// A free slot's worth: that of the page still bound to it; nothing for an
// empty slot or one holding compacted chunks (they are not kept).
double PageHandle::slot_value(int d) const {
  if (dp[d].shared || dp[d].vpid < 0) return -1e300;
  return value_[dp[d].vpid];
}

void PageHandle::push_empty(int d) {
  if (value_.empty()) {
    emptyq_.push_back(d);
    return;
  }
  const double v = slot_value(d);
  auto it = emptyq_.begin();
  while (it != emptyq_.end() && slot_value(*it) <= v) ++it;
  emptyq_.insert(it, d);
}

bool PageHandle::get_lookahead(const uint32_t* datas, int* vpid) {
  if (value_.empty() || emptyq_.empty() || merge_pending() > 0) return false;
  int best = -1;
  for (size_t p = 0; p < vp.size(); ++p) {
    if (vp[p].dpid >= 0 || datas[p] > 0 || value_[p] < lookahead_min_) continue;
    if (best < 0 || value_[p] > value_[best]) best = (int)p;
  }
  if (best < 0) return false;
  // only a slot no later level needs (empty, compacted chunks, or a page
  // valued 0): evicting a page needed later would trade this copy for
  // another one
  const int d = emptyq_.front();
  if (!(slot_value(d) <= 0.0)) return false;
  emptyq_.pop_front();
  unbound(d);
  bound(d, best);
  vp[best].shared = false;
  vp[best].nodenum = page_nodes_[best];
  vp[best].chunknum = 32;
  *vpid = best;
  return true;
}

bool PageHandle::launches_queued() const {
  if (!groupq_.empty()) return true;
  for (const auto& q : execq_)
    if (!q.empty()) return true;
  return false;
}

// The page stays bound to its slot; the next level finds it resident.
void PageHandle::release_lookahead_finished(int vpid) { vp[vpid].cached = true; }

bool PageHandle::get_execute(int* vpid) {
  for (int r = 0; r < maxentry_; ++r) {
    if (!execq_[r].empty()) {
      *vpid = execq_[r].front();
      execq_[r].pop_front();
      hold_[*vpid] = r;
      return true;
    }
  }
  return false;
}

bool PageHandle::shared_find(int required, int* dpid) const {
  for (int d : sharedq_)
    if (dp[d].maxm >= required) {
      *dpid = d;
      return true;
    }
  return false;
}

bool PageHandle::exec_read_kick(int* vpid) {
  for (int r = maxentry_ - 1; r > 0; --r) {
    if (!execq_[r].empty()) {
      *vpid = execq_[r].back();
      execq_[r].pop_back();
      return true;
    }
  }
  return false;
}

void PageHandle::unbound(int d) {
  DPstat& s = dp[d];
  if (s.vpid != -1) {
    VPstat& v = vp[s.vpid];
    v.cached = false;
    v.dpid = -1;
    v.dpoff = 0;
  }
  s.shared = false;
  s.vpid = -1;
  s.maxm = 32;
  s.mark = 0;
}

void PageHandle::bound(int d, int pid) {
  vp[pid].dpid = d;
  vp[pid].dpoff = 0;
  dp[d].vpid = pid;
  dp[d].maxm = 0;
  dp[d].mark = 0xFFFFFFFFu;
}

void PageHandle::make_shared(int d) { dp[d].shared = true; }

bool PageHandle::insert_shared(int d, int pid) {
  const int off = fit_zero_run(dp[d].mark, vp[pid].chunknum);
  if (off < 0) return false;
  dp[d].mark = set_run(dp[d].mark, off, vp[pid].chunknum);
  dp[d].maxm = max_zero_run(dp[d].mark);
  vp[pid].dpid = d;
  vp[pid].dpoff = off;
  return true;
}

void PageHandle::remove_shared(int d, int pid) {
  dp[d].mark = clear_run(dp[d].mark, vp[pid].dpoff, vp[pid].chunknum);
  dp[d].maxm = max_zero_run(dp[d].mark);
  vp[pid].dpid = -1;
  vp[pid].dpoff = 0;
  vp[pid].cached = false;
  if (dp[d].maxm == 32) {
    dp[d].shared = false;
    dp[d].vpid = -1;
    dp[d].mark = 0;
  }
}

bool PageHandle::get_empty(int* srcv) {
  int s;
  if (!merge_query(&s)) return false;
  int d, kick;
  if (vp[s].shared) {
    if (shared_find(vp[s].chunknum, &d)) {
      merge_read(&s);
      insert_shared(d, s);
    } else if (!emptyq_.empty()) {
      d = emptyq_.front();
      emptyq_.pop_front();
      merge_read(&s);
      unbound(d);
      make_shared(d);
      insert_shared(d, s);
      sharedq_.insert(d);
    } else if (exec_read_kick(&kick)) {
      merge_read(&s);
      d = vp[kick].dpid;
      unbound(d);
      make_shared(d);
      insert_shared(d, s);
      sharedq_.insert(d);
    } else {
      return false;
    }
  } else {
    if (!emptyq_.empty()) {
      d = emptyq_.front();
      emptyq_.pop_front();
      merge_read(&s);
      unbound(d);
      bound(d, s);
    } else if (exec_read_kick(&kick)) {
      merge_read(&s);
      d = vp[kick].dpid;
      unbound(d);
      bound(d, s);
    } else {
      return false;
    }
  }
  *srcv = s;
  return true;
}

}  // namespace scaph
