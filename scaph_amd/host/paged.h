// paged.h — state of the out-of-core (paged) BFS mode.
#ifndef SCAPH_HOST_PAGED_H
#define SCAPH_HOST_PAGED_H

#include <memory>
#include <vector>

#include "engine.h"
#include "pagehandle.h"

namespace scaph {

struct Engine::PagedState {
  std::vector<PinnedBuf> h_page;  // page content (Node records + neighbours), used words
  std::vector<unsigned long long> d_hpage;  // the same pages as kernels address them (PCIe)
  std::vector<PinnedBuf> h_tail;  // compacted copy of a cold page (<= thres words)
  DevBuf pool;                    // pool_size slots of P words
  std::vector<std::unique_ptr<Stream>> exec;
  std::unique_ptr<Stream> copy;
  std::vector<std::unique_ptr<Event>> exec_done;
  std::vector<DevBuf> ws;
  size_t ws_bytes = 0;
  std::unique_ptr<Event> copy_beg, copy_end;
  PageHandle handle;
  PinnedBuf h_nodes, h_datas, h_totals, h_front;
  uint64_t thres = 0;  // words (threshold * P/32)
  uint64_t chunk = 0;  // words per chunk (P/32)
  // pages resident at the start of the iteration (page hit accounting)
  std::vector<char> resident_at_reset;
  // grouped launches of the cached active pages (-page_groups)
  std::unique_ptr<Stream> group;
  std::unique_ptr<Event> group_done;
  DevBuf group_ws, d_runs;
  PinnedBuf h_runs;
  size_t group_ws_bytes = 0;
  uint64_t group_nodes_cap = 0;  // active Nodes one group may hold (workspace size)
  // WCC / SSSP: per-exec-stream chunk workspaces, the pages' weights in
  // pinned memory (edgenum words each) and a weight pool with the page
  // pool's slots (a page's weights sit at its slot's edge offsets)
  std::vector<DevBuf> relax_ws;
  std::vector<PinnedBuf> h_wpage;
  std::vector<PinnedBuf> h_wtail;  // compacted weights of a cold page (<= thres words)
  DevBuf wpool;
  // page-activity prediction (-page_predict): per-page count of vertices a
  // later level can still reach, the bitmap of vertices with in-edges, the
  // model's decision record; for the exact bound, per page the levels at
  // which it is active and its data at each level
  DevBuf d_unv, live;
  PinnedBuf h_unv;
  std::vector<uint64_t> level_mask;
  std::vector<std::vector<uint64_t>> level_datas;  // [level][page]
};

}  // namespace scaph

#endif
