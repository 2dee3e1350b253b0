// relax.hip — label-correcting push rounds for the reference's other App
// plugins on the same paged graph:
//   WCC  (app/wcc.cu:14-24, app/wcc_kernel.cuh): value[v] = v, every vertex
//        active, an out-edge (u, v) offers value[u] to v;
//   SSSP (app/sssp.cu:13-31, app/sssp_kernel.cuh): value = 0x7fffffff,
//        value[src] = 0, an out-edge (u, v, w) offers value[u] + w.
// A round takes the active vertices as a list.  One wave takes a vertex and
// its lanes walk the out-list; a value lowered by atomicMin marks the target
// active for the next round.  Lists longer than kRelaxChunk continue as
// chunks in a second, grid-wide pass.  The fixpoint — the minimum over all
// paths — does not depend on the order of the updates, so the values equal
// the reference's asynchronous push for every schedule.
#include "common.h"

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kRelaxChunk = 4096;  // out-edges per wave item; longer lists are chunked
constexpr uint32_t kRelaxSerial = 16;   // lists this short are walked by one lane

struct RelaxChunk {
  uint32_t u;
  uint32_t len;
  uint64_t ptr;  // address of the chunk's first out-edge
};

struct RelaxWs {
  unsigned long long* count;  // chunks appended this round
  RelaxChunk* chunks;
  unsigned long long cap;
  size_t bytes;
};

RelaxWs carve_relax_ws(void* base, unsigned long long max_edges) {
  RelaxWs w;
  w.cap = max_edges / kRelaxChunk + 64;
  w.bytes = 256 + w.cap * sizeof(RelaxChunk);
  char* p = static_cast<char*>(base);
  w.count = reinterpret_cast<unsigned long long*>(p);
  w.chunks = p ? reinterpret_cast<RelaxChunk*>(p + 256) : nullptr;
  return w;
}

// The lane's share of out-edges e[0..len) of a vertex whose value is `val`;
// the weight of edge e[k] sits `wdelta` bytes after it (weights laid out
// like the page pool), or every weight is 0 when wdelta == 0 and !WEIGHTED.
template <int WEIGHTED>
__device__ __forceinline__ void relax_list(const uint32_t* e, uint32_t len, uint32_t val,
                                           long long wdelta, uint32_t n, uint32_t* value,
                                           uint32_t* front_out, unsigned long long& acc_n,
                                           unsigned long long& acc_s, uint32_t lane,
                                           uint32_t stride) {
  const uint32_t* w = WEIGHTED ? reinterpret_cast<const uint32_t*>(
                                     reinterpret_cast<const char*>(e) + wdelta)
                               : nullptr;
  for (uint32_t k = lane; k < len; k += stride) {
    const uint32_t v = e[k];
    ++acc_s;
    if (v >= n) continue;
    const uint32_t cand = WEIGHTED ? val + w[k] : val;
    if (cand < value[v]) {
      const uint32_t old = atomicMin(&value[v], cand);
      if (cand < old) {
        const uint32_t bit = 1u << (v & 31);
        if (!(atomicOr(&front_out[v >> 5], bit) & bit)) ++acc_n;
      }
    }
  }
}

__device__ __forceinline__ void relax_flush(scaph_level_counters* counters,
                                            unsigned long long acc_n, unsigned long long acc_s) {
  acc_n = wave_sum(acc_n);
  acc_s = wave_sum(acc_s);
  if (lane_id() == 0) {
    if (acc_n) atomicAdd(&counters->n_next, acc_n);
    if (acc_s) atomicAdd(&counters->scanned, acc_s);
  }
}

// One wave relaxes vertex u's out-list: the first kRelaxChunk edges here,
// the rest appended as chunks for k_relax_chunks.
template <int WEIGHTED>
__device__ __forceinline__ void relax_node(uint32_t u, NodeRef nr, long long wdelta, uint32_t n,
                                           uint32_t* value, uint32_t* front_out,
                                           scaph_level_counters* counters, const RelaxWs& ws,
                                           unsigned long long& acc_n, unsigned long long& acc_s,
                                           uint32_t lane) {
  const uint32_t val = value[u];
  const uint32_t first = nr.len < kRelaxChunk ? nr.len : kRelaxChunk;
  relax_list<WEIGHTED>(nr.edges, first, val, wdelta, n, value, front_out, acc_n, acc_s, lane,
                       SCAPH_WAVE);
  if (nr.len > kRelaxChunk && lane == 0) {
    const uint32_t rest = nr.len - kRelaxChunk;
    const uint32_t nch = (rest + kRelaxChunk - 1) / kRelaxChunk;
    const unsigned long long at = atomicAdd(ws.count, (unsigned long long)nch);
    if (at + nch <= ws.cap) {
      for (uint32_t c = 0; c < nch; ++c) {
        RelaxChunk ch;
        ch.u = u;
        ch.len = min(kRelaxChunk, rest - c * kRelaxChunk);
        ch.ptr = reinterpret_cast<uint64_t>(nr.edges + kRelaxChunk + (uint64_t)c * kRelaxChunk);
        ws.chunks[at + c] = ch;
      }
    } else {
      counters->reserved = 1;  // overflow: the host reports an error
    }
  }
}

// A wave takes 64 active vertices, one per lane: a lane walks a list of at
// most kRelaxSerial edges itself; longer lists go one at a time to the
// whole wave (relax_node), so short lists do not leave 63 lanes idle.
template <int WEIGHTED>
__global__ void __launch_bounds__(kBlock)
k_relax(scaph_graph_desc g, const uint32_t* q, uint32_t n_q, long long wdelta, uint32_t* value,
        uint32_t* front_out, scaph_level_counters* counters, RelaxWs ws) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const PageTab t = stage_pages(g, dyn_lds);
  __syncthreads();
  const uint32_t n = g.num_vertices;
  const uint32_t lane = lane_id();
  const uint32_t nw = kBlock / SCAPH_WAVE;
  unsigned long long acc_n = 0, acc_s = 0;
  for (uint32_t base = (blockIdx.x * nw + threadIdx.x / SCAPH_WAVE) * SCAPH_WAVE; base < n_q;
       base += gridDim.x * nw * SCAPH_WAVE) {
    const uint32_t i = base + lane;
    uint32_t u = 0;
    NodeRef nr;
    nr.edges = nullptr;
    nr.len = 0;
    bool act = false;
    if (i < n_q) {
      u = q[i];
      if (u < n) {
        nr = load_node(t, g, u);
        act = true;
      }
    }
    if (act && nr.len <= kRelaxSerial)
      relax_list<WEIGHTED>(nr.edges, nr.len, value[u], wdelta, n, value, front_out, acc_n, acc_s,
                           0u, 1u);
    uint64_t ml = __ballot(act && nr.len > kRelaxSerial);
    while (ml) {
      const int j = __ffsll((unsigned long long)ml) - 1;
      ml &= ml - 1;
      NodeRef nj;
      nj.edges = reinterpret_cast<const uint32_t*>(
          wave_bcast64(reinterpret_cast<uint64_t>(nr.edges), j));
      nj.len = wave_bcast(nr.len, j);
      relax_node<WEIGHTED>(wave_bcast(u, j), nj, wdelta, n, value, front_out, counters, ws, acc_n,
                           acc_s, lane);
    }
  }
  relax_flush(counters, acc_n, acc_s);
}

// One page buffer (Node records {id, edge offset, degree} then the
// out-lists, the layout of tools/bin2pagecsr.cpp) of Nodes
// [first_vertex, first_vertex + nodenum): a wave takes 64 consecutive Nodes,
// finds the active ones from front_in and relaxes each in turn.  A
// compacted buffer (first_vertex == SCAPH_PAGE_COMPACTED, src/compress.h
// layout) holds only active Nodes, each naming its vertex.
template <int WEIGHTED>
__global__ void __launch_bounds__(kBlock)
k_relax_page(const uint32_t* buf, uint32_t nodenum, uint32_t first_vertex, const uint32_t* front_in,
             uint32_t n, long long wdelta, uint32_t* value, uint32_t* front_out,
             scaph_level_counters* counters, RelaxWs ws) {
  const uint32_t lane = lane_id();
  const uint32_t nw = kBlock / SCAPH_WAVE;
  const uint32_t* edges = buf + 3ull * nodenum;
  const bool compacted = first_vertex == SCAPH_PAGE_COMPACTED;
  unsigned long long acc_n = 0, acc_s = 0;
  for (uint32_t i0 = (blockIdx.x * nw + threadIdx.x / SCAPH_WAVE) * SCAPH_WAVE; i0 < nodenum;
       i0 += gridDim.x * nw * SCAPH_WAVE) {
    const uint32_t i = i0 + lane;
    bool act = false;
    if (i < nodenum) {
      const uint32_t v = first_vertex + i;
      act = compacted || ((front_in[v >> 5] >> (v & 31)) & 1u);
    }
    uint64_t mask = __ballot(act);
    while (mask) {
      const uint32_t j = (uint32_t)__builtin_ctzll(mask);
      mask &= mask - 1;
      const uint32_t* node = buf + 3ull * (i0 + j);
      NodeRef nr;
      nr.edges = edges + node[1];
      nr.len = node[2];
      const uint32_t u = compacted ? node[0] : first_vertex + i0 + j;
      if (u >= n) continue;
      relax_node<WEIGHTED>(u, nr, wdelta, n, value, front_out, counters, ws, acc_n, acc_s, lane);
    }
  }
  relax_flush(counters, acc_n, acc_s);
}

// The chunks of long lists: one block per chunk, grid-stride.
template <int WEIGHTED>
__global__ void __launch_bounds__(kBlock)
k_relax_chunks(RelaxWs ws, uint32_t n, long long wdelta, uint32_t* value, uint32_t* front_out,
               scaph_level_counters* counters) {
  const unsigned long long cnt = *ws.count;
  const unsigned long long total = cnt < ws.cap ? cnt : ws.cap;
  unsigned long long acc_n = 0, acc_s = 0;
  for (unsigned long long c = blockIdx.x; c < total; c += gridDim.x) {
    const RelaxChunk ch = ws.chunks[c];
    relax_list<WEIGHTED>(reinterpret_cast<const uint32_t*>(ch.ptr), ch.len, value[ch.u], wdelta,
                         n, value, front_out, acc_n, acc_s, threadIdx.x, kBlock);
  }
  relax_flush(counters, acc_n, acc_s);
}

__global__ void k_relax_zero(unsigned long long* p) { *p = 0; }

// Near/far split of a round's active list (SSSP): vertices whose value is
// below `bound` go to `near` for this round; the others keep their bit in
// front_out (the next round's active set) and their least value is kept
// for the next bound.  split[0] = near count, [1] = far count, [2] = least
// far value (split must start as {0, 0, ~0}).
__global__ void __launch_bounds__(kBlock)
k_relax_split(const uint32_t* q, uint32_t n_q, const uint32_t* value, uint32_t bound,
              uint32_t n, uint32_t* near, uint32_t* front_out, unsigned long long* split) {
  const uint32_t stride = gridDim.x * kBlock;
  unsigned long long far_cnt = 0, least = ~0ull;
  for (uint32_t i0 = blockIdx.x * kBlock; i0 < n_q; i0 += stride) {
    const uint32_t i = i0 + threadIdx.x;
    uint32_t u = 0, val = 0;
    bool valid = i < n_q;
    if (valid) {
      u = q[i];
      valid = u < n;
    }
    if (valid) val = value[u];
    const bool is_near = valid && val < bound;
    const uint64_t m = __ballot(is_near);
    unsigned long long base = 0;
    if (lane_id() == 0 && m) base = atomicAdd(&split[0], (unsigned long long)__popcll(m));
    base = wave_bcast64(base, 0);
    if (is_near) near[base + lanes_below(m)] = u;
    if (valid && !is_near) {
      atomicOr(&front_out[u >> 5], 1u << (u & 31));
      ++far_cnt;
      if (val < least) least = val;
    }
  }
  far_cnt = wave_sum(far_cnt);
#pragma unroll
  for (int d = SCAPH_WAVE / 2; d >= 1; d >>= 1) {
    const unsigned long long o = __shfl_xor(least, d, SCAPH_WAVE);
    least = o < least ? o : least;
  }
  if (lane_id() == 0) {
    if (far_cnt) atomicAdd(&split[1], far_cnt);
    if (least != ~0ull) atomicMin(&split[2], least);
  }
}

__global__ void k_relax_split_init(unsigned long long* split) {
  split[0] = 0;
  split[1] = 0;
  split[2] = ~0ull;
}

__global__ void __launch_bounds__(kBlock)
k_relax_init(uint32_t* value, uint32_t n, uint32_t mode, uint32_t src) {
  const uint32_t stride = gridDim.x * kBlock;
  for (uint32_t v = blockIdx.x * kBlock + threadIdx.x; v < n; v += stride)
    value[v] = mode == SCAPH_RELAX_WCC ? v : (v == src ? 0u : SCAPH_RELAX_INF);
}

inline unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  size_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace

extern "C" {

size_t scaph_relax_workspace_bytes(unsigned long long max_edges) {
  return carve_relax_ws(nullptr, max_edges).bytes;
}

int scaph_relax_init(uint32_t* value, uint32_t n, int mode, uint32_t src, void* stream) {
  if (!value || n == 0 || (mode != SCAPH_RELAX_WCC && mode != SCAPH_RELAX_SSSP) ||
      (mode == SCAPH_RELAX_SSSP && src >= n))
    return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_relax_init, dim3(grid_for(n, kBlock, 16384)), dim3(kBlock), 0,
                     as_stream(stream), value, n, (uint32_t)mode, src);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_relax_round(const scaph_graph_desc* graph, const uint32_t* queue, uint32_t n_queue,
                      const uint32_t* weights, const uint32_t* pool_base, uint32_t* value,
                      uint32_t* front_out, scaph_level_counters* counters, void* workspace,
                      size_t workspace_bytes, void* stream) {
  if (!graph || !value || !front_out || !counters || !workspace || (weights && !pool_base))
    return SCAPH_STATUS_INVALID;
  if (n_queue == 0) return 0;
  if (!queue) return SCAPH_STATUS_INVALID;
  RelaxWs ws = carve_relax_ws(workspace, 0);
  // chunk capacity from the workspace the caller sized with
  // scaph_relax_workspace_bytes(graph edges): every chunk of one round fits
  ws.cap = workspace_bytes > 256 ? (workspace_bytes - 256) / sizeof(RelaxChunk) : 0;
  if (ws.cap == 0) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  const long long wdelta =
      weights ? (long long)(reinterpret_cast<uintptr_t>(weights) -
                            reinterpret_cast<uintptr_t>(pool_base))
              : 0ll;
  hipLaunchKernelGGL(k_relax_zero, dim3(1), dim3(1), 0, st, ws.count);
  SCAPH_LAUNCH_CHECK();
  const unsigned grid = grid_for(n_queue, kBlock, 16384);  // a lane per queue entry
  const size_t lds = page_lds_bytes(graph->num_pages);
  if (weights) {
    hipLaunchKernelGGL(k_relax<1>, dim3(grid), dim3(kBlock), lds, st, *graph, queue, n_queue,
                       wdelta, value, front_out, counters, ws);
    SCAPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_relax_chunks<1>, dim3(2048), dim3(kBlock), 0, st, ws,
                       graph->num_vertices, wdelta, value, front_out, counters);
  } else {
    hipLaunchKernelGGL(k_relax<0>, dim3(grid), dim3(kBlock), lds, st, *graph, queue, n_queue,
                       wdelta, value, front_out, counters, ws);
    SCAPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_relax_chunks<0>, dim3(2048), dim3(kBlock), 0, st, ws,
                       graph->num_vertices, wdelta, value, front_out, counters);
  }
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_relax_split(const uint32_t* queue, uint32_t n_queue, const uint32_t* value,
                      uint32_t bound, uint32_t n, uint32_t* near_queue, uint32_t* front_out,
                      unsigned long long* split, void* stream) {
  if (!value || !near_queue || !front_out || !split || (n_queue && !queue))
    return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_relax_split_init, dim3(1), dim3(1), 0, st, split);
  SCAPH_LAUNCH_CHECK();
  if (n_queue == 0) return 0;
  hipLaunchKernelGGL(k_relax_split, dim3(grid_for(n_queue, kBlock, 4096)), dim3(kBlock), 0, st,
                     queue, n_queue, value, bound, n, near_queue, front_out, split);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_relax_page(const uint32_t* buf, uint32_t nodenum, uint32_t first_vertex,
                     const uint32_t* front_in, uint32_t n, const uint32_t* weights,
                     const uint32_t* pool_base, uint32_t* value, uint32_t* front_out,
                     scaph_level_counters* counters, void* workspace, size_t workspace_bytes,
                     void* stream) {
  if (!buf || !front_in || !value || !front_out || !counters || !workspace ||
      (weights && !pool_base) ||
      (first_vertex != SCAPH_PAGE_COMPACTED && (uint64_t)first_vertex + nodenum > n))
    return SCAPH_STATUS_INVALID;
  if (nodenum == 0) return 0;
  RelaxWs ws = carve_relax_ws(workspace, 0);
  ws.cap = workspace_bytes > 256 ? (workspace_bytes - 256) / sizeof(RelaxChunk) : 0;
  if (ws.cap == 0) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  const long long wdelta =
      weights ? (long long)(reinterpret_cast<uintptr_t>(weights) -
                            reinterpret_cast<uintptr_t>(pool_base))
              : 0ll;
  hipLaunchKernelGGL(k_relax_zero, dim3(1), dim3(1), 0, st, ws.count);
  SCAPH_LAUNCH_CHECK();
  const unsigned grid = grid_for(nodenum, kBlock, 8192);
  if (weights) {
    hipLaunchKernelGGL(k_relax_page<1>, dim3(grid), dim3(kBlock), 0, st, buf, nodenum,
                       first_vertex, front_in, n, wdelta, value, front_out, counters, ws);
    SCAPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_relax_chunks<1>, dim3(1024), dim3(kBlock), 0, st, ws, n, wdelta, value,
                       front_out, counters);
  } else {
    hipLaunchKernelGGL(k_relax_page<0>, dim3(grid), dim3(kBlock), 0, st, buf, nodenum,
                       first_vertex, front_in, n, wdelta, value, front_out, counters, ws);
    SCAPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_relax_chunks<0>, dim3(1024), dim3(kBlock), 0, st, ws, n, wdelta, value,
                       front_out, counters);
  }
  SCAPH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
