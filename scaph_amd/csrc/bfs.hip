// bfs.hip — gfx950 BFS hot path: level init, push (top-down) expand,
// pull (bottom-up) expand, frontier conversions, per-page active counts and
// the level-array validator.
//
// The reference engine (src/gpu_engine.cuh:41-169) runs an asynchronous,
// label-correcting push: per edge a CAS-loop writeMin on value[] and a byte
// store into stat[].  Here every level is synchronous: the first edge that
// reaches an unvisited vertex claims it with one atomicOr on a visited
// bitmap (V/8 bytes, small enough to stay in the last-level cache), after
// which the level is written with a plain store.  BFS distances are unique,
// so the final levels equal the reference fixpoint (app/bfs_kernel.cuh:6-16
// relax d > s + 1 -> d = s + 1) for every schedule.
//
// Push load balance: the frontier's adjacency lists are laid end to end
// (block-local compaction + scan of the out-degrees), cut into tiles of
// 2048 edges (1024 on levels routed to XCDs by target), and every 256-thread
// block expands whole tiles — a hub with a million out-edges is spread over
// hundreds of blocks instead of one.
#include <map>
#include <mutex>
#include <tuple>

#include "common.h"

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kTileSmall = 10;               // log2 edges per tile of XCD-sliced levels
constexpr uint32_t kTileLarge = 11;               // log2 edges per tile of other levels
constexpr unsigned long long kSmallEdges = 8192;  // push frontiers this small run in one launch
constexpr uint32_t kPullSerial = 8;                // per-lane scan before wave help
constexpr uint32_t kHeavy = 512;                   // longer in-lists go to k_pull_heavy
constexpr uint32_t kHeavyChunk = 4096;             // in-edges per heavy work item
constexpr uint32_t kMaxHubs = 262144;              // hub frontier bits kept in LDS (<= 32 KB)
constexpr uint32_t kSummaryMaxBits = 1u << 22;     // frontier summary bits (512 KB)
constexpr uint32_t kHubTag = 0x80000000u;          // in-list entry = kHubTag | hub index
constexpr uint32_t kScanThreads = 1024;
constexpr uint32_t kBitsNodes = SCAPH_RUN_SEGMENT;  // Nodes per run segment (32 per thread)
constexpr uint32_t kSmallChunk = SCAPH_WAVE * 32;   // Nodes per wave step of k_push_runs_small
constexpr uint32_t kSlices = 8;                     // target slices of the sliced expand, one per XCD
constexpr uint32_t kMixed = kSlices;                // tiles whose targets span slices

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct PushWs {
  uint64_t* ent_ep;     // [nb*256] edge-list address of each kept entry (block-compacted)
  uint32_t* ent_len;    // [nb*256]
  uint32_t* blk_cnt;    // [nb]
  uint64_t* blk_edges;  // [nb]
  uint32_t* blk_coff;   // [nb] exclusive entry offset
  uint64_t* blk_eoff;   // [nb] exclusive edge offset
  uint64_t* gx;         // [M] global exclusive edge prefix (dense)
  uint64_t* gep;        // [M] edge-list address (dense)
  uint32_t* tstart;     // [tiles+1] entry containing the tile's first edge
  uint8_t* tslice;      // [tiles] target slice of each tile (XCD-sliced expand)
  uint32_t* tlist;      // [tiles] tiles grouped by slice (mixed last)
  uint32_t* tbin;       // [32] tiles per slice, then the scatter cursors
  uint32_t* part_cnt;   // [nb/1024+2] per-chunk sums and offsets of the parallel block scan
  uint64_t* part_edges;
  uint32_t* part_coff;
  uint64_t* part_eoff;
  unsigned long long* totals;  // [4] entries, edges, tiles
  size_t bytes;
};

template <typename Take>
void carve_parts(PushWs& w, size_t nb, Take& take) {
  const size_t parts = nb / kScanThreads + 2;
  w.part_cnt = reinterpret_cast<uint32_t*>(take(parts * 4));
  w.part_edges = reinterpret_cast<uint64_t*>(take(parts * 8));
  w.part_coff = reinterpret_cast<uint32_t*>(take(parts * 4));
  w.part_eoff = reinterpret_cast<uint64_t*>(take(parts * 8));
}

PushWs carve_push_ws(void* base, uint32_t max_entries, unsigned long long max_edges) {
  PushWs w;
  const size_t nb = (max_entries + kBlock - 1) / kBlock + 1;
  const size_t m = (size_t)max_entries + 1;
  const size_t tiles = (size_t)(max_edges >> kTileSmall) + 2;
  size_t off = 0;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* r = p ? p + off : nullptr;
    off += align256(bytes);
    return r;
  };
  w.ent_ep = reinterpret_cast<uint64_t*>(take(nb * kBlock * 8));
  w.ent_len = reinterpret_cast<uint32_t*>(take(nb * kBlock * 4));
  w.blk_cnt = reinterpret_cast<uint32_t*>(take(nb * 4));
  w.blk_edges = reinterpret_cast<uint64_t*>(take(nb * 8));
  w.blk_coff = reinterpret_cast<uint32_t*>(take(nb * 4));
  w.blk_eoff = reinterpret_cast<uint64_t*>(take(nb * 8));
  w.gx = reinterpret_cast<uint64_t*>(take(m * 8));
  w.gep = reinterpret_cast<uint64_t*>(take(m * 8));
  w.tstart = reinterpret_cast<uint32_t*>(take(tiles * 4));
  w.tslice = reinterpret_cast<uint8_t*>(take(tiles));
  w.tlist = reinterpret_cast<uint32_t*>(take(tiles * 4));
  w.tbin = reinterpret_cast<uint32_t*>(take(32 * 4));
  w.totals = reinterpret_cast<unsigned long long*>(take(64));
  carve_parts(w, nb, take);
  w.bytes = off;
  return w;
}

// ------------------------------------------------------------------ init

__global__ void k_fill_words(uint32_t* p, uint32_t v, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

__global__ void k_zero_u64(unsigned long long* p, uint32_t n) {
  for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0;
}


__global__ void __launch_bounds__(kBlock)
k_bfs_init(uint32_t* level, uint32_t* visited, const uint32_t* static_visited, uint32_t n,
           size_t words) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  // 16-byte stores for the aligned bulk of the level array
  const bool al = (reinterpret_cast<uintptr_t>(level) & 15) == 0;
  const size_t n4 = al ? n / 4 : 0;
  uint4* l4 = reinterpret_cast<uint4*>(level);
  const uint4 u4 = make_uint4(SCAPH_UNREACHED, SCAPH_UNREACHED, SCAPH_UNREACHED, SCAPH_UNREACHED);
  for (size_t i = tid; i < n4; i += stride) l4[i] = u4;
  for (size_t v = n4 * 4 + tid; v < n; v += stride) level[v] = SCAPH_UNREACHED;
  for (size_t w = tid; w < words; w += stride) {
    uint32_t x = static_visited ? static_visited[w] : 0u;
    // bits of vertices >= n never become active
    const size_t first = w * 32;
    if (first + 32 > n) {
      const uint32_t valid = first >= n ? 0u : (uint32_t)(n - first);
      x |= valid >= 32 ? 0u : (~0u << valid);
    }
    visited[w] = x;
  }
}

__global__ void k_bfs_set_source(uint32_t* level, uint32_t* visited, uint32_t src,
                                 uint32_t* queue, scaph_level_counters* counters) {
  level[src] = 0;
  visited[src >> 5] |= 1u << (src & 31);
  if (queue) queue[0] = src;
  if (counters) {
    counters->n_next = 0;
    counters->m_next = 0;
    counters->q_len = 0;
    counters->scanned = 0;
    counters->examined = 0;
    counters->reserved = 0;
  }
}

__global__ void __launch_bounds__(kBlock)
k_zero_degree_bitmap(uint32_t* bitmap, const uint32_t* in_degree, uint32_t n, size_t words) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t w = (size_t)blockIdx.x * blockDim.x + threadIdx.x; w < words; w += stride) {
    uint32_t x = 0;
    const size_t first = w * 32;
    for (uint32_t b = 0; b < 32; ++b) {
      const size_t v = first + b;
      if (v >= n || in_degree[v] == 0) x |= 1u << b;
    }
    bitmap[w] = x;
  }
}

// front = visited & ~snapshot: the vertices claimed since the snapshot,
// i.e. the frontier a push level produced, without one atomic per vertex.
__global__ void __launch_bounds__(kBlock)
k_visited_diff(const uint32_t* visited, const uint32_t* snapshot, uint32_t* front, size_t words) {
  const size_t n4 = words / 4;
  const uint4* a4 = reinterpret_cast<const uint4*>(visited);
  const uint4* b4 = reinterpret_cast<const uint4*>(snapshot);
  uint4* f4 = reinterpret_cast<uint4*>(front);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t tid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t i = tid; i < n4; i += stride) {
    const uint4 a = a4[i], b = b4[i];
    f4[i] = make_uint4(a.x & ~b.x, a.y & ~b.y, a.z & ~b.z, a.w & ~b.w);
  }
  for (size_t w = n4 * 4 + tid; w < words; w += stride) front[w] = visited[w] & ~snapshot[w];
}

// A small frontier in one launch: a wave per frontier vertex, its lanes
// striding the out-edges; claims as in k_expand, list appends aggregated
// per wave.  Used when the frontier's out-edges are few (the four-kernel
// tiled expand costs more in launches than in work there).
__global__ void __launch_bounds__(kBlock)
k_push_small(scaph_graph_desc g, const uint32_t* q_in, uint32_t n_in, const uint32_t* degree,
             uint32_t* level, uint32_t* visited, uint32_t* q_out, scaph_level_counters* counters,
             uint32_t next_level) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const PageTab t = stage_pages(g, dyn_lds);
  __syncthreads();
  const uint32_t n = g.num_vertices;
  const uint32_t lane = lane_id();
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  unsigned long long acc_n = 0, acc_m = 0, acc_s = 0;
  for (uint32_t i = blockIdx.x * nw + threadIdx.x / SCAPH_WAVE; i < n_in; i += gridDim.x * nw) {
    const uint32_t u = q_in[i];
    if (u >= n) continue;
    const NodeRef nr = load_node(t, g, u);
    for (uint32_t k0 = 0; k0 < nr.len; k0 += SCAPH_WAVE) {
      const uint32_t k = k0 + lane;
      bool claimed = false;
      uint32_t v = 0;
      if (k < nr.len) {
        ++acc_s;
        v = nr.edges[k];
        if (v < n) {
          const uint32_t w = v >> 5, bit = 1u << (v & 31);
          if (!(visited[w] & bit)) claimed = !(atomicOr(&visited[w], bit) & bit);
        }
      }
      const uint64_t m = __ballot(claimed);
      if (m) {
        unsigned long long base = 0;
        if (lane == 0) base = atomicAdd(&counters->q_len, (unsigned long long)__popcll(m));
        base = __shfl(base, 0, SCAPH_WAVE);
        if (claimed) {
          level[v] = next_level;
          q_out[base + lanes_below(m)] = v;
          acc_n += 1;
          acc_m += degree[v];
        }
      }
    }
  }
  acc_n = wave_sum(acc_n);
  acc_m = wave_sum(acc_m);
  acc_s = wave_sum(acc_s);
  if (lane == 0) {
    if (acc_n) atomicAdd(&counters->n_next, acc_n);
    if (acc_m) atomicAdd(&counters->m_next, acc_m);
    if (acc_s) atomicAdd(&counters->scanned, acc_s);
  }
}

// A frontier of at most kBlock vertices with its out-lists spread over the
// whole grid (a hub's edges no longer fall to one wave): every block reads
// the frontier's Nodes and the prefix of their out-degrees into LDS, and a
// thread takes kCoopEdges consecutive edges of the concatenated lists per
// pass (owner by binary search in the prefix, then advanced), all its
// target loads issued before its claims.  Claims as in k_expand, gathered
// per block in LDS and appended to the list with one atomic per pass.
constexpr uint32_t kCoopEdges = 8;
constexpr unsigned long long kCoopDefault = 1ull << 20;  // out-edge cap of the path

__global__ void __launch_bounds__(kBlock)
k_push_coop(scaph_graph_desc g, const uint32_t* q_in, uint32_t n_in, const uint32_t* degree,
            uint32_t* level, uint32_t* visited, uint32_t* q_out, scaph_level_counters* counters,
            uint32_t next_level) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  __shared__ const uint32_t* s_ep[kBlock];
  __shared__ unsigned long long s_off[kBlock + 1];
  __shared__ uint32_t s_claim[kBlock * kCoopEdges];
  __shared__ unsigned long long wtot[kBlock / SCAPH_WAVE];
  __shared__ uint32_t s_nc;
  __shared__ unsigned long long s_base;
  const PageTab t = stage_pages(g, dyn_lds);
  __syncthreads();
  const uint32_t n = g.num_vertices;
  unsigned long long len = 0;
  const uint32_t* ep = nullptr;
  if (threadIdx.x < n_in) {
    const uint32_t u = q_in[threadIdx.x];
    if (u < n) {
      const NodeRef nr = load_node(t, g, u);
      len = nr.len;
      ep = nr.edges;
    }
  }
  unsigned long long total;
  const unsigned long long incl = block256_incl_scan(len, wtot, &total);
  s_off[threadIdx.x] = incl - len;  // entries past n_in: len 0, offset total
  s_ep[threadIdx.x] = ep;
  if (threadIdx.x == 0) s_off[kBlock] = total;
  unsigned long long acc_s = 0, acc_m = 0, acc_n = 0;
  const uint64_t pass = (uint64_t)gridDim.x * kBlock * kCoopEdges;
  for (uint64_t c0 = (uint64_t)blockIdx.x * kBlock * kCoopEdges; c0 < total; c0 += pass) {
    if (threadIdx.x == 0) s_nc = 0;
    __syncthreads();  // also orders the prefix writes before the first pass's reads
    const uint64_t e0 = c0 + (uint64_t)threadIdx.x * kCoopEdges;
    uint32_t vv[kCoopEdges];
    if (e0 < total) {
      // owner of e0: the first list whose end lies past it
      uint32_t lo = 1, hi = kBlock;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_off[mid] > e0) hi = mid;
        else lo = mid + 1;
      }
      uint32_t own = lo - 1;
#pragma unroll
      for (uint32_t j = 0; j < kCoopEdges; ++j) {
        const uint64_t e = e0 + j;
        vv[j] = 0xFFFFFFFFu;
        if (e < total) {
          while (s_off[own + 1] <= e) ++own;
          vv[j] = s_ep[own][e - s_off[own]];
        }
      }
#pragma unroll
      for (uint32_t j = 0; j < kCoopEdges; ++j) {
        if (e0 + j >= total) continue;
        ++acc_s;
        const uint32_t v = vv[j];
        if (v >= n) continue;
        const uint32_t w = v >> 5, bit = 1u << (v & 31);
        if (visited[w] & bit) continue;
        if (atomicOr(&visited[w], bit) & bit) continue;
        level[v] = next_level;
        s_claim[atomicAdd(&s_nc, 1u)] = v;
        acc_n += 1;
        acc_m += degree[v];
      }
    }
    __syncthreads();
    const uint32_t nc = s_nc;
    if (threadIdx.x == 0) s_base = nc ? atomicAdd(&counters->q_len, (unsigned long long)nc) : 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nc; i += kBlock) q_out[s_base + i] = s_claim[i];
  }
  acc_s = wave_sum(acc_s);
  acc_m = wave_sum(acc_m);
  acc_n = wave_sum(acc_n);
  if (lane_id() == 0) {
    if (acc_n) atomicAdd(&counters->n_next, acc_n);
    if (acc_m) atomicAdd(&counters->m_next, acc_m);
    if (acc_s) atomicAdd(&counters->scanned, acc_s);
  }
}

// Frontier bits of vertices [v, v + c), c in 1..32: bit k is vertex v + k.
__device__ __forceinline__ uint32_t bits_at(const uint32_t* bm, uint32_t v, uint32_t c) {
  const uint32_t w = v >> 5, sh = v & 31;
  uint32_t b = bm[w] >> sh;
  if (sh && c > 32 - sh) b |= bm[w + 1] << (32 - sh);
  return c >= 32 ? b : (b & ((1u << c) - 1u));
}

// ------------------------------------------------------------ push: prep

// Block-local compaction of the entries with len > 0.
__device__ __forceinline__ void prep_store(bool keep, uint64_t ep, uint32_t len, uint64_t* ent_ep,
                                           uint32_t* ent_len, uint32_t* blk_cnt,
                                           uint64_t* blk_edges) {
  __shared__ uint32_t wcnt[kBlock / SCAPH_WAVE];
  __shared__ unsigned long long wedge[kBlock / SCAPH_WAVE];
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  const uint64_t m = __ballot(keep);
  const uint32_t r = lanes_below(m);
  const unsigned long long es = wave_sum((unsigned long long)(keep ? len : 0u));
  if (lane == 0) {
    wcnt[wid] = (uint32_t)__popcll(m);
    wedge[wid] = es;
  }
  __syncthreads();
  uint32_t off = 0, total = 0;
  unsigned long long etot = 0;
#pragma unroll
  for (uint32_t w = 0; w < kBlock / SCAPH_WAVE; ++w) {
    if (w < wid) off += wcnt[w];
    total += wcnt[w];
    etot += wedge[w];
  }
  const size_t base = (size_t)blockIdx.x * kBlock;
  if (keep) {
    ent_ep[base + off + r] = ep;
    ent_len[base + off + r] = len;
  }
  if (threadIdx.x == 0) {
    blk_cnt[blockIdx.x] = total;
    blk_edges[blockIdx.x] = etot;
  }
}

__global__ void __launch_bounds__(kBlock)
k_prep_queue(scaph_graph_desc g, const uint32_t* q, uint32_t n_in, uint64_t* ent_ep,
             uint32_t* ent_len, uint32_t* blk_cnt, uint64_t* blk_edges) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const PageTab s = stage_pages(g, dyn_lds);
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t len = 0;
  uint64_t ep = 0;
  if (i < n_in) {
    const uint32_t v = q[i];
    if (v < g.num_vertices) {
      NodeRef nr = load_node(s, g, v);
      len = nr.len;
      ep = reinterpret_cast<uint64_t>(nr.edges);
    }
  }
  prep_store(len > 0, ep, len, ent_ep, ent_len, blk_cnt, blk_edges);
}

__global__ void __launch_bounds__(kBlock)
k_prep_page(const uint32_t* buf, uint32_t nodenum, const uint32_t* front, uint64_t* ent_ep,
            uint32_t* ent_len, uint32_t* blk_cnt, uint64_t* blk_edges) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  uint32_t len = 0;
  uint64_t ep = 0;
  bool keep = false;
  if (i < nodenum) {
    const uint32_t vtx = buf[3ull * i];
    const uint32_t idx = buf[3ull * i + 1];
    len = buf[3ull * i + 2];
    keep = len > 0 && bit_test(front, vtx);
    ep = reinterpret_cast<uint64_t>(buf + 3ull * nodenum + idx);
  }
  prep_store(keep, ep, keep ? len : 0u, ent_ep, ent_len, blk_cnt, blk_edges);
}

// ------------------------------------------------------ push: page runs
//
// A run is a full page buffer on the device (Node i = vertex first + i).
// A multi-page push takes a table of runs; segment s (kBitsNodes Nodes) of
// the table belongs to the run j with runs[j].seg0 <= s < runs[j+1].seg0.
// The single-page entry passes its one run by value (runs == nullptr).
// Both passes read a segment through its frontier bits, 32 Nodes per
// thread, and load only the records of active Nodes: a sparse frontier
// costs nodenum/8 bytes of bitmap instead of 12 bytes per Node.

__device__ __forceinline__ scaph_page_run run_of(const scaph_page_run* runs, uint32_t num_runs,
                                                 const scaph_page_run& one, uint32_t seg) {
  if (!runs) return one;
  uint32_t lo = 0, hi = num_runs;  // runs[lo].seg0 <= seg < runs[hi].seg0
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (runs[mid].seg0 <= seg) lo = mid; else hi = mid;
  }
  return runs[lo];
}

// This thread's 32 Nodes of segment b: first Node index and frontier bits.
struct SegBits {
  const uint32_t* buf;
  uint32_t nodenum;
  uint32_t i0;
  uint32_t bits;
};

__device__ __forceinline__ SegBits seg_bits(const scaph_page_run& r, uint32_t b,
                                            const uint32_t* front) {
  SegBits s;
  s.buf = reinterpret_cast<const uint32_t*>(r.buf);
  s.nodenum = r.nodenum;
  s.i0 = (b - r.seg0) * kBitsNodes + threadIdx.x * 32;
  s.bits = s.i0 < r.nodenum ? bits_at(front, r.first_vertex + s.i0, min(32u, r.nodenum - s.i0))
                            : 0u;
  return s;
}

// Pass 1: per segment, the active Nodes with a non-empty out-list and their
// out-edges.
__global__ void __launch_bounds__(kBlock)
k_runs_count(const scaph_page_run* runs, uint32_t num_runs, scaph_page_run one,
             const uint32_t* front, uint32_t* blk_cnt, uint64_t* blk_edges) {
  __shared__ uint32_t wc[kBlock / SCAPH_WAVE];
  __shared__ unsigned long long we[kBlock / SCAPH_WAVE];
  const SegBits s = seg_bits(run_of(runs, num_runs, one, blockIdx.x), blockIdx.x, front);
  uint32_t cnt = 0;
  unsigned long long edges = 0;
  for (uint32_t m = s.bits; m; m &= m - 1) {
    const uint32_t len = s.buf[3ull * (s.i0 + __ffs(m) - 1) + 2];
    cnt += len ? 1u : 0u;
    edges += len;
  }
  uint32_t ctot;
  block256_incl_scan(cnt, wc, &ctot);
  unsigned long long etot;
  block256_incl_scan(edges, we, &etot);
  if (threadIdx.x == 0) {
    blk_cnt[blockIdx.x] = ctot;
    blk_edges[blockIdx.x] = etot;
  }
}

// Pass 2: the same Nodes, written densely in segment order — gx (position
// of the first out-edge in the level's edge sequence), gep (address of the
// out-list) — and the first entry of every tile.  Nothing is written when
// k_scan_blocks found more entries than the workspace holds (totals[0] = 0).
__global__ void __launch_bounds__(kBlock)
k_runs_mark(const scaph_page_run* runs, uint32_t num_runs, scaph_page_run one,
            const uint32_t* front, const uint32_t* blk_coff, const uint64_t* blk_eoff,
            uint64_t* gx, uint64_t* gep, uint32_t* tstart, const unsigned long long* totals) {
  __shared__ uint32_t wc[kBlock / SCAPH_WAVE];
  __shared__ unsigned long long we[kBlock / SCAPH_WAVE];
  const unsigned long long n_ent = totals[0];
  const unsigned long long n_tiles = totals[2];
  const uint32_t ts = (uint32_t)totals[4];
  if (n_ent == 0) return;
  const SegBits s = seg_bits(run_of(runs, num_runs, one, blockIdx.x), blockIdx.x, front);
  uint32_t cnt = 0;
  unsigned long long edges = 0;
  for (uint32_t m = s.bits; m; m &= m - 1) {
    const uint32_t len = s.buf[3ull * (s.i0 + __ffs(m) - 1) + 2];
    cnt += len ? 1u : 0u;
    edges += len;
  }
  uint32_t ctot;
  unsigned long long etot;
  uint32_t gi = blk_coff[blockIdx.x] + block256_incl_scan(cnt, wc, &ctot) - cnt;
  unsigned long long g = blk_eoff[blockIdx.x] + block256_incl_scan(edges, we, &etot) - edges;
  const uint32_t* eb = s.buf + 3ull * s.nodenum;
  for (uint32_t m = s.bits; m; m &= m - 1) {
    const uint32_t i = s.i0 + __ffs(m) - 1;
    const uint32_t len = s.buf[3ull * i + 2];
    if (!len || gi >= n_ent) continue;
    gx[gi] = g;
    gep[gi] = reinterpret_cast<uint64_t>(eb + s.buf[3ull * i + 1]);
    const unsigned long long t0 = (g + (1ull << ts) - 1) >> ts;
    unsigned long long t1 = (g + len + (1ull << ts) - 1) >> ts;
    if (t1 > n_tiles) t1 = n_tiles;
    for (unsigned long long t = t0; t < t1; ++t) tstart[t] = gi;
    ++gi;
    g += len;
  }
}

// One wave expands Node i of a page buffer: its lanes cover the out-list.
struct SmallAcc {
  unsigned long long n = 0, m = 0, s = 0;
};

__device__ __forceinline__ void push_node_wave(const uint32_t* buf, uint32_t nodenum, uint32_t i,
                                               const uint32_t* degree, uint32_t* level,
                                               uint32_t* visited, uint32_t* front_out,
                                               uint32_t next_level, uint32_t n, SmallAcc& acc) {
  const uint32_t lane = lane_id();
  const uint32_t idx = buf[3ull * i + 1], len = buf[3ull * i + 2];
  const uint32_t* eb = buf + 3ull * nodenum + idx;
  for (uint32_t k0 = 0; k0 < len; k0 += SCAPH_WAVE) {
    const uint32_t k = k0 + lane;
    if (k >= len) continue;
    ++acc.s;
    const uint32_t v = eb[k];
    if (v >= n) continue;
    const uint32_t w = v >> 5, bit = 1u << (v & 31);
    if (!(visited[w] & bit) && !(atomicOr(&visited[w], bit) & bit)) {
      level[v] = next_level;
      acc.n += 1;
      acc.m += degree[v];
      atomicOr(&front_out[w], bit);
    }
  }
}

__device__ __forceinline__ void small_acc_flush(SmallAcc acc, scaph_level_counters* counters) {
  const unsigned long long a = wave_sum(acc.n), b = wave_sum(acc.m), c = wave_sum(acc.s);
  if (lane_id() == 0) {
    if (a) atomicAdd(&counters->n_next, a);
    if (b) atomicAdd(&counters->m_next, b);
    if (c) atomicAdd(&counters->scanned, c);
  }
}

// One-launch push of runs whose active Nodes have few out-edges: a wave
// step covers kSmallChunk Nodes (32 frontier bits per lane), and the wave
// visits that chunk's active Nodes one after another.
__global__ void __launch_bounds__(kBlock)
k_push_runs_small(const scaph_page_run* runs, uint32_t num_runs, scaph_page_run one,
                  uint32_t num_segs, const uint32_t* front, const uint32_t* degree,
                  uint32_t* level, uint32_t* visited, uint32_t* front_out,
                  scaph_level_counters* counters, uint32_t next_level, uint32_t n) {
  constexpr uint32_t kPerSeg = kBitsNodes / kSmallChunk;
  const uint32_t lane = lane_id();
  const uint32_t nw = kBlock / SCAPH_WAVE;
  const uint32_t waves = gridDim.x * nw;
  const uint32_t units = num_segs * kPerSeg;
  SmallAcc acc;
  for (uint32_t u = blockIdx.x * nw + threadIdx.x / SCAPH_WAVE; u < units; u += waves) {
    const uint32_t seg = u / kPerSeg;
    const scaph_page_run r = run_of(runs, num_runs, one, seg);
    const uint32_t* buf = reinterpret_cast<const uint32_t*>(r.buf);
    const uint32_t c0 = (seg - r.seg0) * kBitsNodes + (u % kPerSeg) * kSmallChunk;
    const uint32_t i0 = c0 + lane * 32;
    const uint32_t bits =
        i0 < r.nodenum ? bits_at(front, r.first_vertex + i0, min(32u, r.nodenum - i0)) : 0u;
    for (uint64_t any = __ballot(bits != 0); any; any &= any - 1) {
      const uint32_t src = (uint32_t)__ffsll((unsigned long long)any) - 1;
      const uint32_t sb = __shfl(bits, (int)src, SCAPH_WAVE);
      for (uint32_t m = sb; m; m &= m - 1)
        push_node_wave(buf, r.nodenum, c0 + src * 32 + __ffs(m) - 1, degree, level, visited,
                       front_out, next_level, n, acc);
    }
  }
  small_acc_flush(acc, counters);
}

// One-launch push of a compacted buffer (every Node active), one Node per
// wave step.
__global__ void __launch_bounds__(kBlock)
k_push_compact_small(const uint32_t* buf, uint32_t nodenum, const uint32_t* degree,
                     uint32_t* level, uint32_t* visited, uint32_t* front_out,
                     scaph_level_counters* counters, uint32_t next_level, uint32_t n) {
  const uint32_t nw = kBlock / SCAPH_WAVE;
  SmallAcc acc;
  for (uint32_t i = blockIdx.x * nw + threadIdx.x / SCAPH_WAVE; i < nodenum; i += gridDim.x * nw)
    push_node_wave(buf, nodenum, i, degree, level, visited, front_out, next_level, n, acc);
  small_acc_flush(acc, counters);
}

// Exclusive scan of the per-block counts (one 1024-thread block).
__global__ void __launch_bounds__(kScanThreads)
k_scan_blocks(const uint32_t* blk_cnt, const uint64_t* blk_edges, uint32_t nb, uint32_t* blk_coff,
              uint64_t* blk_eoff, unsigned long long* totals, unsigned long long tile_cap,
              unsigned long long ent_cap, uint32_t tile_shift) {
  __shared__ uint32_t wc[kScanThreads / SCAPH_WAVE];
  __shared__ unsigned long long we[kScanThreads / SCAPH_WAVE];
  const uint32_t per = (nb + kScanThreads - 1) / kScanThreads;
  const uint32_t b0 = threadIdx.x * per;
  const uint32_t b1 = min(b0 + per, nb);
  uint32_t c = 0;
  unsigned long long e = 0;
  for (uint32_t b = b0; b < b1; ++b) {
    c += blk_cnt[b];
    e += blk_edges[b];
  }
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  const uint32_t ci = wave_incl_scan(c);
  const unsigned long long ei = wave_incl_scan_u64(e);
  if (lane == SCAPH_WAVE - 1) {
    wc[wid] = ci;
    we[wid] = ei;
  }
  __syncthreads();
  uint32_t coff = 0, ctot = 0;
  unsigned long long eoff = 0, etot = 0;
  for (uint32_t w = 0; w < kScanThreads / SCAPH_WAVE; ++w) {
    if (w < wid) {
      coff += wc[w];
      eoff += we[w];
    }
    ctot += wc[w];
    etot += we[w];
  }
  uint32_t rc = coff + ci - c;
  unsigned long long re = eoff + ei - e;
  for (uint32_t b = b0; b < b1; ++b) {
    blk_coff[b] = rc;
    blk_eoff[b] = re;
    rc += blk_cnt[b];
    re += blk_edges[b];
  }
  if (threadIdx.x == 0) {
    // A caller that under-reports the frontier's edges gets a truncated
    // (wrong, detectable) level instead of writes past the tile table.
    if (etot > (tile_cap << tile_shift)) etot = tile_cap << tile_shift;
    // More entries than the dense arrays hold: nothing is expanded.
    if (ctot > ent_cap) ctot = 0, etot = 0;
    totals[0] = ctot;
    totals[1] = etot;
    totals[2] = (etot + (1ull << tile_shift) - 1) >> tile_shift;
    totals[3] = tile_cap;
    totals[4] = tile_shift;
  }
}

// Exclusive scan over a 1024-thread block (16 waves); *total = block sum.
template <typename T>
__device__ __forceinline__ T block1024_excl_scan(T v, T* wtot, T* total) {
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  const T x = wave_incl_scan(v);
  if (lane == SCAPH_WAVE - 1) wtot[wid] = x;
  __syncthreads();
  T off = 0, all = 0;
#pragma unroll
  for (uint32_t w = 0; w < kScanThreads / SCAPH_WAVE; ++w) {
    const T t = wtot[w];
    if (w < wid) off += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return off + x - v;
}

// Parallel block scan for many prep blocks, in three launches: the sums of
// every 1024 consecutive block counts (k_scan_chunk_sums), k_scan_blocks
// over those sums, then each block's offset (k_scan_chunk_down).  All loads
// are coalesced; the single-block k_scan_blocks alone reads nb/1024
// strided values per thread.
__global__ void __launch_bounds__(kScanThreads)
k_scan_chunk_sums(const uint32_t* blk_cnt, const uint64_t* blk_edges, uint32_t nb,
                  uint32_t* part_cnt, uint64_t* part_edges) {
  __shared__ uint32_t wc[kScanThreads / SCAPH_WAVE];
  __shared__ unsigned long long we[kScanThreads / SCAPH_WAVE];
  const uint32_t b = blockIdx.x * kScanThreads + threadIdx.x;
  const uint32_t c = b < nb ? blk_cnt[b] : 0u;
  const unsigned long long e = b < nb ? blk_edges[b] : 0ull;
  uint32_t ct;
  unsigned long long et;
  block1024_excl_scan(c, wc, &ct);
  block1024_excl_scan(e, we, &et);
  if (threadIdx.x == 0) {
    part_cnt[blockIdx.x] = ct;
    part_edges[blockIdx.x] = et;
  }
}

__global__ void __launch_bounds__(kScanThreads)
k_scan_chunk_down(const uint32_t* blk_cnt, const uint64_t* blk_edges, uint32_t nb,
                  const uint32_t* part_coff, const uint64_t* part_eoff, uint32_t* blk_coff,
                  uint64_t* blk_eoff) {
  __shared__ uint32_t wc[kScanThreads / SCAPH_WAVE];
  __shared__ unsigned long long we[kScanThreads / SCAPH_WAVE];
  const uint32_t b = blockIdx.x * kScanThreads + threadIdx.x;
  const uint32_t c = b < nb ? blk_cnt[b] : 0u;
  const unsigned long long e = b < nb ? blk_edges[b] : 0ull;
  uint32_t ct;
  unsigned long long et;
  const uint32_t xc = block1024_excl_scan(c, wc, &ct);
  const unsigned long long xe = block1024_excl_scan(e, we, &et);
  if (b < nb) {
    blk_coff[b] = part_coff[blockIdx.x] + xc;
    blk_eoff[b] = part_eoff[blockIdx.x] + xe;
  }
}

// Dense entry arrays + the first entry of every tile.
__global__ void __launch_bounds__(kBlock)
k_mark_tiles(const uint64_t* ent_ep, const uint32_t* ent_len, const uint32_t* blk_cnt,
             const uint32_t* blk_coff, const uint64_t* blk_eoff, uint64_t* gx, uint64_t* gep,
             uint32_t* tstart, const unsigned long long* totals) {
  __shared__ unsigned long long wtot[kBlock / SCAPH_WAVE];
  const uint32_t b = blockIdx.x;
  const uint32_t j = threadIdx.x;
  const uint32_t cnt = blk_cnt[b];
  const size_t src = (size_t)b * kBlock + j;
  const unsigned long long len = j < cnt ? ent_len[src] : 0ull;
  unsigned long long total;
  const unsigned long long incl = block256_incl_scan(len, wtot, &total);
  if (j < cnt) {
    const uint32_t gi = blk_coff[b] + j;
    const unsigned long long g0 = blk_eoff[b] + incl - len;
    gx[gi] = g0;
    gep[gi] = ent_ep[src];
    const uint32_t ts = (uint32_t)totals[4];
    const unsigned long long t0 = (g0 + (1ull << ts) - 1) >> ts;
    unsigned long long t1 = (g0 + len + (1ull << ts) - 1) >> ts;
    if (t1 > totals[2]) t1 = totals[2];
    for (unsigned long long t = t0; t < t1; ++t) tstart[t] = gi;
  }
}

// --------------------------------------------------------- push: expand

enum { OUT_QUEUE = 0, OUT_BITMAP = 1 };

// Target slice of every tile (slice x = vertex ids [x n/8, (x+1) n/8)): the
// slice of its first and last target when the tile lies inside one out-list
// (lists are sorted by target) and the two agree, else kMixed.  Unsorted
// lists only mislabel tiles, which costs locality, not correctness.
__global__ void __launch_bounds__(kBlock)
k_tile_slices(const uint64_t* gx, const uint64_t* gep, const uint32_t* tstart,
              const unsigned long long* totals, uint8_t* tslice, uint32_t* tbin, uint32_t n) {
  __shared__ uint32_t hist[kSlices + 1];
  if (threadIdx.x <= kSlices) hist[threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long n_ent = totals[0];
  const unsigned long long n_edges = totals[1];
  const unsigned long long n_tiles = totals[2];
  const unsigned long long stride = (unsigned long long)gridDim.x * kBlock;
  for (unsigned long long t = (unsigned long long)blockIdx.x * kBlock + threadIdx.x; t < n_tiles;
       t += stride) {
    const unsigned long long r0 = t << totals[4];
    const unsigned long long r1 = min(r0 + (1ull << totals[4]), n_edges) - 1;
    const uint32_t i = tstart[t];
    uint8_t sl = kMixed;
    if (i + 1ull >= n_ent || gx[i + 1] > r1) {
      const uint32_t* e = reinterpret_cast<const uint32_t*>(gep[i]);
      const unsigned long long b = gx[i];
      const uint32_t a = e[r0 - b], z = e[r1 - b];
      if (a < n && z < n) {
        const uint32_t sa = (uint32_t)((uint64_t)a * kSlices / n);
        const uint32_t sz = (uint32_t)((uint64_t)z * kSlices / n);
        if (sa == sz) sl = (uint8_t)sa;
      }
    }
    tslice[t] = sl;
    atomicAdd(&hist[sl], 1u);
  }
  __syncthreads();
  if (threadIdx.x <= kSlices && hist[threadIdx.x]) atomicAdd(&tbin[threadIdx.x], hist[threadIdx.x]);
}

// Counting-sort scatter of the tiles into tlist, slice 0 first and the
// mixed tiles last (tbin[0..8] = counts, tbin[16..24] = cursors, cleared
// before k_tile_slices).
__global__ void __launch_bounds__(kBlock)
k_tile_scatter(const unsigned long long* totals, const uint8_t* tslice, uint32_t* tbin,
               uint32_t* tlist) {
  __shared__ uint32_t hist[kSlices + 1], base[kSlices + 1];
  const unsigned long long n_tiles = totals[2];
  const unsigned long long t0 = (unsigned long long)blockIdx.x * kBlock;
  if (t0 >= n_tiles) return;  // uniform per block
  if (threadIdx.x <= kSlices) hist[threadIdx.x] = 0;
  __syncthreads();
  const unsigned long long t = t0 + threadIdx.x;
  const bool live = t < n_tiles;
  const uint32_t sl = live ? tslice[t] : 0u;
  const uint32_t rank = live ? atomicAdd(&hist[sl], 1u) : 0u;
  __syncthreads();
  if (threadIdx.x <= kSlices) {
    uint32_t off = 0;
    for (uint32_t b = 0; b < threadIdx.x; ++b) off += tbin[b];
    base[threadIdx.x] =
        off + (hist[threadIdx.x] ? atomicAdd(&tbin[16 + threadIdx.x], hist[threadIdx.x]) : 0u);
  }
  __syncthreads();
  if (live) tlist[base[sl] + rank] = (uint32_t)t;
}

// LAYOUT 0: each thread takes kPerThread consecutive edges of the tile;
// LAYOUT 1: edge j*kBlock + thread (one coalesced 256-edge row per step).
template <int MODE, int LAYOUT, uint32_t kTile>
__global__ void __launch_bounds__(kBlock)
k_expand(const uint64_t* gx, const uint64_t* gep, const uint32_t* tstart,
         const unsigned long long* totals, const uint32_t* tlist, const uint32_t* tbin,
         const uint32_t* degree,
         uint32_t* level, uint32_t* visited, uint32_t* q_out, uint32_t* front_out,
         scaph_level_counters* counters, uint32_t next_level, uint32_t n) {
  constexpr uint32_t kPerThread = kTile / kBlock;  // edges per thread per tile
  static_assert(kTile < 65536, "tile offsets are kept as u16");
  // entry k of the tile: its first edge relative to the tile start (0 for
  // the entry the tile starts inside, whose list address is advanced to
  // match) and its list address; u16 offsets keep the block's LDS small
  // enough for five blocks per CU with 2048-edge tiles
  __shared__ uint16_t s_gx[kTile + 1];
  __shared__ unsigned long long s_ep[kTile + 1];
  __shared__ uint32_t s_claim[kTile];
  __shared__ uint32_t s_cnt;
  __shared__ unsigned long long s_base;
  __shared__ unsigned long long red[3][kBlock / SCAPH_WAVE];
  const unsigned long long n_ent = totals[0];
  const unsigned long long n_edges = totals[1];
  const unsigned long long n_tiles = totals[2];
  unsigned long long acc_n = 0, acc_m = 0;
  // Sliced (tlist != nullptr): block b = 8j + x takes entries j, j + G/8,
  // ... of slice x's part of tlist, then entries b, b + G, ... of the mixed
  // part.  Workgroups are dispatched to the eight XCDs round-robin (block b
  // on XCD (s + b) mod 8 for a per-launch start s), so slice x's visited
  // words stay in one XCD's L2.  Unsliced: tiles b, b + G, ...
  uint32_t sl_beg = 0, sl_cnt = 0, mx_beg = 0, mx_cnt = 0;
  if (tlist) {
    const uint32_t xs = blockIdx.x % kSlices;
    for (uint32_t b = 0; b < kMixed; ++b) {
      const uint32_t c = tbin[b];
      if (b < xs) sl_beg += c;
      if (b == xs) sl_cnt = c;
      mx_beg += c;
    }
    mx_cnt = tbin[kMixed];
  }
  const unsigned long long sl_steps =
      tlist ? (sl_cnt > blockIdx.x / kSlices
                   ? (sl_cnt - blockIdx.x / kSlices + gridDim.x / kSlices - 1) / (gridDim.x / kSlices)
                   : 0)
            : 0;
  const unsigned long long steps =
      tlist ? sl_steps + (mx_cnt > blockIdx.x ? (mx_cnt - blockIdx.x + gridDim.x - 1) / gridDim.x : 0)
            : (n_tiles > blockIdx.x ? (n_tiles - blockIdx.x + gridDim.x - 1) / gridDim.x : 0);
  for (unsigned long long k = 0; k < steps; ++k) {
    unsigned long long t;
    if (!tlist) t = blockIdx.x + k * gridDim.x;
    else if (k < sl_steps) t = tlist[sl_beg + blockIdx.x / kSlices + k * (gridDim.x / kSlices)];
    else t = tlist[mx_beg + blockIdx.x + (k - sl_steps) * gridDim.x];
    const unsigned long long s0 = t * kTile;
    const unsigned long long e0 = min(s0 + kTile, n_edges);
    const uint32_t i0 = tstart[t];
    uint32_t i1 = (t + 1 < n_tiles) ? tstart[t + 1] : (uint32_t)(n_ent - 1);
    if (i1 < i0) i1 = i0;
    uint32_t cnt = i1 - i0 + 1;
    if (cnt > kTile + 1) cnt = kTile + 1;
    for (uint32_t k = threadIdx.x; k < cnt; k += kBlock) {
      const unsigned long long x = gx[i0 + k], ep = gep[i0 + k];
      // entry i0 holds edge s0, so only entry 0 starts at or before the
      // tile start; the others start inside it (x - s0 in [1, kTile]).
      // Keyed on k, not on x < s0: with "x < s0 ? 0 : min(x - s0, kTile)"
      // hipcc -O3 (ROCm 7.2) emitted min(x - s0, kTile) with a wrapping
      // subtract and no select, storing kTile for entry 0.
      const unsigned long long d = x - s0;
      s_gx[k] = k == 0 ? (uint16_t)0 : (uint16_t)(d > kTile ? kTile : d);
      s_ep[k] = k == 0 ? ep + (s0 - x) * sizeof(uint32_t) : ep;
    }
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    // edge offsets relative to the tile start s0
    const uint32_t r0 = LAYOUT == 0 ? threadIdx.x * kPerThread : threadIdx.x;
    const uint32_t rend = (uint32_t)(e0 - s0);
    uint32_t lo = 0, hi = cnt;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_gx[mid] <= r0) lo = mid; else hi = mid;
    }
#pragma unroll
    for (uint32_t j = 0; j < kPerThread; ++j) {
      const uint32_t r = LAYOUT == 0 ? r0 + j : r0 + j * kBlock;
      if (r < rend) {
        if (LAYOUT == 0) {
          while (lo + 1 < cnt && s_gx[lo + 1] <= r) ++lo;
        } else {
          uint32_t h = cnt;  // rows only move forward: search [lo, cnt)
          while (h - lo > 1) {
            const uint32_t mid = (lo + h) >> 1;
            if (s_gx[mid] <= r) lo = mid; else h = mid;
          }
        }
        const uint32_t* ep = reinterpret_cast<const uint32_t*>(s_ep[lo]);
        const uint32_t v = ep[r - s_gx[lo]];
        if (v < n) {
          const uint32_t w = v >> 5, bit = 1u << (v & 31);
          if (!(visited[w] & bit)) {
            const uint32_t old = atomicOr(&visited[w], bit);
            if (!(old & bit)) {
              level[v] = next_level;
              acc_n += 1;
              if (MODE == OUT_BITMAP) {
                acc_m += degree[v];
                atomicOr(&front_out[w], bit);
              } else {
                // the out-degree is read when the tile's claims are written
                // out, off the claim path
                s_claim[atomicAdd(&s_cnt, 1u)] = v;
              }
            }
          }
        }
      }
    }
    __syncthreads();
    if (MODE == OUT_QUEUE) {
      // one global append per tile, then coalesced queue writes
      const uint32_t c = s_cnt;
      if (threadIdx.x == 0 && c) s_base = atomicAdd(&counters->q_len, (unsigned long long)c);
      __syncthreads();
      for (uint32_t k = threadIdx.x; k < c; k += kBlock) {
        const uint32_t v = s_claim[k];
        q_out[s_base + k] = v;
        acc_m += degree[v];
      }
      __syncthreads();
    }
  }
  const unsigned long long sn = wave_sum(acc_n);
  const unsigned long long sm = wave_sum(acc_m);
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  if (lane_id() == 0) {
    red[0][wid] = sn;
    red[1][wid] = sm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0;
    for (uint32_t w = 0; w < kBlock / SCAPH_WAVE; ++w) {
      a += red[0][w];
      b += red[1][w];
    }
    if (a) atomicAdd(&counters->n_next, a);
    if (b) atomicAdd(&counters->m_next, b);
    if (blockIdx.x == 0) atomicAdd(&counters->scanned, n_edges);
  }
}

// ------------------------------------------------------------------ pull
//
// Each wave owns a window of kPullWinWords bitmap words (1024 vertices):
// one coalesced read of the visited words, the unvisited vertices of the
// window compacted into LDS (ballot/prefix over popcounts), then processed
// 64 at a time: a lane reads its vertex's in-Node and probes kPullProbe
// in-neighbours per step against the frontier bitmap (independent loads,
// so several misses are in flight per lane), stopping at the first hit.
// Lists longer than kPullSerial entries finish with the whole wave probing
// 64 in-neighbours per step.  Found vertices set bits in an LDS copy of the
// window, written back as whole words (visited |= found, next = found).

struct HeavyChunk {
  uint32_t v;
  uint32_t len;
  uint64_t ptr;
};

struct PullWs {
  unsigned long long* count;  // heavy chunks appended
  uint32_t* hub_front;        // frontier bits of the hubs (kMaxHubs bits)
  uint32_t* summary;          // frontier summary (kSummaryMaxBits bits)
  HeavyChunk* chunks;
  unsigned long long cap;
  size_t bytes;
  uint32_t summary_shift;     // 0: no summary; else one bit per 2^shift vertices
};

PullWs carve_pull_ws(void* base, unsigned long long m) {
  PullWs w;
  w.cap = m / kHeavyChunk + m / kHeavy + 64;  // >= sum over lists of ceil((len-kHeavy)/chunk)
  const size_t hub_bytes = kMaxHubs / 8;
  const size_t summary_bytes = kSummaryMaxBits / 8;
  w.bytes = 256 + hub_bytes + summary_bytes + w.cap * sizeof(HeavyChunk);
  char* p = static_cast<char*>(base);
  w.count = reinterpret_cast<unsigned long long*>(p);
  w.hub_front = p ? reinterpret_cast<uint32_t*>(p + 256) : nullptr;
  w.summary = p ? reinterpret_cast<uint32_t*>(p + 256 + hub_bytes) : nullptr;
  w.chunks = p ? reinterpret_cast<HeavyChunk*>(p + 256 + hub_bytes + summary_bytes) : nullptr;
  w.summary_shift = 0;
  return w;
}

// Frontier summary: bit i is set when any of vertices [i << shift,
// (i + 1) << shift) is in the frontier (shift >= 5, so a bit covers whole
// frontier words).  Words up to (nbits + 31) / 32 are written; blk is the
// block's index among ceil(nbits / kBlock).
__device__ __forceinline__ void front_summary(uint32_t blk, const uint32_t* front, uint64_t words,
                                              uint32_t* summary, uint32_t nbits, uint32_t shift) {
  const uint32_t i = blk * kBlock + threadIdx.x;
  bool b = false;
  if (i < nbits) {
    const uint64_t w0 = (uint64_t)i << (shift - 5);
    const uint64_t w1e = w0 + (1ull << (shift - 5));
    const uint64_t w1 = w1e < words ? w1e : words;
    uint32_t acc = 0;
    for (uint64_t w = w0; w < w1; ++w) acc |= front[w];
    b = acc != 0;
  }
  const uint64_t m = __ballot(b);
  const uint32_t lane = lane_id();
  const uint32_t base = (blk * kBlock + threadIdx.x - lane) / 32;
  const uint32_t nw = (nbits + 31) / 32;
  if (lane == 0 && base < nw) summary[base] = (uint32_t)m;
  if (lane == 32 && base + 1 < nw) summary[base + 1] = (uint32_t)(m >> 32);
}

// Frontier bits of the hub list, packed (hub i -> bit i); blk is the
// block's index among ceil(num_hubs / kBlock).
__device__ __forceinline__ void hub_gather(uint32_t blk, const uint32_t* hub_list,
                                           uint32_t num_hubs, const uint32_t* front,
                                           uint32_t* hub_front) {
  const uint32_t i = blk * kBlock + threadIdx.x;
  const bool b = i < num_hubs && bit_test(front, hub_list[i]);
  const uint64_t m = __ballot(b);
  const uint32_t lane = lane_id();
  const uint32_t base = (blk * kBlock + threadIdx.x - lane) / 32;
  if (lane == 0 && base < kMaxHubs / 32) hub_front[base] = (uint32_t)m;
  if (lane == 32 && base + 1 < kMaxHubs / 32) hub_front[base + 1] = (uint32_t)(m >> 32);
}

// What a pull level needs before its expand, in one launch: the heavy-chunk
// counter zeroed, the hubs' frontier bits gathered (blocks [0, hub_blocks))
// and the frontier summary built (the nbits > 0 blocks after them).
__global__ void __launch_bounds__(kBlock)
k_pull_prologue(unsigned long long* count, const uint32_t* hub_list, uint32_t num_hubs,
                uint32_t hub_blocks, const uint32_t* front, uint64_t words, uint32_t* hub_front,
                uint32_t* summary, uint32_t nbits, uint32_t shift) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *count = 0ull;
  if (blockIdx.x < hub_blocks)
    hub_gather(blockIdx.x, hub_list, num_hubs, front, hub_front);
  else if (nbits)
    front_summary(blockIdx.x - hub_blocks, front, words, summary, nbits, shift);
}

// The hub bits sit in dynamic LDS behind the page table; a launch passes
// page_lds_bytes(num_pages) + hub_lds_bytes(num_hubs) bytes.
inline size_t hub_lds_bytes(uint32_t num_hubs) { return ((size_t)(num_hubs + 127) / 128) * 16; }

__device__ __forceinline__ uint32_t* hub_lds(uint8_t* dyn_lds, uint32_t num_pages) {
  const uint32_t np = num_pages < SCAPH_LDS_PAGES ? num_pages : SCAPH_LDS_PAGES;
  return reinterpret_cast<uint32_t*>(dyn_lds + (size_t)np * 16);
}

__device__ __forceinline__ void stage_hubs(uint32_t* s_hub, const uint32_t* hub_front,
                                           uint32_t num_hubs) {
  for (uint32_t i = threadIdx.x; i < (num_hubs + 31) / 32; i += blockDim.x) s_hub[i] = hub_front[i];
}

// Is in-list entry u (a vertex id, or kHubTag | hub index) in the frontier?
__device__ __forceinline__ bool in_front(uint32_t u, const uint32_t* s_hub,
                                         const uint32_t* front) {
  if (u & kHubTag) {
    const uint32_t h = u & ~kHubTag;
    return (s_hub[h >> 5] >> (h & 31)) & 1u;
  }
  return bit_test(front, u);
}

__device__ __forceinline__ uint32_t page_left(const PageTab& t, const scaph_graph_desc& g,
                                              uint32_t p) {
  return t.in_lds ? t.left[p] : g.pages[p].left;
}

// Node of v, starting the page search at `hint` (a page with left <= v).
__device__ __forceinline__ NodeRef load_node_hint(const PageTab& t, const scaph_graph_desc& g,
                                                  uint32_t v, uint32_t hint) {
  uint32_t p = hint;
  while (p + 1 < g.num_pages && page_left(t, g, p + 1) <= v) ++p;
  uint32_t left, nn;
  const uint32_t* base;
  if (t.in_lds) {
    left = t.left[p];
    nn = t.nodenum[p];
    base = reinterpret_cast<const uint32_t*>(t.ptr[p]);
  } else {
    left = g.pages[p].left;
    nn = g.pages[p].nodenum;
    base = reinterpret_cast<const uint32_t*>(g.page_ptr[p]);
  }
  const uint32_t i = v - left;
  NodeRef r;
  const uint32_t idx = base[3u * i + 1u];
  r.len = base[3u * i + 2u];
  r.edges = base + 3ull * nn + idx;
  return r;
}

// The edge array of v's page (its list starts at the Node idx there),
// from the page table alone: no Node record is read.
__device__ __forceinline__ const uint32_t* page_edges_hint(const PageTab& t,
                                                           const scaph_graph_desc& g, uint32_t v,
                                                           uint32_t hint) {
  uint32_t p = hint;
  while (p + 1 < g.num_pages && page_left(t, g, p + 1) <= v) ++p;
  if (t.in_lds)
    return reinterpret_cast<const uint32_t*>(t.ptr[p]) + 3ull * t.nodenum[p];
  return reinterpret_cast<const uint32_t*>(g.page_ptr[p]) + 3ull * g.pages[p].nodenum;
}

constexpr uint32_t kPullWinWords = 32;
constexpr uint32_t kPullWin = kPullWinWords * 32;
constexpr uint32_t kPullProbe = 4;
constexpr uint32_t kPullWaves = 8;  // waves per pull block by default

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// What one pull batch needs besides its vertices.
struct PullCtx {
  scaph_graph_desc g;
  PageTab s;
  const uint32_t* s_hub;
  const uint32_t* front_in;
  const uint4* head;
  uint32_t head_outdeg, head_indexed;
  PullWs hw;
  scaph_level_counters* counters;
};

// in_front, with the frontier summary (when the context has one) tested
// before a non-hub entry's frontier word: a clear summary bit answers from a
// small, cache-resident array instead of a random line of the bitmap.
__device__ __forceinline__ bool probe(const PullCtx& pc, uint32_t u) {
  if (u & kHubTag) {
    const uint32_t h = u & ~kHubTag;
    return (pc.s_hub[h >> 5] >> (h & 31)) & 1u;
  }
  if (pc.hw.summary_shift && !bit_test(pc.hw.summary, u >> pc.hw.summary_shift)) return false;
  return bit_test(pc.front_in, u);
}

// probe() in two steps, so that a batch's loads can be requested one batch
// before its answers are needed: probe_word requests the word that decides
// entry u (none for a hub: its bit is in LDS), probe_answer reads it.
// With a summary the word is the summary's; a set bit then still needs the
// frontier word, read at once.
__device__ __forceinline__ uint32_t probe_word(const PullCtx& pc, uint32_t u, bool need) {
  if (!need || (u & kHubTag)) return 0u;
  if (pc.hw.summary_shift) return pc.hw.summary[(u >> pc.hw.summary_shift) >> 5];
  return pc.front_in[u >> 5];
}

__device__ __forceinline__ bool probe_answer(const PullCtx& pc, uint32_t u, bool need,
                                             uint32_t w) {
  if (!need) return false;
  if (u & kHubTag) {
    const uint32_t h = u & ~kHubTag;
    return (pc.s_hub[h >> 5] >> (h & 31)) & 1u;
  }
  if (pc.hw.summary_shift) {
    if (!((w >> ((u >> pc.hw.summary_shift) & 31)) & 1u)) return false;
    return bit_test(pc.front_in, u);
  }
  return (w >> (u & 31)) & 1u;
}

// The words of one head record's entries (head_probe_words) and the batch's
// answer from them (head_probe_answer): PHASE 1 of pull_batch split in two.
struct HeadWords {
  uint32_t w0, w1, w2;
};

__device__ __forceinline__ uint32_t head_len(const PullCtx& pc, const uint4& hd) {
  return pc.head_indexed ? (hd.z & 0xFFu) : pc.head_outdeg ? hd.z : hd.w;
}

__device__ __forceinline__ HeadWords head_probe_words(const PullCtx& pc, bool act,
                                                      const uint4& hd) {
  const uint32_t cap = pc.head_outdeg ? 2u : 3u;
  const uint32_t len = head_len(pc, hd);
  const uint32_t l0 = act ? (len < cap ? len : cap) : 0u;
  HeadWords w;
  w.w0 = probe_word(pc, hd.x, l0 > 0);
  w.w1 = probe_word(pc, hd.y, l0 > 1);
  w.w2 = probe_word(pc, hd.z, l0 > 2);
  return w;
}

__device__ __forceinline__ bool head_probe_answer(const PullCtx& pc, bool act, const uint4& hd,
                                                  const HeadWords& w, uint32_t& outdeg,
                                                  unsigned long long& acc_s, bool& more) {
  const uint32_t cap = pc.head_outdeg ? 2u : 3u;
  const uint32_t len = head_len(pc, hd);
  const uint32_t l0 = act ? (len < cap ? len : cap) : 0u;
  outdeg = pc.head_outdeg ? hd.w : 0xFFFFFFFFu;
  const bool found = probe_answer(pc, hd.x, l0 > 0, w.w0) |
                     probe_answer(pc, hd.y, l0 > 1, w.w1) |
                     probe_answer(pc, hd.z, l0 > 2, w.w2);
  more = act && !found && len > cap;
  acc_s += l0 < kPullSerial ? l0 : kPullSerial;
  return found;
}

// One pull batch, one vertex per lane (act): does v have an in-neighbour in
// the frontier?  hd is v's head record when the context has heads; hint is
// a page at or before v's.  outdeg gets v's out-degree from the head record
// (0xFFFFFFFF when it has none); acc_s counts the in-edges scanned.  Called
// by the whole wave (the long lists are probed cooperatively).
// PHASE 0 probes the head and then the list; with head records, PHASE 1
// probes only the head (more: the list must still be read) and PHASE 2
// only the list past the head, for vertices PHASE 1 left undecided.
template <int PHASE>
__device__ __forceinline__ bool pull_batch(const PullCtx& pc, bool act, uint32_t v, const uint4& hd,
                                           uint32_t hint, uint32_t& outdeg,
                                           unsigned long long& acc_s, bool& more) {
  const scaph_graph_desc& g = pc.g;
  const PageTab& s = pc.s;
  const uint4* head = pc.head;
  const uint32_t head_outdeg = pc.head_outdeg, head_indexed = pc.head_indexed;
  const PullWs& hw = pc.hw;
  scaph_level_counters* counters = pc.counters;
  const uint32_t lane = lane_id();
  bool found = false, longl = false;
  NodeRef nr;
  nr.edges = nullptr;
  nr.len = 0;
  outdeg = 0xFFFFFFFFu;
  more = false;
  if (act) {
    uint32_t k = 0, k_head = 0;
    if (head) {
      // the first three in-list entries and the in-degree sit in a
      // 16-byte record read in vertex order: a vertex found there (the
      // lists start with their hubs) never touches its page
      // (head_outdeg: {e0, e1, in-degree, out-degree} instead, so a
      // found vertex also needs no degree[] load)
      // (head_indexed: {e0, e1, Node idx << 8 | min(in-degree, 255),
      // out-degree}, so a list past its head needs no Node record)
      const uint32_t cap = head_outdeg ? 2u : 3u;
      const uint32_t len = head_indexed ? (hd.z & 0xFFu) : head_outdeg ? hd.z : hd.w;
      if (head_outdeg) outdeg = hd.w;
      nr.len = len;
      const uint32_t l0 = len < cap ? len : cap;
      k = l0;
      if (PHASE != 2) {
        const bool h0 = l0 > 0 && probe(pc, hd.x);
        const bool h1 = l0 > 1 && probe(pc, hd.y);
        const bool h2 = l0 > 2 && probe(pc, hd.z);
        found = h0 | h1 | h2;
      }
      if (PHASE == 1) {
        more = !found && len > cap;
        acc_s += l0 < kPullSerial ? l0 : kPullSerial;
        return found;
      }
      k_head = l0 < kPullSerial ? l0 : kPullSerial;  // counted by PHASE 1
      if (!found && len > cap) {
        if (head_indexed && len < 255u)
          nr.edges = page_edges_hint(s, g, v, hint) + (hd.z >> 8);
        else
          nr = load_node_hint(s, g, v, hint);
      }
    } else {
      nr = load_node_hint(s, g, v, hint);
    }
    const uint32_t lim = nr.len < kPullSerial ? nr.len : kPullSerial;
    for (; !found && k < lim; k += kPullProbe) {
      uint32_t u[kPullProbe];
#pragma unroll
      for (uint32_t j = 0; j < kPullProbe; ++j) u[j] = (k + j < lim) ? nr.edges[k + j] : 0u;
      bool h = false;
#pragma unroll
      for (uint32_t j = 0; j < kPullProbe; ++j) h |= (k + j < lim) && probe(pc, u[j]);
      if (h) {
        found = true;
        k += kPullProbe;
        break;
      }
    }
    acc_s += (k < lim ? k : lim) - (PHASE == 2 ? (k_head < lim ? k_head : lim) : 0u);
    longl = !found && nr.len > kPullSerial;
  }
  if (PHASE == 1) return found;
  // lists past the serial probes: the whole wave probes 64 at a time up
  // to kHeavy entries; what is left after that is deferred
  uint64_t ml = __ballot(longl);
  while (ml) {
    const int j = __ffsll((unsigned long long)ml) - 1;
    ml &= ml - 1;
    const uint64_t ep = wave_bcast64(reinterpret_cast<uint64_t>(nr.edges), j);
    const uint32_t len = wave_bcast(nr.len, j);
    const uint32_t stop = len < kHeavy ? len : kHeavy;
    const uint32_t* e = reinterpret_cast<const uint32_t*>(ep);
    bool hitj = false;
    for (uint32_t k = kPullSerial; k < stop; k += SCAPH_WAVE) {
      const uint32_t kk = k + lane;
      const bool hit = kk < stop && probe(pc, e[kk]);
      const uint64_t mh = __ballot(hit);
      if (lane == 0) acc_s += min(stop - k, (uint32_t)SCAPH_WAVE);
      if (mh) {
        hitj = true;
        break;
      }
    }
    if (hitj) {
      if ((int)lane == j) found = true;
    } else if (len > kHeavy && (int)lane == j) {
      // grid-wide follow-up: one work item per kHeavyChunk in-edges
      const uint32_t rest = len - kHeavy;
      const uint32_t nch = (rest + kHeavyChunk - 1) / kHeavyChunk;
      const unsigned long long at = atomicAdd(hw.count, (unsigned long long)nch);
      if (at + nch <= hw.cap) {
        for (uint32_t c = 0; c < nch; ++c) {
          HeavyChunk h;
          h.v = v;
          h.len = min(kHeavyChunk, rest - c * kHeavyChunk);
          h.ptr = reinterpret_cast<uint64_t>(e + kHeavy + (uint64_t)c * kHeavyChunk);
          hw.chunks[at + c] = h;
        }
      } else {
        counters->reserved = 1;  // overflow: the host reports an error
      }
    }
  }
  return found;
}

// WAVES waves per block share one copy of the hub bits; the window queue
// holds 10-bit window offsets as u16, so 8-wave blocks fit three to a CU.
// MINW > 1 asks the compiler for registers that let MINW waves share a SIMD.
// TWO_PHASE (head records required): a window's vertices are first probed by
// their head records alone, then the undecided ones by their lists, so that
// a batch's lanes wait for chains of like length.
template <int PREFETCH, int WAVES, int MINW, int TWO_PHASE>
__global__ void __launch_bounds__(WAVES * SCAPH_WAVE, MINW)
k_pull(scaph_graph_desc g, const uint4* head, const uint32_t* no_in, const uint32_t* live_base,
       uint32_t head_outdeg, uint32_t head_indexed, const uint32_t* degree, uint32_t* level,
       uint32_t* visited, const uint32_t* front_in, uint32_t* front_out,
       scaph_level_counters* counters, uint32_t next_level, PullWs hw, uint32_t num_hubs) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  __shared__ uint16_t s_q[WAVES][kPullWin];
  __shared__ uint32_t s_fw[WAVES][kPullWinWords];
  uint32_t* const s_hub = hub_lds(dyn_lds, g.num_pages);
  __shared__ unsigned long long red[4][WAVES];
  // head records by live rank: per window word, its live (in-edge) bits and
  // the live vertices before it in the window
  __shared__ uint32_t s_lw[WAVES][kPullWinWords];
  __shared__ uint32_t s_lb[WAVES][kPullWinWords];
  stage_hubs(s_hub, hw.hub_front, num_hubs);
  const PageTab s = stage_pages(g, dyn_lds);
  __syncthreads();
  const PullCtx pc{g, s, s_hub, front_in, head, head_outdeg, head_indexed, hw, counters};
  const uint32_t n = g.num_vertices;
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  const uint64_t per_block = (uint64_t)kPullWinWords * WAVES;
  const uint64_t iters = (words + per_block - 1) / per_block;
  unsigned long long acc_n = 0, acc_m = 0, acc_s = 0, acc_x = 0;
  uint16_t* q = s_q[wid];
  uint32_t* fw = s_fw[wid];
  // visited word of this lane in the wave's window of iteration bi, and
  // (live-rank heads) its no-in-edge word and the window's first rank
  auto visited_word = [&](uint64_t bi) -> uint32_t {
    const uint64_t ww = bi * per_block + (uint64_t)wid * kPullWinWords + lane;
    return (lane < kPullWinWords && ww < words) ? visited[ww] : ~0u;
  };
  auto no_in_word = [&](uint64_t bi) -> uint32_t {
    const uint64_t ww = bi * per_block + (uint64_t)wid * kPullWinWords + lane;
    return (live_base && lane < kPullWinWords && ww < words) ? no_in[ww] : ~0u;
  };
  auto window_base = [&](uint64_t bi) -> uint32_t {
    const uint64_t ww = bi * per_block + (uint64_t)wid * kPullWinWords;
    return (live_base && ww < words) ? live_base[ww / kPullWinWords] : 0u;
  };
  const bool first = blockIdx.x < iters;
  uint32_t vw_next = PREFETCH && first ? visited_word(blockIdx.x) : 0u;
  uint32_t nw_next = PREFETCH && first ? no_in_word(blockIdx.x) : 0u;
  uint32_t wb_next = PREFETCH && first ? window_base(blockIdx.x) : 0u;
  for (uint64_t bi = blockIdx.x; bi < iters; bi += gridDim.x) {
    const uint64_t w0 = bi * per_block + (uint64_t)wid * kPullWinWords;
    const uint64_t w = w0 + lane;
    const bool own = lane < kPullWinWords && w < words;
    uint32_t vw, nw, wb;
    if (PREFETCH) {
      // this window's words were requested one window earlier; request the
      // next window's
      vw = vw_next;
      nw = nw_next;
      wb = wb_next;
      if (bi + gridDim.x < iters) {
        vw_next = visited_word(bi + gridDim.x);
        nw_next = no_in_word(bi + gridDim.x);
        wb_next = window_base(bi + gridDim.x);
      }
    } else {
      vw = own ? visited[w] : ~0u;
      nw = no_in_word(bi);
      wb = window_base(bi);
    }
    uint32_t unv = ~vw;
    const uint32_t c = __popc(unv);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t tot = wave_bcast(incl, SCAPH_WAVE - 1);
    uint32_t o = incl - c;
    while (unv) {
      const uint32_t b = __ffs(unv) - 1;
      unv &= unv - 1;
      q[o++] = (uint16_t)((lane << 5) | b);
    }
    if (lane < kPullWinWords) fw[lane] = 0u;
    uint32_t wbase = 0;
    if (live_base && tot) {
      const uint32_t lw = own ? ~nw : 0u;
      const uint32_t lc = __popc(lw);
      const uint32_t li = wave_incl_scan(lc);
      if (lane < kPullWinWords) {
        s_lw[wid][lane] = lw;
        s_lb[wid][lane] = li - lc;
      }
      wbase = wb;
    }
    wave_lds_sync();
    // index of the head record of window vertex loc
    auto hidx = [&](uint32_t loc) -> uint32_t {
      if (!live_base) return (uint32_t)(w0 * 32) + loc;
      const uint32_t j = loc >> 5;
      return wbase + s_lb[wid][j] + __popc(s_lw[wid][j] & ((1u << (loc & 31)) - 1u));
    };
    const uint32_t hint = tot ? find_page(s, g, (uint32_t)(w0 * 32)) : 0u;
    // the head records of the next batch are requested before this batch
    // is probed (one extra load in flight per lane)
    uint32_t nq2 = 0;  // two phases: vertices moved to the list phase (wave-uniform)
    if (TWO_PHASE == 2) {
      // head phase, pipelined: while batch b is answered, the words that
      // answer batch b + 1 and the head records of batch b + 2 are in flight
      auto head_of = [&](uint32_t i) -> uint4 {
        return i < tot ? head[hidx(q[i])] : make_uint4(0u, 0u, 0u, 0u);
      };
      uint4 hd_c = head_of(lane);
      uint4 hd_n = head_of(SCAPH_WAVE + lane);
      HeadWords pw_c = head_probe_words(pc, lane < tot, hd_c);
      for (uint32_t i0 = 0; i0 < tot; i0 += SCAPH_WAVE) {
        const uint32_t i = i0 + lane;
        const bool act = i < tot;
        const uint32_t loc = act ? q[i] : 0u;
        const uint32_t v = (uint32_t)(w0 * 32) + loc;
        const uint4 hd_n2 = head_of(i + 2 * SCAPH_WAVE);
        const HeadWords pw_n = head_probe_words(pc, i + SCAPH_WAVE < tot, hd_n);
        if (act) ++acc_x;
        uint32_t outdeg;
        bool more;
        const bool found = head_probe_answer(pc, act, hd_c, pw_c, outdeg, acc_s, more);
        if (found) {
          level[v] = next_level;
          acc_n += 1;
          acc_m += outdeg != 0xFFFFFFFFu ? outdeg : degree[v];
          atomicOr(&fw[loc >> 5], 1u << (loc & 31));
        }
        const uint64_t mm = __ballot(more);
        if (more) q[nq2 + lanes_below(mm)] = (uint16_t)loc;
        nq2 += (uint32_t)__popcll(mm);
        hd_c = hd_n;
        hd_n = hd_n2;
        pw_c = pw_n;
      }
    }
    uint4 hd_next = make_uint4(0u, 0u, 0u, 0u);
    if (TWO_PHASE != 2 && PREFETCH && head && lane < tot) hd_next = head[hidx(q[lane])];
    for (uint32_t i0 = 0; TWO_PHASE != 2 && i0 < tot; i0 += SCAPH_WAVE) {
      const uint32_t i = i0 + lane;
      const bool act = i < tot;
      const uint32_t loc = act ? q[i] : 0u;
      const uint32_t v = (uint32_t)(w0 * 32) + loc;
      const uint4 hd_cur = hd_next;
      if (PREFETCH && head && i + SCAPH_WAVE < tot) hd_next = head[hidx(q[i + SCAPH_WAVE])];
      uint4 hd = hd_cur;
      if (!PREFETCH && head && act) hd = head[hidx(loc)];
      if (act) ++acc_x;
      uint32_t outdeg;
      bool more;
      const bool found = TWO_PHASE ? pull_batch<1>(pc, act, v, hd, hint, outdeg, acc_s, more)
                                   : pull_batch<0>(pc, act, v, hd, hint, outdeg, acc_s, more);
      if (found) {
        level[v] = next_level;
        acc_n += 1;
        acc_m += outdeg != 0xFFFFFFFFu ? outdeg : degree[v];
        atomicOr(&fw[loc >> 5], 1u << (loc & 31));
      }
      if (TWO_PHASE) {
        // the vertices whose list must still be read move to the front of
        // the window queue (below every index this or a later batch reads)
        const uint64_t mm = __ballot(more);
        if (more) q[nq2 + lanes_below(mm)] = (uint16_t)loc;
        nq2 += (uint32_t)__popcll(mm);
      }
    }
    if (TWO_PHASE && nq2) {
      // their lists, in batches of vertices that all need one
      wave_lds_sync();
      uint4 hd2_next = make_uint4(0u, 0u, 0u, 0u);
      if (lane < nq2) hd2_next = head[hidx(q[lane])];
      for (uint32_t i0 = 0; i0 < nq2; i0 += SCAPH_WAVE) {
        const uint32_t i = i0 + lane;
        const bool act = i < nq2;
        const uint32_t loc = act ? q[i] : 0u;
        const uint32_t v = (uint32_t)(w0 * 32) + loc;
        const uint4 hd = hd2_next;
        if (i + SCAPH_WAVE < nq2) hd2_next = head[hidx(q[i + SCAPH_WAVE])];
        uint32_t outdeg;
        bool more;
        const bool found = pull_batch<2>(pc, act, v, hd, hint, outdeg, acc_s, more);
        if (found) {
          level[v] = next_level;
          acc_n += 1;
          acc_m += outdeg != 0xFFFFFFFFu ? outdeg : degree[v];
          atomicOr(&fw[loc >> 5], 1u << (loc & 31));
        }
      }
    }
    wave_lds_sync();
    if (own) {
      const uint32_t f = fw[lane];
      visited[w] = vw | f;
      front_out[w] = f;
    }
    wave_lds_sync();
  }
  const unsigned long long sn = wave_sum(acc_n);
  const unsigned long long sm = wave_sum(acc_m);
  const unsigned long long ss = wave_sum(acc_s);
  const unsigned long long sx = wave_sum(acc_x);
  if (lane == 0) {
    red[0][wid] = sn;
    red[1][wid] = sm;
    red[2][wid] = ss;
    red[3][wid] = sx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b = 0, c = 0, x = 0;
    for (uint32_t w = 0; w < WAVES; ++w) {
      a += red[0][w];
      b += red[1][w];
      c += red[2][w];
      x += red[3][w];
    }
    if (a) atomicAdd(&counters->n_next, a);
    if (b) atomicAdd(&counters->m_next, b);
    if (c) atomicAdd(&counters->scanned, c);
    if (x) atomicAdd(&counters->examined, x);
  }
}

// Sparse pull levels (few unvisited vertices left).  k_pull gives a wave
// one 1024-vertex window at a time; with one or two unvisited vertices per
// window its lanes idle and the wave runs those vertices' dependent chains
// (head record, list, probe) one window after the other.  Here a wave
// gathers the unvisited vertices of its windows, with their head-record
// index, into an LDS buffer and probes them 64 at a time once the buffer
// fills or its windows end.  Found vertices are claimed with atomics on
// visited and on the next frontier, which the launch before clears.
constexpr uint32_t kSparseBuf = 256;  // vertices a wave holds before it probes them

template <int WAVES>
__global__ void __launch_bounds__(WAVES * SCAPH_WAVE)
k_pull_sparse(scaph_graph_desc g, const uint4* head, const uint32_t* no_in,
              const uint32_t* live_base, uint32_t head_outdeg, uint32_t head_indexed,
              const uint32_t* degree, uint32_t* level, uint32_t* visited,
              const uint32_t* front_in, uint32_t* front_out, scaph_level_counters* counters,
              uint32_t next_level, PullWs hw, uint32_t num_hubs) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  __shared__ uint32_t s_v[WAVES][kSparseBuf];
  __shared__ uint32_t s_h[WAVES][kSparseBuf];
  __shared__ unsigned long long red[4][WAVES];
  uint32_t* const s_hub = hub_lds(dyn_lds, g.num_pages);
  stage_hubs(s_hub, hw.hub_front, num_hubs);
  const PageTab s = stage_pages(g, dyn_lds);
  __syncthreads();
  const PullCtx pc{g, s, s_hub, front_in, head, head_outdeg, head_indexed, hw, counters};
  const uint32_t n = g.num_vertices;
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  uint32_t* const sv = s_v[wid];
  uint32_t* const sh = s_h[wid];
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  // the walk takes two 1024-vertex windows per step, one word per lane
  constexpr uint32_t kSpan = 2 * kPullWinWords;
  const uint64_t windows = (words + kSpan - 1) / kSpan;
  const uint64_t nwaves = (uint64_t)gridDim.x * WAVES;
  unsigned long long acc_n = 0, acc_m = 0, acc_s = 0, acc_x = 0;
  uint32_t held = 0;  // wave-uniform
  auto drain = [&]() {
    wave_lds_sync();
    for (uint32_t i0 = 0; i0 < held; i0 += SCAPH_WAVE) {
      const uint32_t i = i0 + lane;
      const bool act = i < held;
      const uint32_t v = act ? sv[i] : 0u;
      uint4 hd = make_uint4(0u, 0u, 0u, 0u);
      if (head && act) hd = head[sh[i]];
      const uint32_t hint = act ? find_page(s, g, v) : 0u;
      if (act) ++acc_x;
      uint32_t outdeg;
      bool more;
      const bool found = pull_batch<0>(pc, act, v, hd, hint, outdeg, acc_s, more);
      if (found) {
        const uint32_t bit = 1u << (v & 31);
        level[v] = next_level;
        acc_n += 1;
        acc_m += outdeg != 0xFFFFFFFFu ? outdeg : degree[v];
        atomicOr(&visited[v >> 5], bit);
        atomicOr(&front_out[v >> 5], bit);
      }
    }
    wave_lds_sync();
    held = 0;
  };
  const uint64_t first = (uint64_t)blockIdx.x * WAVES + wid;
  // a window's visited word, in-edge word and live base are requested one
  // window ahead, so the walk waits on none of them
  uint32_t vw_next = ~0u, nw_next = ~0u, base_next = 0u;
  auto request = [&](uint64_t win) {
    const uint64_t w = win * kSpan + lane;
    const bool own = w < words;
    vw_next = own ? visited[w] : ~0u;
    if (live_base) {
      nw_next = own ? no_in[w] : ~0u;
      base_next = own ? live_base[w / kPullWinWords] : 0u;  // this lane's window
    }
  };
  if (first < windows) request(first);
  for (uint64_t win = first; win < windows; win += nwaves) {
    const uint64_t w = win * kSpan + lane;
    const uint32_t vw = vw_next, nw = nw_next, wbase = base_next;
    if (win + nwaves < windows) request(win + nwaves);
    const uint32_t unv = ~vw;
    const uint32_t c = __popc(unv);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t tot = wave_bcast(incl, SCAPH_WAVE - 1);
    if (tot == 0) continue;
    // head index = live rank, as k_pull's hidx computes it: the rank
    // restarts at each window's first word (lane 32 starts the second)
    uint32_t lw = 0, lb = 0;
    if (live_base) {
      lw = ~nw;
      const uint32_t lc = __popc(lw);
      const uint32_t li = wave_incl_scan(lc);
      const uint32_t first_half = wave_bcast(li, kPullWinWords - 1);
      lb = li - lc - (lane >= kPullWinWords ? first_half : 0u);
    }
    // the window's vertices in order, kSparseBuf at a time
    for (uint32_t lo = 0; lo < tot; lo += kSparseBuf) {
      const uint32_t take = tot - lo < kSparseBuf ? tot - lo : kSparseBuf;
      if (held + take > kSparseBuf) drain();
      uint32_t o = incl - c;
      for (uint32_t m = unv; m; m &= m - 1, ++o) {
        if (o < lo || o >= lo + take) continue;
        const uint32_t b = __ffs(m) - 1;
        sv[held + o - lo] = (uint32_t)(w * 32 + b);
        sh[held + o - lo] =
            live_base ? wbase + lb + __popc(lw & ((1u << b) - 1u)) : (uint32_t)(w * 32 + b);
      }
      held += take;
    }
  }
  if (held) drain();
  const unsigned long long sn = wave_sum(acc_n);
  const unsigned long long sm = wave_sum(acc_m);
  const unsigned long long ss = wave_sum(acc_s);
  const unsigned long long sx = wave_sum(acc_x);
  if (lane == 0) {
    red[0][wid] = sn;
    red[1][wid] = sm;
    red[2][wid] = ss;
    red[3][wid] = sx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long a = 0, b2 = 0, c2 = 0, x = 0;
    for (uint32_t i = 0; i < WAVES; ++i) {
      a += red[0][i];
      b2 += red[1][i];
      c2 += red[2][i];
      x += red[3][i];
    }
    if (a) atomicAdd(&counters->n_next, a);
    if (b2) atomicAdd(&counters->m_next, b2);
    if (c2) atomicAdd(&counters->scanned, c2);
    if (x) atomicAdd(&counters->examined, x);
  }
}

// Heavy in-lists deferred by k_pull: each block takes one chunk of
// kHeavyChunk in-edges of one vertex and probes them 256 at a time against
// the frontier, stopping at the first hit; the first block to hit claims
// the vertex (atomicOr on the visited word decides).
__global__ void __launch_bounds__(kBlock)
k_pull_heavy(const HeavyChunk* chunks, const unsigned long long* count, const uint32_t* degree,
             uint32_t* level, uint32_t* visited, const uint32_t* front_in, uint32_t* front_out,
             scaph_level_counters* counters, uint32_t next_level, const uint32_t* hub_front,
             uint32_t num_hubs) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  __shared__ HeavyChunk s_c;
  __shared__ int s_skip;
  uint32_t* const s_hub = hub_lds(dyn_lds, 0);
  const unsigned long long total = *count;
  // most levels defer nothing: leave before staging the hub bits (a
  // block-uniform exit)
  if (blockIdx.x >= total) return;
  stage_hubs(s_hub, hub_front, num_hubs);
  __syncthreads();
  unsigned long long scanned = 0;
  for (unsigned long long e = blockIdx.x; e < total; e += gridDim.x) {
    if (threadIdx.x == 0) {
      s_c = chunks[e];
      s_skip = bit_test(visited, s_c.v) ? 1 : 0;  // another chunk already won
    }
    __syncthreads();
    const HeavyChunk c = s_c;
    const int skip = s_skip;
    __syncthreads();
    if (skip) continue;
    const uint32_t* ep = reinterpret_cast<const uint32_t*>(c.ptr);
    bool hit = false;
    for (uint32_t k0 = 0; k0 < c.len; k0 += kBlock) {
      const uint32_t k = k0 + threadIdx.x;
      const bool h = k < c.len && in_front(ep[k], s_hub, front_in);
      scanned += (k < c.len) ? 1 : 0;
      if (__syncthreads_or(h)) {
        hit = true;
        break;
      }
      if ((k0 / kBlock) % 4 == 3) {
        // another chunk of the same list may have won meanwhile; every wave
        // reads the flag before any wave can leave the loop and overwrite
        // it for the next chunk
        if (threadIdx.x == 0)
          s_skip = (int)((__hip_atomic_load(&visited[c.v >> 5], __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT) >> (c.v & 31)) & 1u);
        __syncthreads();
        const int won = s_skip;
        __syncthreads();
        if (won) break;
      }
    }
    if (hit && threadIdx.x == 0) {
      const uint32_t v = c.v, w = v >> 5, bit = 1u << (v & 31);
      const uint32_t old = atomicOr(&visited[w], bit);
      if (!(old & bit)) {
        level[v] = next_level;
        atomicOr(&front_out[w], bit);
        atomicAdd(&counters->n_next, 1ull);
        atomicAdd(&counters->m_next, (unsigned long long)degree[v]);
      }
    }
  }
  scanned = wave_sum(scanned);
  if (lane_id() == 0 && scanned) atomicAdd(&counters->scanned, scanned);
}

// ------------------------------------------------------- conversions

__global__ void __launch_bounds__(kBlock)
k_queue_to_bitmap(const uint32_t* q, uint32_t nq, uint32_t* bitmap, uint32_t n) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += stride) {
    const uint32_t v = q[i];
    if (v < n) atomicOr(&bitmap[v >> 5], 1u << (v & 31));
  }
}

// Frontier bitmap -> vertex list.  Each thread owns kQueueWords consecutive
// words (its block kBlock * kQueueWords), requested at once as 16-byte loads
// and kept in registers.  One block scan gives every thread its place; the
// block takes its slice of the list with a single atomic (few blocks, so
// the shared counter is not a queue of atomics) and each thread writes its
// vertices in order.
constexpr uint32_t kQueueWords = 32;

__global__ void __launch_bounds__(kBlock)
k_bitmap_to_queue(const uint32_t* bitmap, uint32_t n, uint32_t* q,
                  scaph_level_counters* counters) {
  __shared__ uint32_t wtot[kBlock / SCAPH_WAVE];
  __shared__ unsigned long long s_base;
  const uint64_t words = ((uint64_t)n + 31) / 32;
  const uint64_t w0 = ((uint64_t)blockIdx.x * kBlock + threadIdx.x) * kQueueWords;
  uint32_t bits[kQueueWords];
  if (w0 + kQueueWords <= words) {
    const uint4* p4 = reinterpret_cast<const uint4*>(bitmap + w0);
#pragma unroll
    for (uint32_t j = 0; j < kQueueWords / 4; ++j) {
      const uint4 t = p4[j];
      bits[4 * j] = t.x;
      bits[4 * j + 1] = t.y;
      bits[4 * j + 2] = t.z;
      bits[4 * j + 3] = t.w;
    }
  } else {
#pragma unroll
    for (uint32_t j = 0; j < kQueueWords; ++j) bits[j] = w0 + j < words ? bitmap[w0 + j] : 0u;
  }
  uint32_t cnt = 0;
#pragma unroll
  for (uint32_t j = 0; j < kQueueWords; ++j) {
    const uint64_t w = w0 + j;
    if (w < words && w * 32 + 32 > n) {
      const uint32_t valid = (uint32_t)(n - w * 32);
      bits[j] &= valid >= 32 ? ~0u : ((1u << valid) - 1u);
    }
    cnt += __popc(bits[j]);
  }
  uint32_t total;
  const uint32_t incl = block256_incl_scan(cnt, wtot, &total);
  if (threadIdx.x == 0) s_base = total ? atomicAdd(&counters->q_len, (unsigned long long)total) : 0;
  __syncthreads();
  if (total == 0) return;
  unsigned long long o = s_base + incl - cnt;
#pragma unroll
  for (uint32_t j = 0; j < kQueueWords; ++j) {
    uint32_t x = bits[j];
    while (x) {
      const uint32_t b = __ffs(x) - 1;
      x &= x - 1;
      q[o++] = (uint32_t)((w0 + j) * 32 + b);
    }
  }
}

// --------------------------------------------------------- page counts

__global__ void __launch_bounds__(kBlock)
k_page_counts(const scaph_page* pages, uint32_t num_pages, const uint32_t* front,
              const uint32_t* degree, int weighted, uint32_t* nodes, uint32_t* datas,
              unsigned long long* totals) {
  __shared__ unsigned long long red[2][kBlock / SCAPH_WAVE];
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  for (uint32_t p = blockIdx.x; p < num_pages; p += gridDim.x) {
    const uint32_t left = pages[p].left, right = pages[p].right;
    const uint32_t w0 = left >> 5, w1 = right >> 5;
    unsigned long long cn = 0, cd = 0;
    for (uint32_t w = w0 + threadIdx.x; w <= w1; w += kBlock) {
      uint32_t bits = front[w];
      if (w == w0) bits &= ~0u << (left & 31);
      if (w == w1) bits &= (right & 31) == 31 ? ~0u : ((1u << ((right & 31) + 1)) - 1u);
      cn += __popc(bits);
      while (bits) {
        const uint32_t b = __ffs(bits) - 1;
        bits &= bits - 1;
        const uint32_t d = degree[w * 32 + b];
        cd += weighted ? (2ull * d + 3ull) : (d + 3ull);
      }
    }
    cn = wave_sum(cn);
    cd = wave_sum(cd);
    if (lane == 0) {
      red[0][wid] = cn;
      red[1][wid] = cd;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long a = 0, b = 0;
      for (uint32_t w = 0; w < kBlock / SCAPH_WAVE; ++w) {
        a += red[0][w];
        b += red[1][w];
      }
      nodes[p] = (uint32_t)a;
      datas[p] = (uint32_t)b;
      if (a) atomicAdd(&totals[0], a);
      if (b) atomicAdd(&totals[1], b);
    }
    __syncthreads();
  }
}


// Per-page count of vertices that can still become active: set in `live`
// (NULL: every vertex) and clear in `visited`.  A page whose count is 0
// holds no vertex any later level can reach.
__global__ void __launch_bounds__(kBlock)
k_page_unvisited(const scaph_page* pages, uint32_t num_pages, const uint32_t* visited,
                 const uint32_t* live, uint32_t* out) {
  __shared__ uint32_t red[kBlock / SCAPH_WAVE];
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  for (uint32_t p = blockIdx.x; p < num_pages; p += gridDim.x) {
    const uint32_t left = pages[p].left, right = pages[p].right;
    uint32_t c = 0;
    if (pages[p].nodenum) {
      const uint32_t w0 = left >> 5, w1 = right >> 5;
      for (uint32_t w = w0 + threadIdx.x; w <= w1; w += kBlock) {
        uint32_t bits = ~visited[w] & (live ? live[w] : ~0u);
        if (w == w0) bits &= ~0u << (left & 31);
        if (w == w1) bits &= (right & 31) == 31 ? ~0u : ((1u << ((right & 31) + 1)) - 1u);
        c += __popc(bits);
      }
    }
    c = wave_sum(c);
    if (lane_id() == 0) red[wid] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t t = 0;
      for (uint32_t w = 0; w < kBlock / SCAPH_WAVE; ++w) t += red[w];
      out[p] = t;
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ validation

__global__ void __launch_bounds__(kBlock)
k_validate_edges(scaph_graph_desc g, const uint32_t* level, unsigned long long* out) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const PageTab s = stage_pages(g, dyn_lds);
  const uint32_t n = g.num_vertices;
  unsigned long long bad = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t u = blockIdx.x * blockDim.x + threadIdx.x; u < n; u += stride) {
    const uint32_t lu = level[u];
    if (lu == SCAPH_UNREACHED) continue;
    NodeRef nr = load_node(s, g, u);
    for (uint32_t k = 0; k < nr.len; ++k) {
      const uint32_t lv = level[nr.edges[k]];
      if (lv == SCAPH_UNREACHED || lv > lu + 1) ++bad;
    }
  }
  bad = wave_sum(bad);
  if (lane_id() == 0 && bad) atomicAdd(&out[0], bad);
}

__global__ void __launch_bounds__(kBlock)
k_validate_parents(scaph_graph_desc gin, const uint32_t* hub_list, const uint32_t* level,
                   uint32_t src, unsigned long long* out) {
  extern __shared__ __align__(16) uint8_t dyn_lds[];
  const PageTab s = stage_pages(gin, dyn_lds);
  const uint32_t n = gin.num_vertices;
  unsigned long long bad = 0;
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const uint32_t lv = level[v];
    if (lv == SCAPH_UNREACHED) continue;
    if (v == src) {
      if (lv != 0) ++bad;
      continue;
    }
    if (lv == 0) { ++bad; continue; }
    NodeRef nr = load_node(s, gin, v);
    bool ok = false;
    for (uint32_t k = 0; k < nr.len && !ok; ++k) {
      uint32_t u = nr.edges[k];
      if (u & kHubTag) u = hub_list[u & ~kHubTag];
      ok = level[u] == lv - 1;
    }
    if (!ok) ++bad;
  }
  bad = wave_sum(bad);
  if (lane_id() == 0 && bad) atomicAdd(&out[1], bad);
}

// Compute units of the current device (cached per process).
inline unsigned cu_count() {
  static unsigned cus = 0;
  if (!cus) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
      cus = (unsigned)v;
    else
      cus = 256;
  }
  return cus;
}

inline unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  size_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

// Blocks of `kernel` that stay resident on one CU at this block size and
// dynamic LDS (the runtime's occupancy calculation, cached per process).
unsigned resident_blocks(const void* kernel, unsigned threads, size_t lds) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, unsigned, size_t>, unsigned> cache;
  const auto key = std::make_tuple(kernel, threads, lds);
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, (int)threads, lds) != hipSuccess ||
      nb < 1)
    nb = 1;
  cache.emplace(key, (unsigned)nb);
  return (unsigned)nb;
}

// Offsets of the prep blocks' entries and edges, and the level totals.
hipError_t launch_block_scan(const PushWs& w, uint32_t nb, unsigned long long tile_cap,
                             unsigned long long ent_cap, uint32_t tile_shift,
                             const scaph_push_tuning* tune, hipStream_t st) {
  const uint32_t single = tune && tune->scan_single ? tune->scan_single : kScanThreads;
  if (nb <= single) {
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, st, w.blk_cnt, w.blk_edges,
                       nb, w.blk_coff, w.blk_eoff, w.totals, tile_cap, ent_cap, tile_shift);
    return hipGetLastError();
  }
  const uint32_t parts = (nb + kScanThreads - 1) / kScanThreads;
  hipLaunchKernelGGL(k_scan_chunk_sums, dim3(parts), dim3(kScanThreads), 0, st, w.blk_cnt,
                     w.blk_edges, nb, w.part_cnt, w.part_edges);
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(kScanThreads), 0, st, w.part_cnt, w.part_edges,
                     parts, w.part_coff, w.part_eoff, w.totals, tile_cap, ent_cap, tile_shift);
  hipLaunchKernelGGL(k_scan_chunk_down, dim3(parts), dim3(kScanThreads), 0, st, w.blk_cnt,
                     w.blk_edges, nb, w.part_coff, w.part_eoff, w.blk_coff, w.blk_eoff);
  return hipGetLastError();
}

// Levels routed to XCDs by target use 1024-edge tiles (a tile inside one
// sorted list then spans fewer targets); the others 2048-edge tiles (fewer
// per-tile stagings and appends).
bool sliced_level(unsigned long long max_edges, const scaph_push_tuning* tune) {
  return tune && tune->xcd_slices && (max_edges >> kTileSmall) >= tune->xcd_slices;
}

uint32_t tile_shift_for(unsigned long long max_edges, const scaph_push_tuning* tune) {
  return sliced_level(max_edges, tune) ? kTileSmall : kTileLarge;
}

int launch_expand(const PushWs& w, unsigned long long max_edges, const uint32_t* degree,
                  uint32_t* level, uint32_t* visited, uint32_t* q_out, uint32_t* front_out,
                  scaph_level_counters* counters, uint32_t next_level, uint32_t n, int mode,
                  uint32_t tile_shift, const scaph_push_tuning* tune, hipStream_t st) {
  const unsigned long long tiles = (max_edges + (1ull << tile_shift) - 1) >> tile_shift;
  unsigned grid = grid_for((size_t)tiles, 1, 8192);
  if (tune && tune->grid_blocks) grid = tune->grid_blocks < grid ? tune->grid_blocks : grid;
  const bool rows = tune && tune->edge_rows;
  const uint32_t* tlist = nullptr;
  if (tile_shift == kTileSmall && sliced_level(max_edges, tune)) {
    hipLaunchKernelGGL(k_fill_words, dim3(1), dim3(64), 0, st, w.tbin, 0u, 32u);
    SCAPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_slices, dim3(grid_for((size_t)tiles, kBlock, 4096)), dim3(kBlock), 0,
                       st, w.gx, w.gep, w.tstart, w.totals, w.tslice, w.tbin, n);
    SCAPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_tile_scatter, dim3(grid_for((size_t)tiles, kBlock, 1u << 30)),
                       dim3(kBlock), 0, st, w.totals, w.tslice, w.tbin, w.tlist);
    SCAPH_LAUNCH_CHECK();
    tlist = w.tlist;
    grid = grid < kSlices ? kSlices : grid / kSlices * kSlices;
  }
#define SCAPH_EXPAND(M, L, T)                                                                    \
  hipLaunchKernelGGL((k_expand<M, L, T>), dim3(grid), dim3(kBlock), 0, st, w.gx, w.gep,          \
                     w.tstart, w.totals, tlist, w.tbin, degree, level, visited, q_out,          \
                     front_out, counters, next_level, n)
#define SCAPH_EXPAND_T(M, L)                                                                     \
  if (tile_shift == kTileSmall) SCAPH_EXPAND(M, L, 1u << kTileSmall);                          \
  else SCAPH_EXPAND(M, L, 1u << kTileLarge)
  if (mode == OUT_QUEUE) {
    if (rows) { SCAPH_EXPAND_T(OUT_QUEUE, 1); }
    else { SCAPH_EXPAND_T(OUT_QUEUE, 0); }
  } else {
    if (rows) { SCAPH_EXPAND_T(OUT_BITMAP, 1); }
    else { SCAPH_EXPAND_T(OUT_BITMAP, 0); }
  }
#undef SCAPH_EXPAND_T
#undef SCAPH_EXPAND
  SCAPH_LAUNCH_CHECK();
  return 0;
}

// Entries prepared by k_prep_queue / k_prep_page: block scan, dense arrays
// and tiles, expand.
int launch_push_tail(const PushWs& w, uint32_t nb, unsigned long long max_edges,
                     const uint32_t* degree, uint32_t* level, uint32_t* visited, uint32_t* q_out,
                     uint32_t* front_out, scaph_level_counters* counters, uint32_t next_level,
                     uint32_t n, int mode, const scaph_push_tuning* tune, hipStream_t st) {
  const uint32_t ts = tile_shift_for(max_edges, tune);
  const unsigned long long tile_cap = (max_edges + (1ull << ts) - 1) >> ts;
  SCAPH_TRY(launch_block_scan(w, nb, tile_cap, ~0ull, ts, tune, st));
  hipLaunchKernelGGL(k_mark_tiles, dim3(nb), dim3(kBlock), 0, st, w.ent_ep, w.ent_len, w.blk_cnt,
                     w.blk_coff, w.blk_eoff, w.gx, w.gep, w.tstart, w.totals);
  SCAPH_LAUNCH_CHECK();
  return launch_expand(w, max_edges, degree, level, visited, q_out, front_out, counters,
                       next_level, n, mode, ts, tune, st);
}

// Workspace of a run push: per-segment counts and offsets, the dense entry
// arrays (one entry per active Node with a non-empty out-list) and tiles.
PushWs carve_runs_ws(void* base, uint32_t num_segs, unsigned long long max_entries,
                     unsigned long long max_edges) {
  PushWs w{};
  const size_t nb = (size_t)num_segs + 1;
  const size_t m = (size_t)max_entries + 1;
  const size_t tiles = (size_t)(max_edges >> kTileSmall) + 2;
  size_t off = 0;
  char* p = static_cast<char*>(base);
  auto take = [&](size_t bytes) {
    char* r = p ? p + off : nullptr;
    off += align256(bytes);
    return r;
  };
  w.blk_cnt = reinterpret_cast<uint32_t*>(take(nb * 4));
  w.blk_edges = reinterpret_cast<uint64_t*>(take(nb * 8));
  w.blk_coff = reinterpret_cast<uint32_t*>(take(nb * 4));
  w.blk_eoff = reinterpret_cast<uint64_t*>(take(nb * 8));
  w.gx = reinterpret_cast<uint64_t*>(take(m * 8));
  w.gep = reinterpret_cast<uint64_t*>(take(m * 8));
  w.tstart = reinterpret_cast<uint32_t*>(take(tiles * 4));
  w.tslice = reinterpret_cast<uint8_t*>(take(tiles));
  w.tlist = reinterpret_cast<uint32_t*>(take(tiles * 4));
  w.tbin = reinterpret_cast<uint32_t*>(take(32 * 4));
  w.totals = reinterpret_cast<unsigned long long*>(take(64));
  carve_parts(w, nb, take);
  w.bytes = off;
  return w;
}

// Push of full page runs: the table in device memory (runs != nullptr) or
// one run passed by value.
int push_runs(const scaph_page_run* runs, uint32_t num_runs, const scaph_page_run& one,
              uint32_t num_segs, unsigned long long active_nodes,
              unsigned long long active_edges, uint32_t n, const uint32_t* degree,
              uint32_t* level, uint32_t* visited, const uint32_t* front_in, uint32_t* front_out,
              scaph_level_counters* counters, uint32_t next_level, void* workspace,
              size_t workspace_bytes, const scaph_push_tuning* tune, hipStream_t st) {
  if (num_segs == 0 || active_nodes == 0 || active_edges == 0) return 0;
  const unsigned long long small = tune && tune->small_edges ? tune->small_edges : kSmallEdges;
  if (active_edges <= small) {
    const size_t units = (size_t)num_segs * (kBitsNodes / kSmallChunk);
    hipLaunchKernelGGL(k_push_runs_small, dim3(grid_for(units, kBlock / SCAPH_WAVE, 1024)),
                       dim3(kBlock), 0, st, runs, num_runs, one, num_segs, front_in, degree, level,
                       visited, front_out, counters, next_level, n);
    SCAPH_LAUNCH_CHECK();
    return 0;
  }
  PushWs w = carve_runs_ws(workspace, num_segs, active_nodes, active_edges);
  if (w.bytes > workspace_bytes) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_runs_count, dim3(num_segs), dim3(kBlock), 0, st, runs, num_runs, one,
                     front_in, w.blk_cnt, w.blk_edges);
  SCAPH_LAUNCH_CHECK();
  const uint32_t ts = tile_shift_for(active_edges, tune);
  const unsigned long long tile_cap = (active_edges + (1ull << ts) - 1) >> ts;
  SCAPH_TRY(launch_block_scan(w, num_segs, tile_cap, active_nodes, ts, tune, st));
  hipLaunchKernelGGL(k_runs_mark, dim3(num_segs), dim3(kBlock), 0, st, runs, num_runs, one,
                     front_in, w.blk_coff, w.blk_eoff, w.gx, w.gep, w.tstart, w.totals);
  SCAPH_LAUNCH_CHECK();
  return launch_expand(w, active_edges, degree, level, visited, nullptr, front_out, counters,
                       next_level, n, OUT_BITMAP, ts, tune, st);
}

}  // namespace

// ------------------------------------------------- the no-alloc push level
// scaph_bfs_push_noalloc takes its workspace from `scratch`, a file-scope
// device array, instead of a caller buffer or an allocation.  Every device
// the library runs on holds its own copy (kNoAllocScratchBytes) until the
// process ends.  Host code never uses the array's name as an address: the
// device address of each device's copy comes from hipGetSymbolAddress, once
// per device with that device current (scaph_init, or the first call on a
// device), and is kept per device.  Every call on one device shares that
// copy, so calls that overlap on a device race; the push kernels initialise
// the parts of the workspace they read before reading them.
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// One push level from a frontier queue with the given workspace (the body of
// scaph_bfs_push and scaph_bfs_push_noalloc).
int push_queue(const scaph_graph_desc* out_graph, const uint32_t* degree, uint32_t* level,
               uint32_t* visited, const uint32_t* q_in, uint32_t n_in, unsigned long long m_in,
               uint32_t* q_out, scaph_level_counters* counters, uint32_t cur_level,
               void* workspace, size_t workspace_bytes, const scaph_push_tuning* tune,
               hipStream_t st) {
  if (!out_graph || !degree || !level || !visited || !q_out || !counters || !workspace)
    return SCAPH_STATUS_INVALID;
  if (n_in == 0) return 0;
  if (!q_in) return SCAPH_STATUS_INVALID;
  const unsigned long long coop = tune && tune->coop_edges ? tune->coop_edges : kCoopDefault;
  const unsigned long long small = tune && tune->small_edges ? tune->small_edges : kSmallEdges;
  // a wave per vertex (k_push_small) is quicker for a few short lists; the
  // grid-wide spread pays once lists average more than a wave's width
  if (n_in <= kBlock && m_in <= coop && (m_in > small || m_in > (unsigned long long)SCAPH_WAVE * n_in)) {
    // at least one block (m_in only sizes the grid; the blocks stride over
    // the frontier's true edge count)
    const unsigned long long per = (unsigned long long)kBlock * kCoopEdges;
    const unsigned long long nb = (m_in + per - 1) / per;
    const unsigned grid = nb == 0 ? 1u : (unsigned)nb;
    hipLaunchKernelGGL(k_push_coop, dim3(grid), dim3(kBlock), page_lds_bytes(out_graph->num_pages),
                       st, *out_graph, q_in, n_in, degree, level, visited, q_out, counters,
                       cur_level + 1);
    SCAPH_LAUNCH_CHECK();
    return 0;
  }
  if (m_in <= small && n_in <= 65536) {
    const unsigned grid = grid_for(n_in, kBlock / SCAPH_WAVE, 64);
    hipLaunchKernelGGL(k_push_small, dim3(grid), dim3(kBlock), page_lds_bytes(out_graph->num_pages),
                       st, *out_graph, q_in, n_in, degree, level, visited, q_out, counters,
                       cur_level + 1);
    SCAPH_LAUNCH_CHECK();
    return 0;
  }
  PushWs w = carve_push_ws(workspace, n_in, m_in);
  if (w.bytes > workspace_bytes) return SCAPH_STATUS_INVALID;
  const uint32_t nb = (n_in + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_prep_queue, dim3(nb), dim3(kBlock), page_lds_bytes(out_graph->num_pages), st,
                     *out_graph, q_in, n_in,
                     w.ent_ep, w.ent_len, w.blk_cnt, w.blk_edges);
  SCAPH_LAUNCH_CHECK();
  return launch_push_tail(w, nb, m_in, degree, level, visited, q_out, nullptr, counters,
                          cur_level + 1, out_graph->num_vertices, OUT_QUEUE, tune, st);
}

}  // namespace

// Device address of `scratch` on device `dev`, which must be current.
int scaph_noalloc_scratch_setup(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return SCAPH_STATUS_INVALID;
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return SCAPH_STATUS_INVALID;
    g_scratch_ptr[dev] = p;
  }
  if (out) *out = g_scratch_ptr[dev];
  return 0;
}

extern "C" {

int scaph_bfs_init(uint32_t* level, uint32_t* visited, const uint32_t* static_visited, uint32_t n,
                   uint32_t src, uint32_t* queue, scaph_level_counters* counters, void* stream) {
  if (!level || !visited || n == 0 || src >= n) return SCAPH_STATUS_INVALID;
  const size_t words = (((size_t)n + 63) / 64) * 2;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_bfs_init, dim3(grid_for(n, kBlock * 16, 16384)), dim3(kBlock), 0, st, level,
                     visited, static_visited, n, words);
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_bfs_set_source, dim3(1), dim3(1), 0, st, level, visited, src, queue,
                     counters);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_bitmap_from_zero_degree(uint32_t* bitmap, const uint32_t* in_degree, uint32_t n,
                                  void* stream) {
  if (!bitmap || !in_degree) return SCAPH_STATUS_INVALID;
  const size_t words = (((size_t)n + 63) / 64) * 2;
  hipLaunchKernelGGL(k_zero_degree_bitmap, dim3(grid_for(words, kBlock, 16384)), dim3(kBlock), 0,
                     as_stream(stream), bitmap, in_degree, n, words);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

size_t scaph_push_workspace_bytes(uint32_t max_entries, unsigned long long max_edges) {
  return carve_push_ws(nullptr, max_entries, max_edges).bytes;
}

int scaph_bfs_push(const scaph_graph_desc* out_graph, const uint32_t* degree, uint32_t* level,
                   uint32_t* visited, const uint32_t* q_in, uint32_t n_in,
                   unsigned long long m_in, uint32_t* q_out, scaph_level_counters* counters,
                   uint32_t cur_level, void* workspace, size_t workspace_bytes,
                   const scaph_push_tuning* tune, void* stream) {
  if (!workspace) return SCAPH_STATUS_INVALID;
  return push_queue(out_graph, degree, level, visited, q_in, n_in, m_in, q_out, counters,
                    cur_level, workspace, workspace_bytes, tune, as_stream(stream));
}

size_t scaph_bfs_push_noalloc_max_bytes(void) { return kNoAllocScratchBytes; }

int scaph_bfs_push_noalloc(const scaph_graph_desc* out_graph, const uint32_t* degree,
                           uint32_t* level, uint32_t* visited, const uint32_t* q_in, uint32_t n_in,
                           unsigned long long m_in, uint32_t* q_out,
                           scaph_level_counters* counters, uint32_t cur_level,
                           const scaph_push_tuning* tune, void* stream) {
  if (!out_graph || !degree || !level || !visited || !q_out || !counters)
    return SCAPH_STATUS_INVALID;
  if (n_in == 0) return 0;
  // the frontier's workspace must fit the file-scope array: refused before
  // anything is launched
  if (carve_push_ws(nullptr, n_in, m_in).bytes > kNoAllocScratchBytes)
    return SCAPH_STATUS_SCRATCH_TOO_SMALL;
  int dev = 0;
  SCAPH_TRY(hipGetDevice(&dev));
  void* ws = nullptr;
  const int rc = scaph_noalloc_scratch_setup(dev, &ws);
  if (rc) return rc;
  return push_queue(out_graph, degree, level, visited, q_in, n_in, m_in, q_out, counters,
                    cur_level, ws, kNoAllocScratchBytes, tune, as_stream(stream));
}

int scaph_bfs_push_page(const uint32_t* page_buf, uint32_t nodenum, unsigned long long edgenum,
                          uint32_t n, const uint32_t* degree, uint32_t* level, uint32_t* visited,
                          const uint32_t* front_in, uint32_t* front_out,
                          scaph_level_counters* counters, uint32_t cur_level, void* workspace,
                          size_t workspace_bytes, const scaph_page_hint* hint,
                          const scaph_push_tuning* tune, void* stream) {
  if (!page_buf || !degree || !level || !visited || !front_in || !front_out || !counters ||
      !workspace)
    return SCAPH_STATUS_INVALID;
  if (nodenum == 0) return 0;
  hipStream_t st = as_stream(stream);
  if (hint && hint->first_vertex != SCAPH_PAGE_COMPACTED) {
    if ((unsigned long long)hint->first_vertex + nodenum > n) return SCAPH_STATUS_INVALID;
    scaph_page_run one{};
    one.buf = reinterpret_cast<unsigned long long>(page_buf);
    one.first_vertex = hint->first_vertex;
    one.nodenum = nodenum;
    return push_runs(nullptr, 1, one, (nodenum + kBitsNodes - 1) / kBitsNodes,
                     hint->active_nodes, hint->active_edges, n, degree, level, visited, front_in,
                     front_out, counters, cur_level + 1, workspace, workspace_bytes, tune, st);
  }
  if (hint) {
    // compacted buffer: every Node is active
    if (hint->active_edges == 0) return 0;
    edgenum = hint->active_edges;
    const unsigned long long small = tune && tune->small_edges ? tune->small_edges : kSmallEdges;
    if (edgenum <= small) {
      hipLaunchKernelGGL(k_push_compact_small, dim3(grid_for(nodenum, kBlock / SCAPH_WAVE, 64)),
                         dim3(kBlock), 0, st, page_buf, nodenum, degree, level, visited, front_out,
                         counters, cur_level + 1, n);
      SCAPH_LAUNCH_CHECK();
      return 0;
    }
  }
  PushWs w = carve_push_ws(workspace, nodenum, edgenum);
  if (w.bytes > workspace_bytes) return SCAPH_STATUS_INVALID;
  const uint32_t nb = (nodenum + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_prep_page, dim3(nb), dim3(kBlock), 0, st, page_buf, nodenum, front_in,
                     w.ent_ep, w.ent_len, w.blk_cnt, w.blk_edges);
  SCAPH_LAUNCH_CHECK();
  return launch_push_tail(w, nb, edgenum, degree, level, visited, nullptr, front_out, counters,
                          cur_level + 1, n, OUT_BITMAP, tune, st);
}

size_t scaph_push_runs_workspace_bytes(uint32_t num_segs, unsigned long long active_nodes,
                                       unsigned long long active_edges) {
  return carve_runs_ws(nullptr, num_segs, active_nodes, active_edges).bytes;
}

int scaph_bfs_push_runs(const scaph_page_run* runs, uint32_t num_runs, uint32_t num_segs,
                        unsigned long long active_nodes, unsigned long long active_edges,
                        uint32_t n, const uint32_t* degree, uint32_t* level, uint32_t* visited,
                        const uint32_t* front_in, uint32_t* front_out,
                        scaph_level_counters* counters, uint32_t cur_level, void* workspace,
                        size_t workspace_bytes, const scaph_push_tuning* tune, void* stream) {
  if (!runs || !degree || !level || !visited || !front_in || !front_out || !counters ||
      num_runs == 0)
    return SCAPH_STATUS_INVALID;
  if (active_edges > 0 && !workspace) return SCAPH_STATUS_INVALID;
  const scaph_page_run none{};
  return push_runs(runs, num_runs, none, num_segs, active_nodes, active_edges, n, degree, level,
                   visited, front_in, front_out, counters, cur_level + 1, workspace,
                   workspace_bytes, tune, as_stream(stream));
}

size_t scaph_pull_workspace_bytes(unsigned long long max_edges) {
  return carve_pull_ws(nullptr, max_edges).bytes;
}

int scaph_bfs_pull_sparse(const scaph_graph_desc* in_graph, const scaph_pull_heads* heads,
                          const uint32_t* hub_list, uint32_t num_hubs, const uint32_t* degree,
                          uint32_t* level, uint32_t* visited, const uint32_t* front_in,
                          uint32_t* front_out, scaph_level_counters* counters, uint32_t cur_level,
                          unsigned long long max_edges, void* workspace, size_t workspace_bytes,
                          const scaph_pull_tuning* tune, void* stream) {
  if (!in_graph || !degree || !level || !visited || !front_in || !front_out || !counters ||
      !workspace || num_hubs > kMaxHubs || (num_hubs && !hub_list))
    return SCAPH_STATUS_INVALID;
  PullWs hw = carve_pull_ws(workspace, max_edges);
  if (hw.bytes > workspace_bytes) return SCAPH_STATUS_INVALID;
  const uint4* head = heads ? reinterpret_cast<const uint4*>(heads->head) : nullptr;
  const uint32_t* no_in = head ? heads->no_in_edges : nullptr;
  const uint32_t* live_base = no_in ? heads->live_base : nullptr;
  if (head && heads->live_base && !no_in) return SCAPH_STATUS_INVALID;
  const uint32_t indexed_heads = tune ? tune->head_indexed : 0u;
  const uint32_t outdeg_heads = tune ? (tune->head_outdeg | indexed_heads) : 0u;
  hipStream_t st = as_stream(stream);
  const uint64_t words = (((uint64_t)in_graph->num_vertices + 63) / 64) * 2;
  {
    const uint32_t hub_blocks = (num_hubs + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_pull_prologue, dim3(hub_blocks ? hub_blocks : 1u), dim3(kBlock), 0, st,
                       hw.count, hub_list, num_hubs, hub_blocks, front_in, words, hw.hub_front,
                       nullptr, 0u, 5u);
    SCAPH_LAUNCH_CHECK();
  }
  SCAPH_TRY((hipError_t)scaph_fill_u32(front_out, 0u, words, stream));
  const size_t lds = page_lds_bytes(in_graph->num_pages) + hub_lds_bytes(num_hubs);
  const unsigned threads = 8 * SCAPH_WAVE;
  const unsigned grid =
      grid_for(words, (size_t)kPullWinWords * 8,
               resident_blocks(reinterpret_cast<const void*>(k_pull_sparse<8>), threads, lds) *
                   cu_count());
  hipLaunchKernelGGL(k_pull_sparse<8>, dim3(grid), dim3(threads), lds, st, *in_graph, head, no_in,
                     live_base, outdeg_heads, indexed_heads, degree, level, visited, front_in,
                     front_out, counters, cur_level + 1, hw, num_hubs);
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pull_heavy, dim3(2048), dim3(kBlock), hub_lds_bytes(num_hubs), st,
                     hw.chunks, hw.count, degree, level, visited, front_in, front_out, counters,
                     cur_level + 1, hw.hub_front, num_hubs);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_bfs_pull(const scaph_graph_desc* in_graph, const scaph_pull_heads* heads,
                   const uint32_t* hub_list, uint32_t num_hubs,
                   const uint32_t* degree, uint32_t* level, uint32_t* visited,
                   const uint32_t* front_in, uint32_t* front_out, scaph_level_counters* counters,
                   uint32_t cur_level, unsigned long long max_edges, void* workspace,
                   size_t workspace_bytes, const scaph_pull_tuning* tune, void* stream) {
  if (!in_graph || !degree || !level || !visited || !front_in || !front_out || !counters ||
      !workspace || num_hubs > kMaxHubs || (num_hubs && !hub_list))
    return SCAPH_STATUS_INVALID;
  PullWs hw = carve_pull_ws(workspace, max_edges);
  if (hw.bytes > workspace_bytes) return SCAPH_STATUS_INVALID;
  const uint64_t words = (((uint64_t)in_graph->num_vertices + 63) / 64) * 2;
  const uint32_t shift = tune ? tune->summary_shift : 0u;
  uint64_t summary_bits = 0;
  if (shift) {
    if (shift < 5 || shift > 31) return SCAPH_STATUS_INVALID;
    summary_bits = ((uint64_t)in_graph->num_vertices + (1ull << shift) - 1) >> shift;
    if (summary_bits > kSummaryMaxBits) return SCAPH_STATUS_INVALID;
  }
  hipStream_t st = as_stream(stream);
  {
    const uint32_t hub_blocks = (num_hubs + kBlock - 1) / kBlock;
    const uint32_t blocks = hub_blocks + (uint32_t)((summary_bits + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_pull_prologue, dim3(blocks ? blocks : 1u), dim3(kBlock), 0, st, hw.count,
                       hub_list, num_hubs, hub_blocks, front_in, words, hw.hub_front, hw.summary,
                       (uint32_t)summary_bits, shift ? shift : 5u);
    SCAPH_LAUNCH_CHECK();
    if (summary_bits) hw.summary_shift = shift;
  }
  const uint32_t waves = tune && tune->block_waves ? tune->block_waves : kPullWaves;
  if (waves != 4 && waves != 8) return SCAPH_STATUS_INVALID;
  const uint32_t indexed_heads = tune ? tune->head_indexed : 0u;
  const uint32_t outdeg_heads = tune ? (tune->head_outdeg | indexed_heads) : 0u;
  const size_t lds = page_lds_bytes(in_graph->num_pages) + hub_lds_bytes(num_hubs);
  const uint4* head = heads ? reinterpret_cast<const uint4*>(heads->head) : nullptr;
  const uint32_t* no_in = head ? heads->no_in_edges : nullptr;
  const uint32_t* live_base = no_in ? heads->live_base : nullptr;
  if (head && heads->live_base && !no_in) return SCAPH_STATUS_INVALID;
  const bool prefetch = !tune || !tune->no_prefetch;
  const uint32_t simd = tune ? tune->waves_per_simd : 0u;
  if (simd != 0 && simd != 8) return SCAPH_STATUS_INVALID;
  // two phases need head records
  const uint32_t two = tune && head ? tune->two_phase : 0u;
  if (two > 2) return SCAPH_STATUS_INVALID;
  const int variant =
      (prefetch ? 1 : 0) | (waves == 8 ? 2 : 0) | (simd == 8 ? 4 : 0) | (int)(two * 8);
  static const void* const kerns[24] = {
      (const void*)k_pull<0, 4, 1, 0>, (const void*)k_pull<1, 4, 1, 0>,
      (const void*)k_pull<0, 8, 1, 0>, (const void*)k_pull<1, 8, 1, 0>,
      (const void*)k_pull<0, 4, 8, 0>, (const void*)k_pull<1, 4, 8, 0>,
      (const void*)k_pull<0, 8, 8, 0>, (const void*)k_pull<1, 8, 8, 0>,
      (const void*)k_pull<0, 4, 1, 1>, (const void*)k_pull<1, 4, 1, 1>,
      (const void*)k_pull<0, 8, 1, 1>, (const void*)k_pull<1, 8, 1, 1>,
      (const void*)k_pull<0, 4, 8, 1>, (const void*)k_pull<1, 4, 8, 1>,
      (const void*)k_pull<0, 8, 8, 1>, (const void*)k_pull<1, 8, 8, 1>,
      (const void*)k_pull<0, 4, 1, 2>, (const void*)k_pull<1, 4, 1, 2>,
      (const void*)k_pull<0, 8, 1, 2>, (const void*)k_pull<1, 8, 1, 2>,
      (const void*)k_pull<0, 4, 8, 2>, (const void*)k_pull<1, 4, 8, 2>,
      (const void*)k_pull<0, 8, 8, 2>, (const void*)k_pull<1, 8, 8, 2>};
  const void* kern = kerns[variant];
  // persistent-style: as many blocks as stay resident on every CU, each
  // walking its share of the window groups (a launch larger than what is
  // resident leaves a tail of late blocks; measured best on s26 among
  // 512..4096 256-thread blocks)
  const unsigned threads = waves * SCAPH_WAVE;
  unsigned grid = grid_for(words, (size_t)kPullWinWords * waves,
                           resident_blocks(kern, threads, lds) * cu_count());
  if (tune && tune->grid_blocks) grid = tune->grid_blocks < grid ? tune->grid_blocks : grid;
#define SCAPH_PULL_LAUNCH(P, W, M, T)                                                        \
  hipLaunchKernelGGL((k_pull<P, W, M, T>), dim3(grid), dim3(threads), lds, st, *in_graph, head, \
                     no_in, live_base, outdeg_heads, indexed_heads, degree, level, visited,       \
                     front_in, front_out, counters, cur_level + 1, hw, num_hubs)
  switch (variant) {
    case 0: SCAPH_PULL_LAUNCH(0, 4, 1, 0); break;
    case 1: SCAPH_PULL_LAUNCH(1, 4, 1, 0); break;
    case 2: SCAPH_PULL_LAUNCH(0, 8, 1, 0); break;
    case 3: SCAPH_PULL_LAUNCH(1, 8, 1, 0); break;
    case 4: SCAPH_PULL_LAUNCH(0, 4, 8, 0); break;
    case 5: SCAPH_PULL_LAUNCH(1, 4, 8, 0); break;
    case 6: SCAPH_PULL_LAUNCH(0, 8, 8, 0); break;
    case 7: SCAPH_PULL_LAUNCH(1, 8, 8, 0); break;
    case 8: SCAPH_PULL_LAUNCH(0, 4, 1, 1); break;
    case 9: SCAPH_PULL_LAUNCH(1, 4, 1, 1); break;
    case 10: SCAPH_PULL_LAUNCH(0, 8, 1, 1); break;
    case 11: SCAPH_PULL_LAUNCH(1, 8, 1, 1); break;
    case 12: SCAPH_PULL_LAUNCH(0, 4, 8, 1); break;
    case 13: SCAPH_PULL_LAUNCH(1, 4, 8, 1); break;
    case 14: SCAPH_PULL_LAUNCH(0, 8, 8, 1); break;
    case 16: SCAPH_PULL_LAUNCH(0, 4, 1, 2); break;
    case 17: SCAPH_PULL_LAUNCH(1, 4, 1, 2); break;
    case 18: SCAPH_PULL_LAUNCH(0, 8, 1, 2); break;
    case 19: SCAPH_PULL_LAUNCH(1, 8, 1, 2); break;
    case 20: SCAPH_PULL_LAUNCH(0, 4, 8, 2); break;
    case 21: SCAPH_PULL_LAUNCH(1, 4, 8, 2); break;
    case 22: SCAPH_PULL_LAUNCH(0, 8, 8, 2); break;
    case 23: SCAPH_PULL_LAUNCH(1, 8, 8, 2); break;
    default: SCAPH_PULL_LAUNCH(1, 8, 8, 1); break;
  }
#undef SCAPH_PULL_LAUNCH
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pull_heavy, dim3(2048), dim3(kBlock), hub_lds_bytes(num_hubs), st,
                     hw.chunks, hw.count, degree,
                     level, visited, front_in, front_out, counters, cur_level + 1, hw.hub_front,
                     num_hubs);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_frontier_queue_to_bitmap(const uint32_t* queue, uint32_t n_queue, uint32_t* bitmap,
                                   uint32_t n, void* stream) {
  if (!bitmap) return SCAPH_STATUS_INVALID;
  const size_t words = (((size_t)n + 63) / 64) * 2;
  int rc = scaph_fill_u32(bitmap, 0u, words, stream);
  if (rc) return rc;
  if (n_queue == 0) return 0;
  if (!queue) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_queue_to_bitmap, dim3(grid_for(n_queue, kBlock, 16384)), dim3(kBlock), 0,
                     as_stream(stream), queue, n_queue, bitmap, n);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_frontier_from_visited(const uint32_t* visited, const uint32_t* snapshot,
                                uint32_t* bitmap, uint32_t n, void* stream) {
  if (!visited || !snapshot || !bitmap) return SCAPH_STATUS_INVALID;
  if ((reinterpret_cast<uintptr_t>(visited) | reinterpret_cast<uintptr_t>(snapshot) |
       reinterpret_cast<uintptr_t>(bitmap)) & 15)
    return SCAPH_STATUS_INVALID;  // the bulk uses 16-byte accesses
  const size_t words = (((size_t)n + 63) / 64) * 2;
  hipLaunchKernelGGL(k_visited_diff, dim3(grid_for(words / 4 + 1, kBlock, 16384)), dim3(kBlock), 0,
                     as_stream(stream), visited, snapshot, bitmap, words);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_frontier_bitmap_to_queue(const uint32_t* bitmap, uint32_t n, uint32_t* queue,
                                   scaph_level_counters* counters, void* stream) {
  if (!bitmap || !queue || !counters || (reinterpret_cast<uintptr_t>(bitmap) & 15u))
    return SCAPH_STATUS_INVALID;
  const size_t words = ((size_t)n + 31) / 32;
  const size_t per_block = (size_t)kBlock * kQueueWords;
  const uint32_t grid = (uint32_t)((words + per_block - 1) / per_block);
  if (grid == 0) return 0;
  hipLaunchKernelGGL(k_bitmap_to_queue, dim3(grid), dim3(kBlock), 0, as_stream(stream), bitmap, n,
                     queue, counters);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_page_counts(const scaph_page* pages, uint32_t num_pages, const uint32_t* front,
                      const uint32_t* degree, int weighted, uint32_t* nodes, uint32_t* datas,
                      unsigned long long* totals, void* stream) {
  if (!pages || !front || !degree || !nodes || !datas || !totals) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, totals, 2u);
  SCAPH_LAUNCH_CHECK();
  if (num_pages == 0) return 0;
  hipLaunchKernelGGL(k_page_counts, dim3(grid_for(num_pages, 1, 4096)), dim3(kBlock), 0, st, pages,
                     num_pages, front, degree, weighted, nodes, datas, totals);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_page_unvisited(const scaph_page* pages, uint32_t num_pages, const uint32_t* visited,
                         const uint32_t* live, uint32_t* out, void* stream) {
  if (num_pages == 0) return 0;
  if (!pages || !visited || !out) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_page_unvisited, dim3(grid_for(num_pages, 1, 4096)), dim3(kBlock), 0,
                     as_stream(stream), pages, num_pages, visited, live, out);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_bfs_validate(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                       const uint32_t* hub_list, const uint32_t* level, uint32_t src,
                       unsigned long long* out, void* stream) {
  if (!out_graph || !level || !out) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, st, out, 2u);
  SCAPH_LAUNCH_CHECK();
  const unsigned grid = grid_for(out_graph->num_vertices, kBlock, 16384);
  hipLaunchKernelGGL(k_validate_edges, dim3(grid), dim3(kBlock), page_lds_bytes(out_graph->num_pages),
                     st, *out_graph, level, out);
  SCAPH_LAUNCH_CHECK();
  if (in_graph) {
    hipLaunchKernelGGL(k_validate_parents, dim3(grid), dim3(kBlock),
                       page_lds_bytes(in_graph->num_pages), st, *in_graph, hub_list, level, src,
                       out);
    SCAPH_LAUNCH_CHECK();
  }
  return 0;
}

}  // extern "C"
