// batch.hip — BFS over a batch of small independent graphs in one
// persistent launch, and the device generator of such a batch.  No
// reference counterpart (the reference runs one query on one graph).
//
// The batch is stored as one disjoint-union CSR: graph g owns the global
// vertex range [vbase_g, vbase_g + n_g) and its neighbour ids are global.
// Every workgroup loops: take the next graph from an atomic work counter
// (graphs are listed largest first so the long ones start early), run a
// level-synchronous BFS on it with the visited bits in LDS (n_g <= 65536:
// 8 KB) and the frontier queue in the workgroup's slice of global memory,
// and write the graph's levels.  Each level walks the frontier in chunks of
// one vertex per thread; a block-wide prefix of their degrees spreads the
// chunk's edges evenly over the threads.
#include "common.h"
#include "rmat_spec.h"

namespace {

constexpr uint32_t kMaxBatchVertices = 65536;  // per graph: the LDS visited bitmap
constexpr uint32_t kMaxBlock = 1024;
constexpr uint32_t kRun = 4;  // consecutive edges per thread per expansion round

// Exclusive block-wide prefix of v (any multiple of 64 up to 1024 threads);
// *total receives the sum.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* wtot,
                                                    uint32_t* total) {
  const uint32_t lane = lane_id(), wid = threadIdx.x / SCAPH_WAVE;
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  const uint32_t incl = wave_incl_scan(v);
  if (lane == SCAPH_WAVE - 1) wtot[wid] = incl;
  __syncthreads();
  uint32_t off = 0, all = 0;
  for (uint32_t w = 0; w < nw; ++w) {
    const uint32_t t = wtot[w];
    if (w < wid) off += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return off + incl - v;
}

__global__ void __launch_bounds__(kMaxBlock)
k_batch_bfs(const uint64_t* offsets, const uint32_t* ngh, const scaph_batch_entry* graphs,
            const uint32_t* order, uint32_t count, uint32_t* levels, uint32_t* work,
            uint32_t* qbuf, uint32_t qcap) {
  __shared__ uint32_t vis[kMaxBatchVertices / 32];
  __shared__ uint32_t c_v[kMaxBlock];
  __shared__ uint32_t c_pre[kMaxBlock];
  __shared__ uint64_t c_off[kMaxBlock];
  __shared__ uint32_t wtot[kMaxBlock / SCAPH_WAVE];
  __shared__ uint32_t s_g, s_tail;
  uint32_t* q = qbuf + (size_t)blockIdx.x * qcap;
  const uint32_t B = blockDim.x, tid = threadIdx.x;
  for (;;) {
    if (tid == 0) s_g = atomicAdd(work, 1u);
    __syncthreads();
    const uint32_t gi = s_g;
    __syncthreads();
    if (gi >= count) break;
    const scaph_batch_entry G = graphs[order[gi]];
    const uint32_t n = G.n;
    const uint64_t vb = G.vbase;
    const uint64_t* off = offsets + vb;
    uint32_t* lev = levels + vb;
    for (uint32_t i = tid; i < (n + 31) / 32; i += B) vis[i] = 0;
    for (uint32_t i = tid; i < n; i += B) lev[i] = SCAPH_UNREACHED;
    __syncthreads();
    if (tid == 0) {
      vis[G.src >> 5] |= 1u << (G.src & 31);
      lev[G.src] = 0;
      q[0] = G.src;
      s_tail = 1;
    }
    __syncthreads();
    uint32_t head = 0, tail = 1, L = 0;
    while (head < tail) {
      for (uint32_t c = head; c < tail; c += B) {
        const uint32_t i = c + tid;
        uint32_t deg = 0;
        if (i < tail) {
          const uint32_t u = q[i];
          const uint64_t o = off[u];
          deg = (uint32_t)(off[u + 1] - o);
          c_v[tid] = u;
          c_off[tid] = o;
        }
        uint32_t T;
        const uint32_t pre = block_excl_scan(deg, wtot, &T);
        c_pre[tid] = i < tail ? pre : 0xFFFFFFFFu;
        __syncthreads();
        const uint32_t rows = tail - c < B ? tail - c : B;
        // each thread takes kRun consecutive edges of the chunk: one binary
        // search over the rows' prefixes for the first, then a forward walk
        for (uint32_t e0 = 0; e0 < T; e0 += B * kRun) {
          const uint32_t ef = e0 + tid * kRun;
          uint32_t lo = 0;
          if (ef < T) {
            uint32_t hi = rows;  // last row whose prefix <= ef
            while (hi - lo > 1) {
              const uint32_t mid = (lo + hi) >> 1;
              if (c_pre[mid] <= ef) lo = mid; else hi = mid;
            }
          }
          // the run's neighbour loads go out together, then its claims
          uint32_t vr[kRun];
#pragma unroll
          for (uint32_t j = 0; j < kRun; ++j) {
            const uint32_t e = ef + j;
            vr[j] = 0;
            if (e < T) {
              while (lo + 1 < rows && c_pre[lo + 1] <= e) ++lo;
              vr[j] = (uint32_t)(ngh[c_off[lo] + (e - c_pre[lo])] - vb);
            }
          }
#pragma unroll
          for (uint32_t j = 0; j < kRun; ++j) {
            const uint32_t e = ef + j;
            bool claimed = false;
            const uint32_t v = vr[j];
            if (e < T) {
              const uint32_t bit = 1u << (v & 31);
              if (!(vis[v >> 5] & bit)) claimed = !(atomicOr(&vis[v >> 5], bit) & bit);
            }
            const uint64_t m = __ballot(claimed);
            if (m) {
              uint32_t base = 0;
              if (lane_id() == 0) base = atomicAdd(&s_tail, (uint32_t)__popcll(m));
              base = __shfl(base, 0, SCAPH_WAVE);
              if (claimed) {
                q[base + lanes_below(m)] = v;
                lev[v] = L + 1;
              }
            }
          }
        }
        __syncthreads();
      }
      head = tail;
      tail = s_tail;
      ++L;
      __syncthreads();
    }
  }
}

__global__ void k_batch_perm_keys(const scaph_batch_entry* graphs, uint32_t count,
                                  scaph_spec::PermKey* keys) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g < count) keys[g] = scaph_spec::make_perm_key(graphs[g].seed, graphs[g].scale);
}

// Edge e of the batch: graph by binary search over the edge bases, then the
// graph's own R-MAT stream (rmat_spec.h) and relabelling, as global ids.
__global__ void __launch_bounds__(256)
k_batch_rmat_keys(uint64_t* keys, const scaph_batch_entry* graphs, uint32_t count,
                  const scaph_spec::PermKey* pk, uint64_t total, uint32_t thr_a, uint32_t thr_ab,
                  uint32_t thr_abc) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
    uint32_t lo = 0, hi = count;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (graphs[mid].ebase <= i) lo = mid; else hi = mid;
    }
    const scaph_batch_entry& G = graphs[lo];
    uint32_t u, v;
    scaph_spec::rmat_edge(G.seed, i - G.ebase, G.scale, thr_a, thr_ab, thr_abc, &u, &v);
    const uint64_t n = 1ull << G.scale;
    u = (uint32_t)scaph_spec::perm_range(pk[lo], u, n);
    v = (uint32_t)scaph_spec::perm_range(pk[lo], v, n);
    keys[i] = ((G.vbase + u) << 32) | (G.vbase + v);
  }
}

inline unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  size_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace

extern "C" {

size_t scaph_batch_perm_ws_bytes(uint32_t count) {
  return (size_t)count * sizeof(scaph_spec::PermKey);
}

int scaph_batch_gen_rmat_keys(uint64_t* keys, const scaph_batch_entry* graphs, uint32_t count,
                              unsigned long long total_edges, uint32_t thr_a, uint32_t thr_ab,
                              uint32_t thr_abc, void* perm_ws, void* stream) {
  if (!keys || !graphs || !perm_ws || count == 0) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  auto* pk = static_cast<scaph_spec::PermKey*>(perm_ws);
  hipLaunchKernelGGL(k_batch_perm_keys, dim3((count + 255) / 256), dim3(256), 0, st, graphs, count,
                     pk);
  SCAPH_LAUNCH_CHECK();
  if (total_edges == 0) return 0;
  hipLaunchKernelGGL(k_batch_rmat_keys, dim3(grid_for(total_edges, 256 * 8, 65536)), dim3(256), 0,
                     st, keys, graphs, count, pk, (uint64_t)total_edges, thr_a, thr_ab, thr_abc);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_batch_bfs(const uint64_t* offsets, const uint32_t* ngh, const scaph_batch_entry* graphs,
                    const uint32_t* order, uint32_t count, uint32_t* levels, uint32_t* work,
                    uint32_t* qbuf, uint32_t qcap, uint32_t block, uint32_t grid, void* stream) {
  if (!offsets || !ngh || !graphs || !order || !levels || !work || !qbuf || grid == 0 ||
      qcap == 0 || qcap > kMaxBatchVertices || (block != 256 && block != 512 && block != 1024))
    return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  // the work counter is cleared by a kernel so the pair can be captured
  SCAPH_TRY((hipError_t)scaph_fill_u32(work, 0u, 1, stream));
  hipLaunchKernelGGL(k_batch_bfs, dim3(grid), dim3(block), 0, st, offsets, ngh, graphs, order,
                     count, levels, work, qbuf, qcap);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
