// msbfs.hip — multi-source bitset BFS: up to 256 queries traverse one graph
// together, one bit per query in 4 x 64-bit words per vertex.  No reference
// counterpart (the reference answers one query per traversal).
//
// Per level L the frontier words front[L][v] hold the queries that reached v
// at exactly level L; seen[v] holds every query that reached v so far.
//   push (small frontiers): every frontier vertex u ORs front[L][u] & ~seen[v]
//        into front[L+1][v] of each out-neighbour v (64-bit atomics), then a
//        finalise pass masks front[L+1] with ~seen and updates seen.
//   pull (large frontiers): every vertex v still missing some query ORs the
//        frontier words of its in-neighbours (only those whose bit is set in
//        the level's membership bitmap are loaded) until all missing queries
//        are found or the in-list ends; one lane per vertex, in-lists longer
//        than kPullSerial finish on the whole wave.
// The direction is a pure function of the level's counters, evaluated by
// each kernel on the device, so a level is a fixed launch sequence.
#include <hip/hip_cooperative_groups.h>

#include "common.h"

namespace {

constexpr uint32_t kBlock = 256;
constexpr uint32_t kHubTag = 0x80000000u;
constexpr uint32_t kPushChunk = 2048;  // out-edges one wave takes per frontier vertex
constexpr uint32_t kPullSerial = 8;    // in-edges one lane probes before the wave helps
constexpr uint32_t kMsProbe = 4;       // in-edges a lane probes per step
constexpr uint32_t kMaxWaves = 16;     // 1024-thread blocks

struct W4 {
  uint64_t w[4];
};

__device__ __forceinline__ W4 ld4(const uint64_t* p) {
  const ulonglong2* q = reinterpret_cast<const ulonglong2*>(p);
  const ulonglong2 a = q[0], b = q[1];
  W4 r;
  r.w[0] = a.x;
  r.w[1] = a.y;
  r.w[2] = b.x;
  r.w[3] = b.y;
  return r;
}

__device__ __forceinline__ void st4(uint64_t* p, const W4& x) {
  ulonglong2* q = reinterpret_cast<ulonglong2*>(p);
  q[0] = make_ulonglong2(x.w[0], x.w[1]);
  q[1] = make_ulonglong2(x.w[2], x.w[3]);
}

__device__ __forceinline__ bool any4(const W4& x) {
  return (x.w[0] | x.w[1] | x.w[2] | x.w[3]) != 0;
}

__device__ __forceinline__ uint32_t popc4(const W4& x) {
  return __popcll(x.w[0]) + __popcll(x.w[1]) + __popcll(x.w[2]) + __popcll(x.w[3]);
}

__device__ __forceinline__ W4 query_mask(uint32_t q) {
  W4 m;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int lo = 64 * k;
    m.w[k] = (int)q >= lo + 64 ? ~0ull : ((int)q <= lo ? 0ull : ((1ull << (q - lo)) - 1ull));
  }
  return m;
}

__device__ __forceinline__ bool use_push(const scaph_ms_counters* c, uint32_t n,
                                         const scaph_ms_params& p) {
  return c->n_active * (unsigned long long)p.push_div < (unsigned long long)n &&
         c->m_active * (unsigned long long)p.push_mdiv < p.m;
}

__device__ __forceinline__ uint64_t* front_of(const scaph_ms_state& s, uint32_t L) {
  return s.front + (size_t)L * s.n * 4;
}
__device__ __forceinline__ uint64_t* act_of(const scaph_ms_state& s, uint32_t L) {
  return s.act + (size_t)L * s.act_words;
}

// Adds per-thread counts into ctr (one atomic per block and counter).
__device__ __forceinline__ void flush_counts(scaph_ms_counters* ctr, unsigned long long bits,
                                             unsigned long long nact,
                                             unsigned long long mact) {
  __shared__ unsigned long long red[3][kMaxWaves];
  const uint32_t wid = threadIdx.x / SCAPH_WAVE, nw = blockDim.x / SCAPH_WAVE;
  bits = wave_sum(bits);
  nact = wave_sum(nact);
  mact = wave_sum(mact);
  if (lane_id() == 0) {
    red[0][wid] = bits;
    red[1][wid] = nact;
    red[2][wid] = mact;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    unsigned long long t = 0;
    for (uint32_t w = 0; w < nw; ++w) t += red[threadIdx.x][w];
    if (t) {
      unsigned long long* dst =
          threadIdx.x == 0 ? &ctr->bits : (threadIdx.x == 1 ? &ctr->n_active : &ctr->m_active);
      atomicAdd(dst, t);
    }
  }
}

__global__ void k_ms_zero(uint64_t* p, size_t words) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += stride) p[i] = 0;
}

__global__ void k_ms_sources(scaph_ms_state s, const uint32_t* sources, uint32_t count,
                             const uint32_t* degree) {
  const uint32_t q = threadIdx.x;
  if (q >= count) return;
  const uint32_t v = sources[q];
  scaph_ms_counters* c = &s.ctr[0];
  if (v >= s.n) {
    atomicAdd(&c->overflow, 1ull);
    return;
  }
  const uint64_t bit = 1ull << (q & 63);
  atomicOr(reinterpret_cast<unsigned long long*>(s.front + (size_t)v * 4 + (q >> 6)), bit);
  atomicOr(reinterpret_cast<unsigned long long*>(s.seen + (size_t)v * 4 + (q >> 6)), bit);
  const uint64_t vb = 1ull << (v & 63);
  const uint64_t old = atomicOr(reinterpret_cast<unsigned long long*>(s.act + (v >> 6)), vb);
  atomicAdd(&c->bits, 1ull);
  if (!(old & vb)) {
    atomicAdd(&c->n_active, 1ull);
    atomicAdd(&c->m_active, (unsigned long long)degree[v]);
  }
}

// Each phase of a level is a device function over blocks bid of nblk, run
// by its own launch (the per-level sequence) or by every block of the
// persistent kernel between grid barriers.

// Push preparation: clear level L+1's words and list the frontier.
__device__ __forceinline__ void ms_prep(const scaph_ms_state& s, uint32_t L, uint32_t bid,
                                        uint32_t nblk) {
  const size_t stride = (size_t)nblk * blockDim.x;
  const size_t tid = (size_t)bid * blockDim.x + threadIdx.x;
  ulonglong2* fn = reinterpret_cast<ulonglong2*>(front_of(s, L + 1));
  for (size_t i = tid; i < (size_t)s.n * 2; i += stride) fn[i] = make_ulonglong2(0, 0);
  uint64_t* an = act_of(s, L + 1);
  for (size_t i = tid; i < s.act_words; i += stride) an[i] = 0;
  // frontier list: one membership word per lane, wave-aggregated appends
  const uint64_t* a = act_of(s, L);
  const size_t words_pad = (s.act_words + SCAPH_WAVE - 1) / SCAPH_WAVE * SCAPH_WAVE;
  for (size_t i = tid; i < words_pad; i += stride) {
    const uint64_t bits = i < s.act_words ? a[i] : 0ull;
    const uint32_t cnt = __popcll(bits);
    const uint32_t incl = wave_incl_scan(cnt);
    const uint32_t total = __shfl(incl, SCAPH_WAVE - 1, SCAPH_WAVE);
    unsigned long long base = 0;
    if (lane_id() == SCAPH_WAVE - 1 && total)
      base = atomicAdd(&s.ctr[L].q_len, (unsigned long long)total);
    base = __shfl(base, SCAPH_WAVE - 1, SCAPH_WAVE);
    uint32_t o = (uint32_t)(base + incl - cnt);
    uint64_t b = bits;
    while (b) {
      const uint32_t t = __ffsll((long long)b) - 1;
      b &= b - 1;
      s.queue[o++] = (uint32_t)(i * 64 + t);
    }
  }
}

__global__ void __launch_bounds__(kBlock)
k_ms_prep(scaph_ms_state s, uint32_t L, scaph_ms_params prm) {
  const scaph_ms_counters* c = &s.ctr[L];
  if (c->n_active == 0 || !use_push(c, s.n, prm)) return;
  ms_prep(s, L, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void push_edges(const uint32_t* edges, uint32_t begin, uint32_t end,
                                           const W4& f, const uint64_t* seen, uint64_t* fn) {
  for (uint32_t j = begin + lane_id(); j < end; j += SCAPH_WAVE) {
    const uint32_t v = edges[j];
    const uint64_t* sv = seen + (size_t)v * 4;
    unsigned long long* fv = reinterpret_cast<unsigned long long*>(fn + (size_t)v * 4);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      if (f.w[k]) {
        const uint64_t w = f.w[k] & ~sv[k];
        if (w) atomicOr(fv + k, w);
      }
    }
  }
}

__device__ __forceinline__ void ms_push(const PageTab& t, const scaph_graph_desc& g,
                                        const scaph_ms_state& s, uint32_t L, uint32_t bid,
                                        uint32_t nblk) {
  const unsigned long long q_len = s.ctr[L].q_len;
  const uint64_t* fr = front_of(s, L);
  uint64_t* fn = front_of(s, L + 1);
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  const unsigned long long waves = (unsigned long long)nblk * nw;
  for (unsigned long long i = (unsigned long long)bid * nw + threadIdx.x / SCAPH_WAVE;
       i < q_len; i += waves) {
    const uint32_t u = s.queue[i];
    const W4 f = ld4(fr + (size_t)u * 4);
    const NodeRef nd = load_node(t, g, u);
    push_edges(nd.edges, 0, nd.len < kPushChunk ? nd.len : kPushChunk, f, s.seen, fn);
    if (nd.len > kPushChunk) {
      const uint32_t chunks = (nd.len - 1) / kPushChunk;  // beyond the first
      unsigned long long base = 0;
      if (lane_id() == 0) base = atomicAdd(&s.ctr[L].heavy_len, (unsigned long long)chunks);
      base = __shfl(base, 0, SCAPH_WAVE);
      for (uint32_t k = lane_id(); k < chunks; k += SCAPH_WAVE) {
        if (base + k < s.heavy_cap)
          s.heavy[base + k] = ((uint64_t)u << 32) | (uint64_t)(kPushChunk * (k + 1));
        else
          atomicAdd(&s.ctr[L].overflow, 1ull);
      }
    }
  }
}

__global__ void __launch_bounds__(kBlock)
k_ms_push(scaph_graph_desc g, scaph_ms_state s, uint32_t L, scaph_ms_params prm) {
  extern __shared__ __align__(16) unsigned char lds[];
  const scaph_ms_counters* c = &s.ctr[L];
  if (c->n_active == 0 || !use_push(c, s.n, prm)) return;
  const PageTab t = stage_pages(g, lds);
  __syncthreads();
  ms_push(t, g, s, L, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void ms_push_heavy(const PageTab& t, const scaph_graph_desc& g,
                                              const scaph_ms_state& s, uint32_t L, uint32_t bid,
                                              uint32_t nblk) {
  unsigned long long hl = s.ctr[L].heavy_len;
  if (hl > s.heavy_cap) hl = s.heavy_cap;
  const uint64_t* fr = front_of(s, L);
  uint64_t* fn = front_of(s, L + 1);
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  const unsigned long long waves = (unsigned long long)nblk * nw;
  for (unsigned long long i = (unsigned long long)bid * nw + threadIdx.x / SCAPH_WAVE;
       i < hl; i += waves) {
    const uint64_t item = s.heavy[i];
    const uint32_t u = (uint32_t)(item >> 32), start = (uint32_t)item;
    const W4 f = ld4(fr + (size_t)u * 4);
    const NodeRef nd = load_node(t, g, u);
    const uint32_t end = nd.len - start < kPushChunk ? nd.len : start + kPushChunk;
    push_edges(nd.edges, start, end, f, s.seen, fn);
  }
}

__global__ void __launch_bounds__(kBlock)
k_ms_push_heavy(scaph_graph_desc g, scaph_ms_state s, uint32_t L, scaph_ms_params prm) {
  extern __shared__ __align__(16) unsigned char lds[];
  const scaph_ms_counters* c = &s.ctr[L];
  if (c->n_active == 0 || !use_push(c, s.n, prm) || c->heavy_len == 0) return;
  const PageTab t = stage_pages(g, lds);
  __syncthreads();
  ms_push_heavy(t, g, s, L, blockIdx.x, gridDim.x);
}

// After push: front[L+1] &= ~seen, seen |= front[L+1], membership bits and
// counters of level L+1.
__device__ __forceinline__ void ms_finalize(const scaph_ms_state& s, uint32_t L,
                                            const uint32_t* degree, uint32_t bid, uint32_t nblk) {
  uint64_t* fn = front_of(s, L + 1);
  uint64_t* an = act_of(s, L + 1);
  unsigned long long bits = 0, nact = 0, mact = 0;
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  const size_t wstride = (size_t)nblk * nw * SCAPH_WAVE;
  for (size_t base = ((size_t)bid * nw + threadIdx.x / SCAPH_WAVE) * SCAPH_WAVE;
       base < s.n; base += wstride) {
    const size_t v = base + lane_id();
    bool hit = false;
    if (v < s.n) {
      const W4 x = ld4(fn + v * 4);
      if (any4(x)) {
        W4 sv = ld4(s.seen + v * 4);
        W4 nx;
#pragma unroll
        for (int k = 0; k < 4; ++k) nx.w[k] = x.w[k] & ~sv.w[k];
        st4(fn + v * 4, nx);
        if (any4(nx)) {
#pragma unroll
          for (int k = 0; k < 4; ++k) sv.w[k] |= nx.w[k];
          st4(s.seen + v * 4, sv);
          hit = true;
          bits += popc4(nx);
          nact += 1;
          mact += degree[v];
        }
      }
    }
    const uint64_t b = __ballot(hit);
    if (lane_id() == 0) an[base >> 6] = b;
  }
  flush_counts(&s.ctr[L + 1], bits, nact, mact);
}

__global__ void __launch_bounds__(kBlock)
k_ms_finalize(scaph_ms_state s, uint32_t L, scaph_ms_params prm, const uint32_t* degree) {
  const scaph_ms_counters* c = &s.ctr[L];
  if (c->n_active == 0 || !use_push(c, s.n, prm)) return;
  ms_finalize(s, L, degree, blockIdx.x, gridDim.x);
}

// The frontier words of in-list entry e at the level being pulled, or
// nullptr when e is not in that frontier: a hub (kHubTag | hub index) is
// tested in the LDS copy of the hub bits and read from the hubs' compact
// copy of their words, any other vertex through the level's bitmap.
__device__ __forceinline__ const uint64_t* ms_front_of(uint32_t e, const uint32_t* s_hub,
                                                       const uint64_t* hub_front,
                                                       const uint64_t* act, const uint64_t* fr) {
  if (e & kHubTag) {
    const uint32_t h = e & ~kHubTag;
    return ((s_hub[h >> 5] >> (h & 31)) & 1u) ? hub_front + (size_t)h * 4 : nullptr;
  }
  return ((act[e >> 6] >> (e & 63)) & 1ull) ? fr + (size_t)e * 4 : nullptr;
}

// Pull levels only: the frontier words of the hubs (hub i = hub_list[i])
// copied to hub_front[i] where hub i is in the level's frontier, and those
// memberships as bits (hub_bits, whole words; the pull stages them in LDS).
// (256-thread blocks.)
__device__ __forceinline__ void ms_hub_gather(const scaph_ms_state& s, uint32_t L, uint32_t nh,
                                              const uint32_t* hub_list, uint32_t bid,
                                              uint32_t nblk) {
  const uint64_t* fr = front_of(s, L);
  const uint64_t* act = act_of(s, L);
  const uint32_t lane = lane_id();
  // base is uniform per block, and a multiple of 256 below nh <= the hub
  // capacity (a multiple of 256), so every word written is in range
  for (uint32_t base = bid * kBlock; base < nh; base += nblk * kBlock) {
    const uint32_t i = base + threadIdx.x;
    bool a = false;
    if (i < nh) {
      const uint32_t u = hub_list[i];
      a = (act[u >> 6] >> (u & 63)) & 1ull;
      if (a) st4(s.hub_front + (size_t)i * 4, ld4(fr + (size_t)u * 4));
    }
    const uint64_t b = __ballot(a);
    const uint32_t w = (i - lane) / 32;
    if (lane == 0) s.hub_bits[w] = (uint32_t)b;
    if (lane == 32) s.hub_bits[w + 1] = (uint32_t)(b >> 32);
  }
}

__global__ void __launch_bounds__(kBlock)
k_ms_hub_gather(scaph_ms_state s, uint32_t L, scaph_ms_params prm, const uint32_t* hub_list) {
  const scaph_ms_counters* c = &s.ctr[L];
  if (c->n_active == 0 || use_push(c, s.n, prm)) return;
  ms_hub_gather(s, L, prm.num_hubs, hub_list, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void stage_hub_bits(uint32_t* s_hub, const uint32_t* hub_bits,
                                               uint32_t nh) {
  for (uint32_t i = threadIdx.x; i < (nh + 31) / 32; i += blockDim.x) s_hub[i] = hub_bits[i];
}

// LDS bytes of the hub bits, behind the page table.
inline size_t ms_hub_lds_bytes(uint32_t num_hubs) { return ((size_t)(num_hubs + 127) / 128) * 16; }

__device__ __forceinline__ uint64_t wave_or64(uint64_t x) {
#pragma unroll
  for (int d = SCAPH_WAVE / 2; d >= 1; d >>= 1) x |= __shfl_xor(x, d, SCAPH_WAVE);
  return x;
}

__device__ __forceinline__ void ms_pull(const PageTab& t, const scaph_graph_desc& g,
                                        const uint32_t* s_hub, const scaph_ms_state& s,
                                        uint32_t L, const scaph_ms_params& prm,
                                        const uint32_t* degree, uint32_t bid, uint32_t nblk) {
  const uint64_t* fr = front_of(s, L);
  const uint64_t* act = act_of(s, L);
  uint64_t* fn = front_of(s, L + 1);
  uint64_t* an = act_of(s, L + 1);
  const W4 qm = query_mask(prm.num_queries);
  unsigned long long bits = 0, nact = 0, mact = 0;
  const uint32_t nw = blockDim.x / SCAPH_WAVE;
  const size_t wstride = (size_t)nblk * nw * SCAPH_WAVE;
  for (size_t base = ((size_t)bid * nw + threadIdx.x / SCAPH_WAVE) * SCAPH_WAVE;
       base < s.n; base += wstride) {
    const size_t v = base + lane_id();
    const bool valid = v < s.n;
    W4 sv = {{~0ull, ~0ull, ~0ull, ~0ull}}, need, acc = {{0, 0, 0, 0}};
    if (valid) sv = ld4(s.seen + v * 4);
#pragma unroll
    for (int k = 0; k < 4; ++k) need.w[k] = ~sv.w[k] & qm.w[k];
    const bool want = valid && any4(need);
    const uint32_t* edges = nullptr;
    uint32_t len = 0;
    bool done = !want;
    if (want) {
      const NodeRef nd = load_node(t, g, (uint32_t)v);
      edges = nd.edges;
      len = nd.len;
      const uint32_t lim = len < kPullSerial ? len : kPullSerial;
      // kMsProbe in-edges per step, each stage's loads independent of one
      // another (entries, hub ids, membership bits, frontier words), so a
      // lane keeps several misses in flight; the test for "every missing
      // query found" runs once per step
      for (uint32_t i = 0; i < lim; i += kMsProbe) {
        uint32_t u[kMsProbe];
        const uint64_t* fp[kMsProbe];
#pragma unroll
        for (uint32_t j = 0; j < kMsProbe; ++j) u[j] = i + j < lim ? edges[i + j] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < kMsProbe; ++j)
          fp[j] = i + j < lim ? ms_front_of(u[j], s_hub, s.hub_front, act, fr) : nullptr;
#pragma unroll
        for (uint32_t j = 0; j < kMsProbe; ++j) {
          if (fp[j]) {
            const W4 f = ld4(fp[j]);
#pragma unroll
            for (int k = 0; k < 4; ++k) acc.w[k] |= f.w[k] & need.w[k];
          }
        }
        if (acc.w[0] == need.w[0] && acc.w[1] == need.w[1] && acc.w[2] == need.w[2] &&
            acc.w[3] == need.w[3]) {
          done = true;
          break;
        }
      }
    }
    // long in-lists of this window, one at a time on the whole wave
    uint64_t hm = __ballot(!done && len > kPullSerial);
    while (hm) {
      const int l = __ffsll((long long)hm) - 1;
      hm &= hm - 1;
      const uint64_t ep = wave_bcast64(reinterpret_cast<uint64_t>(edges), l);
      const uint32_t* el = reinterpret_cast<const uint32_t*>(ep);
      const uint32_t ll = __shfl(len, l, SCAPH_WAVE);
      W4 nl, al, a;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        nl.w[k] = wave_bcast64(need.w[k], l);
        al.w[k] = wave_bcast64(acc.w[k], l);
        a.w[k] = 0;
      }
      for (uint32_t i0 = kPullSerial; i0 < ll; i0 += 8 * SCAPH_WAVE) {
        const uint32_t i1 = i0 + 8 * SCAPH_WAVE < ll ? i0 + 8 * SCAPH_WAVE : ll;
        for (uint32_t i = i0 + lane_id(); i < i1; i += SCAPH_WAVE) {
          const uint64_t* p = ms_front_of(el[i], s_hub, s.hub_front, act, fr);
          if (p) {
            const W4 f = ld4(p);
#pragma unroll
            for (int k = 0; k < 4; ++k) a.w[k] |= f.w[k] & nl.w[k];
          }
        }
        bool full = true;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          a.w[k] = wave_or64(a.w[k]);
          full &= (a.w[k] | al.w[k]) == nl.w[k];
        }
        if (full) break;
      }
      // every lane holds the wave-wide OR after the last round
      if ((int)lane_id() == l) {
#pragma unroll
        for (int k = 0; k < 4; ++k) acc.w[k] |= a.w[k];
      }
    }
    bool hit = false;
    if (valid) {
      st4(fn + v * 4, acc);
      if (any4(acc)) {
#pragma unroll
        for (int k = 0; k < 4; ++k) sv.w[k] |= acc.w[k];
        st4(s.seen + v * 4, sv);
        hit = true;
        bits += popc4(acc);
        nact += 1;
        mact += degree[v];
      }
    }
    const uint64_t b = __ballot(hit);
    if (lane_id() == 0) an[base >> 6] = b;
  }
  flush_counts(&s.ctr[L + 1], bits, nact, mact);
}

__global__ void __launch_bounds__(1024)
k_ms_pull(scaph_graph_desc g, scaph_ms_state s, uint32_t L, scaph_ms_params prm,
          const uint32_t* degree) {
  extern __shared__ __align__(16) unsigned char lds[];
  const scaph_ms_counters* c = &s.ctr[L];
  if (c->n_active == 0 || use_push(c, s.n, prm)) return;
  const uint32_t np = g.num_pages < SCAPH_LDS_PAGES ? g.num_pages : SCAPH_LDS_PAGES;
  uint32_t* const s_hub = reinterpret_cast<uint32_t*>(lds + (size_t)np * 16);
  stage_hub_bits(s_hub, s.hub_bits, prm.num_hubs);
  const PageTab t = stage_pages(g, lds);
  __syncthreads();
  ms_pull(t, g, s_hub, s, L, prm, degree, blockIdx.x, gridDim.x);
}

// Device bytes of a page table staged in LDS (page_lds_bytes on the host).
__device__ __forceinline__ size_t page_lds_bytes_dev(uint32_t num_pages) {
  return (size_t)(num_pages < SCAPH_LDS_PAGES ? num_pages : SCAPH_LDS_PAGES) * 16;
}

__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Every level of one multi-source traversal in one cooperative launch of
// 256-thread blocks, all resident: the phases of a level (push: prep, push,
// heavy push, finalize; pull: hub gather, pull) are separated by grid
// barriers, and each block takes the level's direction and the loop's end
// from the counters, read at device scope after a barrier, so every block
// makes the same decisions and reaches the same barriers.  Levels
// first .. max_levels - 2 at most: a traversal still active after the last
// level buffer stops there (the host reports it).
__global__ void __launch_bounds__(kBlock)
k_ms_persistent(scaph_graph_desc gout, scaph_graph_desc gin, const uint32_t* hub_list,
                const uint32_t* degree, scaph_ms_state s, scaph_ms_params prm,
                uint32_t first_level) {
  extern __shared__ __align__(16) unsigned char lds[];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const PageTab tout = stage_pages(gout, lds);
  unsigned char* lin = lds + page_lds_bytes_dev(gout.num_pages);
  const PageTab tin = stage_pages(gin, lin);
  uint32_t* const s_hub = reinterpret_cast<uint32_t*>(lin + page_lds_bytes_dev(gin.num_pages));
  __syncthreads();
  const uint32_t bid = blockIdx.x, nblk = gridDim.x;
  for (uint32_t L = first_level; L + 1 < s.max_levels; ++L) {
    const scaph_ms_counters* c = &s.ctr[L];
    const unsigned long long na = ld_agent(&c->n_active);
    if (na == 0) break;
    const bool push = na * (unsigned long long)prm.push_div < (unsigned long long)s.n &&
                      ld_agent(&c->m_active) * (unsigned long long)prm.push_mdiv < prm.m;
    if (push) {
      ms_prep(s, L, bid, nblk);
      grid.sync();
      ms_push(tout, gout, s, L, bid, nblk);
      grid.sync();
      if (ld_agent(&c->heavy_len)) {
        ms_push_heavy(tout, gout, s, L, bid, nblk);
        grid.sync();
      }
      ms_finalize(s, L, degree, bid, nblk);
    } else {
      if (prm.num_hubs) {
        ms_hub_gather(s, L, prm.num_hubs, hub_list, bid, nblk);
        grid.sync();
        stage_hub_bits(s_hub, s.hub_bits, prm.num_hubs);
        __syncthreads();
      }
      ms_pull(tin, gin, s_hub, s, L, prm, degree, bid, nblk);
    }
    grid.sync();
  }
}

__global__ void __launch_bounds__(kBlock)
k_ms_levels_u32(scaph_ms_state s, uint32_t depth, uint32_t q, uint32_t* out) {
  const uint32_t k = q >> 6, b = q & 63;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < s.n; v += stride) {
    uint32_t lvl = SCAPH_UNREACHED;
    for (uint32_t L = 0; L < depth; ++L) {
      if ((front_of(s, L)[v * 4 + k] >> b) & 1ull) {
        lvl = L;
        break;
      }
    }
    out[v] = lvl;
  }
}

// Byte masks of the 16 nibble values: bit i of the nibble -> byte i.
__device__ __forceinline__ uint32_t nibble_bytes(uint32_t nib) {
  return (nib & 1u ? 0x000000FFu : 0u) | (nib & 2u ? 0x0000FF00u : 0u) |
         (nib & 4u ? 0x00FF0000u : 0u) | (nib & 8u ? 0xFF000000u : 0u);
}

__global__ void __launch_bounds__(kBlock)
k_ms_levels_u8(scaph_ms_state s, uint32_t depth, uint8_t* out) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = (size_t)blockIdx.x * blockDim.x + threadIdx.x; v < s.n; v += stride) {
    uint32_t r[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) r[i] = 0xFFFFFFFFu;
    for (uint32_t L = 0; L < depth; ++L) {
      const W4 f = ld4(front_of(s, L) + v * 4);
      if (!any4(f)) continue;
      const uint32_t lv = L * 0x01010101u;
#pragma unroll
      for (int k = 0; k < 4; ++k) {
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const uint32_t m = nibble_bytes((uint32_t)(f.w[k] >> (4 * j)) & 0xFu);
          r[16 * k + j] = (r[16 * k + j] & ~m) | (lv & m);
        }
      }
    }
    uint4* o = reinterpret_cast<uint4*>(out + v * 256);
#pragma unroll
    for (int i = 0; i < 16; ++i) o[i] = make_uint4(r[4 * i], r[4 * i + 1], r[4 * i + 2], r[4 * i + 3]);
  }
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

inline unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  size_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

}  // namespace

extern "C" {

size_t scaph_ms_state_bytes(uint32_t n, unsigned long long m, uint32_t max_levels) {
  const size_t act_words = ((size_t)n + 63) / 64;
  const size_t heavy_cap = m / kPushChunk + 1;
  return align256((size_t)n * 32) + align256((size_t)max_levels * n * 32) +
         align256((size_t)max_levels * act_words * 8) +
         align256(((size_t)max_levels + 1) * sizeof(scaph_ms_counters)) +
         align256((size_t)n * 4 + 4) + align256(heavy_cap * 8) +
         align256((size_t)SCAPH_MS_MAX_HUBS * 32) + align256(SCAPH_MS_MAX_HUBS / 8);
}

int scaph_ms_state_carve(void* base, uint32_t n, unsigned long long m, uint32_t max_levels,
                         scaph_ms_state* st) {
  if (!base || !st || n == 0 || max_levels < 2) return SCAPH_STATUS_INVALID;
  char* p = static_cast<char*>(base);
  st->n = n;
  st->max_levels = max_levels;
  st->act_words = ((size_t)n + 63) / 64;
  st->heavy_cap = m / kPushChunk + 1;
  st->seen = reinterpret_cast<uint64_t*>(p);
  p += align256((size_t)n * 32);
  st->front = reinterpret_cast<uint64_t*>(p);
  p += align256((size_t)max_levels * n * 32);
  st->act = reinterpret_cast<uint64_t*>(p);
  p += align256((size_t)max_levels * st->act_words * 8);
  st->ctr = reinterpret_cast<scaph_ms_counters*>(p);
  p += align256(((size_t)max_levels + 1) * sizeof(scaph_ms_counters));
  st->queue = reinterpret_cast<uint32_t*>(p);
  p += align256((size_t)n * 4 + 4);
  st->heavy = reinterpret_cast<uint64_t*>(p);
  p += align256(st->heavy_cap * 8);
  st->hub_front = reinterpret_cast<uint64_t*>(p);
  p += align256((size_t)SCAPH_MS_MAX_HUBS * 32);
  st->hub_bits = reinterpret_cast<uint32_t*>(p);
  return 0;
}

int scaph_ms_init(const scaph_ms_state* st, const uint32_t* sources, uint32_t count,
                  const uint32_t* degree, void* stream) {
  if (!st || !sources || !degree || count == 0 || count > SCAPH_MS_MAX_QUERIES)
    return SCAPH_STATUS_INVALID;
  hipStream_t sm = as_stream(stream);
  const scaph_ms_state s = *st;
  auto zero = [&](uint64_t* p, size_t words) -> int {
    hipLaunchKernelGGL(k_ms_zero, dim3(grid_for(words, kBlock * 4, 8192)), dim3(kBlock), 0, sm,
                       p, words);
    SCAPH_LAUNCH_CHECK();
    return 0;
  };
  SCAPH_TRY((hipError_t)zero(s.seen, (size_t)s.n * 4));
  SCAPH_TRY((hipError_t)zero(s.front, (size_t)s.n * 4));
  SCAPH_TRY((hipError_t)zero(s.act, s.act_words));
  SCAPH_TRY((hipError_t)zero(reinterpret_cast<uint64_t*>(s.ctr),
                             ((size_t)s.max_levels + 1) * sizeof(scaph_ms_counters) / 8));
  hipLaunchKernelGGL(k_ms_sources, dim3(1), dim3(SCAPH_MS_MAX_QUERIES), 0, sm, s, sources, count,
                     degree);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

static bool ms_args_ok(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                       const uint32_t* hub_list, const uint32_t* degree, const scaph_ms_state* st,
                       uint32_t level, const scaph_ms_params* prm) {
  return out_graph && in_graph && degree && st && prm && level + 1 < st->max_levels &&
         prm->num_queries != 0 && prm->num_queries <= SCAPH_MS_MAX_QUERIES &&
         prm->push_div != 0 && prm->push_mdiv != 0 && prm->num_hubs <= SCAPH_MS_MAX_HUBS &&
         (prm->num_hubs == 0 || hub_list);
}

int scaph_ms_pull_blocks_per_cu(uint32_t pull_block, uint32_t in_pages, uint32_t num_hubs,
                                 uint32_t* out) {
  if (!out || (pull_block != 256 && pull_block != 512 && pull_block != 1024))
    return SCAPH_STATUS_INVALID;
  int nb = 0;
  SCAPH_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &nb, reinterpret_cast<const void*>(k_ms_pull), (int)pull_block,
      page_lds_bytes(in_pages) + ms_hub_lds_bytes(num_hubs)));
  *out = nb > 0 ? (uint32_t)nb : 1u;
  return 0;
}

int scaph_ms_run_persistent(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                            const uint32_t* hub_list, const uint32_t* degree,
                            const scaph_ms_state* st, uint32_t first_level,
                            const scaph_ms_params* prm, void* stream) {
  if (!ms_args_ok(out_graph, in_graph, hub_list, degree, st, first_level, prm))
    return SCAPH_STATUS_INVALID;
  const size_t lds = page_lds_bytes(out_graph->num_pages) + page_lds_bytes(in_graph->num_pages) +
                     ms_hub_lds_bytes(prm->num_hubs);
  int dev = 0, cus = 0, per_cu = 0;
  SCAPH_TRY(hipGetDevice(&dev));
  SCAPH_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // every block must be resident for the grid barriers: the grid is what the
  // occupancy calculation allows, and the cooperative launch refuses more
  SCAPH_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(k_ms_persistent), (int)kBlock, lds));
  if (per_cu < 1 || cus < 1) return SCAPH_STATUS_INVALID;
  scaph_graph_desc go = *out_graph, gi = *in_graph;
  scaph_ms_state s = *st;
  scaph_ms_params p = *prm;
  uint32_t first = first_level;
  void* args[] = {&go, &gi, const_cast<uint32_t**>(&hub_list), const_cast<uint32_t**>(&degree),
                  &s, &p, &first};
  SCAPH_TRY(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(k_ms_persistent),
                                       dim3((unsigned)(per_cu * cus)), dim3(kBlock), args,
                                       (unsigned)lds, as_stream(stream)));
  return 0;
}

int scaph_ms_level(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                   const uint32_t* hub_list, const uint32_t* degree, const scaph_ms_state* st,
                   uint32_t level, const scaph_ms_params* prm, void* stream) {
  if (!ms_args_ok(out_graph, in_graph, hub_list, degree, st, level, prm))
    return SCAPH_STATUS_INVALID;
  const uint32_t pb = prm->pull_block;
  if (pb != 256 && pb != 512 && pb != 1024) return SCAPH_STATUS_INVALID;
  hipStream_t sm = as_stream(stream);
  const scaph_ms_state s = *st;
  const scaph_ms_params p = *prm;
  const unsigned g_push = 2048;
  hipLaunchKernelGGL(k_ms_prep, dim3(grid_for((size_t)s.n * 2, kBlock * 8, 8192)), dim3(kBlock),
                     0, sm, s, level, p);
  SCAPH_LAUNCH_CHECK();
  const size_t lo = page_lds_bytes(out_graph->num_pages);
  hipLaunchKernelGGL(k_ms_push, dim3(g_push), dim3(kBlock), lo, sm, *out_graph, s, level, p);
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ms_push_heavy, dim3(g_push), dim3(kBlock), lo, sm, *out_graph, s, level, p);
  SCAPH_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_ms_finalize, dim3(grid_for(s.n, kBlock, 8192)), dim3(kBlock), 0, sm, s,
                     level, p, degree);
  SCAPH_LAUNCH_CHECK();
  if (p.num_hubs) {
    hipLaunchKernelGGL(k_ms_hub_gather, dim3(grid_for(p.num_hubs, kBlock, 1024)), dim3(kBlock), 0,
                       sm, s, level, p, hub_list);
    SCAPH_LAUNCH_CHECK();
  }
  const size_t li = page_lds_bytes(in_graph->num_pages) + ms_hub_lds_bytes(p.num_hubs);
  unsigned pgrid = grid_for(s.n, pb, 16384);
  if (p.pull_grid && p.pull_grid < pgrid) pgrid = p.pull_grid;
  hipLaunchKernelGGL(k_ms_pull, dim3(pgrid), dim3(pb), li, sm, *in_graph, s, level, p, degree);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_ms_levels_u32(const scaph_ms_state* st, uint32_t depth, uint32_t query, uint32_t* out,
                        void* stream) {
  if (!st || !out || query >= SCAPH_MS_MAX_QUERIES || depth > st->max_levels)
    return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_ms_levels_u32, dim3(grid_for(st->n, kBlock, 8192)), dim3(kBlock), 0,
                     as_stream(stream), *st, depth, query, out);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_ms_levels_u8(const scaph_ms_state* st, uint32_t depth, uint8_t* out, void* stream) {
  if (!st || !out || depth > st->max_levels || depth >= 255) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_ms_levels_u8, dim3(grid_for(st->n, kBlock, 8192)), dim3(kBlock), 0,
                     as_stream(stream), *st, depth, out);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

}  // extern "C"
