// common.h — device-side helpers shared by the gfx950 kernels.
// Wave64 throughout: a wavefront is 64 lanes, a 256-thread block is 4 waves.
#ifndef SCAPH_COMMON_H
#define SCAPH_COMMON_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "scaph_hip.h"

#define SCAPH_WAVE 64
#define SCAPH_TRY(expr)                           \
  do {                                            \
    hipError_t _e = (expr);                       \
    if (_e != hipSuccess) return (int)_e;         \
  } while (0)
#define SCAPH_LAUNCH_CHECK() SCAPH_TRY(hipGetLastError())

static inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Device address of the no-alloc push workspace on device `dev` (which must
// be the current device), looked up once per device (bfs.hip).
int scaph_noalloc_scratch_setup(int dev, void** out);

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

// Number of set bits of `mask` below this lane (ballot prefix).
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

template <typename T>
__device__ __forceinline__ T wave_bcast(T v, int src) {
  return __shfl(v, src, SCAPH_WAVE);
}

__device__ __forceinline__ uint64_t wave_bcast64(uint64_t v, int src) {
  uint32_t lo = __shfl((uint32_t)v, src, SCAPH_WAVE);
  uint32_t hi = __shfl((uint32_t)(v >> 32), src, SCAPH_WAVE);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sum across the 64 lanes of a wave.
template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < SCAPH_WAVE; d <<= 1) {
    T o = __shfl_up(v, d, SCAPH_WAVE);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

__device__ __forceinline__ unsigned long long wave_incl_scan_u64(unsigned long long v) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < SCAPH_WAVE; d <<= 1) {
    unsigned long long o = __shfl_up(v, d, SCAPH_WAVE);
    if (lane >= (uint32_t)d) v += o;
  }
  return v;
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int d = SCAPH_WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, SCAPH_WAVE);
  return v;
}

// Block-wide inclusive scan for blockDim.x == 256 (4 waves).  `wtot` is an
// LDS scratch of 4 entries.  Returns the inclusive prefix; *total gets the
// block total.
template <typename T>
__device__ __forceinline__ T block256_incl_scan(T v, T* wtot, T* total) {
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x / SCAPH_WAVE;
  T x = wave_incl_scan(v);
  if (lane == SCAPH_WAVE - 1) wtot[wid] = x;
  __syncthreads();
  T off = 0;
  T all = 0;
#pragma unroll
  for (uint32_t w = 0; w < 4; ++w) {
    T t = wtot[w];
    if (w < wid) off += t;
    all += t;
  }
  __syncthreads();
  *total = all;
  return x + off;
}

__device__ __forceinline__ bool bit_test(const uint32_t* bm, uint32_t v) {
  return (bm[v >> 5] >> (v & 31)) & 1u;
}

// Page table staged in LDS: lefts, node counts and buffer addresses of up
// to SCAPH_LDS_PAGES pages, in the kernel's dynamic LDS (launch with
// page_lds_bytes(num_pages) bytes); larger page sets are searched in HBM.
#define SCAPH_LDS_PAGES 1024

static inline size_t page_lds_bytes(uint32_t num_pages) {
  const size_t s = num_pages < SCAPH_LDS_PAGES ? num_pages : SCAPH_LDS_PAGES;
  return s * 16;
}

struct PageTab {
  const uint64_t* ptr;
  const uint32_t* left;
  const uint32_t* nodenum;
  bool in_lds;
};

__device__ __forceinline__ PageTab stage_pages(const scaph_graph_desc& g, void* lds) {
  PageTab t;
  t.in_lds = g.num_pages <= SCAPH_LDS_PAGES;
  if (!t.in_lds) {
    t.ptr = nullptr;
    t.left = nullptr;
    t.nodenum = nullptr;
    return t;
  }
  const uint32_t np = g.num_pages;
  uint64_t* ptr = static_cast<uint64_t*>(lds);
  uint32_t* left = reinterpret_cast<uint32_t*>(ptr + np);
  uint32_t* nn = left + np;
  for (uint32_t p = threadIdx.x; p < np; p += blockDim.x) {
    ptr[p] = g.page_ptr[p];
    left[p] = g.pages[p].left;
    nn[p] = g.pages[p].nodenum;
  }
  __syncthreads();
  t.ptr = ptr;
  t.left = left;
  t.nodenum = nn;
  return t;
}

// Page holding vertex v: the last page whose left <= v.
__device__ __forceinline__ uint32_t find_page(const PageTab& t, const scaph_graph_desc& g,
                                              uint32_t v) {
  uint32_t lo = 0, hi = g.num_pages;  // invariant: left[lo] <= v < left[hi]
  if (t.in_lds) {
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (t.left[mid] <= v) lo = mid; else hi = mid;
    }
  } else {
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (g.pages[mid].left <= v) lo = mid; else hi = mid;
    }
  }
  return lo;
}

struct NodeRef {
  const uint32_t* edges;  // first neighbour of v
  uint32_t len;           // out- (or in-) degree of v
};

__device__ __forceinline__ NodeRef load_node(const PageTab& t, const scaph_graph_desc& g,
                                             uint32_t v) {
  const uint32_t p = find_page(t, g, v);
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

#endif
