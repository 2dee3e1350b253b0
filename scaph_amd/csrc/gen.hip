// gen.hip — synthetic graph construction on the device: R-MAT edge keys,
// sort + dedup into a flat CSR (out- or in-edges), and page buffers in the
// tools/bin2pagecsr.cpp:221-242 layout.  None of this is timed as part of a
// BFS; it exists so a 2-billion-edge graph is built in seconds on the GPU
// instead of minutes on the host.  The edge stream is defined in
// rmat_spec.h and reproduced bit for bit by the host generator.
#include <cstring>

#include <rocprim/rocprim.hpp>

#include "common.h"
#include "rmat_spec.h"

namespace {

constexpr uint32_t kBlock = 256;

inline unsigned grid_for(size_t items, size_t per_block, unsigned cap) {
  size_t b = (items + per_block - 1) / per_block;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return (unsigned)b;
}

__global__ void __launch_bounds__(kBlock)
k_rmat_keys(uint64_t* keys, uint64_t e_begin, uint64_t count, uint32_t scale, uint32_t thr_a,
            uint32_t thr_ab, uint32_t thr_abc, uint64_t seed, scaph_spec::PermKey pk) {
  const uint64_t n = 1ull << scale;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    uint32_t u, v;
    scaph_spec::rmat_edge(seed, e_begin + i, scale, thr_a, thr_ab, thr_abc, &u, &v);
    u = (uint32_t)scaph_spec::perm_range(pk, u, n);
    v = (uint32_t)scaph_spec::perm_range(pk, v, n);
    if (u == v) u = v = (uint32_t)(n - 1);
    keys[i] = ((uint64_t)u << 32) | v;
  }
}

// Directed G(n, m): endpoints floor(w * n / 2^64) of SplitMix64 words 2e
// and 2e+1 (Lemire's multiply-high mapping), as the host generator.
__global__ void __launch_bounds__(kBlock)
k_er_keys(uint64_t* keys, uint64_t e_begin, uint64_t count, uint32_t n, uint64_t seed) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t e = e_begin + i;
    const uint64_t wu = scaph_spec::word(seed, 2 * e), wv = scaph_spec::word(seed, 2 * e + 1);
    const uint32_t u = (uint32_t)__umul64hi(wu, (uint64_t)n);
    const uint32_t v = (uint32_t)__umul64hi(wv, (uint64_t)n);
    keys[i] = ((uint64_t)u << 32) | v;
  }
}

// Per vertex: the first three entries of its list (0xFFFFFFFF past the end)
// and its length, as one 16-byte record; with out_degree, the first two
// entries, the length and the vertex's out-degree.
__global__ void __launch_bounds__(kBlock)
k_list_heads(const uint64_t* offsets, const uint32_t* ngh, uint32_t n, const uint32_t* out_degree,
             uint4* head) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    const uint64_t o = offsets[v];
    const uint32_t d = (uint32_t)(offsets[v + 1] - o);
    uint4 r;
    r.x = d > 0 ? ngh[o] : 0xFFFFFFFFu;
    r.y = d > 1 ? ngh[o + 1] : 0xFFFFFFFFu;
    if (out_degree) {
      r.z = d;
      r.w = out_degree[v];
    } else {
      r.z = d > 2 ? ngh[o + 2] : 0xFFFFFFFFu;
      r.w = d;
    }
    head[v] = r;
  }
}

// Vertices with in-edges per window of 1024 ids (32 bitmap words): the
// complement of the no-in-edge bitmap (whose padding bits past n are set).
__global__ void __launch_bounds__(kBlock)
k_window_live(const uint32_t* no_in, uint64_t words, uint32_t windows, uint32_t* cnt) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < windows; k += stride) {
    uint32_t c = 0;
    for (uint64_t w = (uint64_t)k * 32; w < (uint64_t)k * 32 + 32 && w < words; ++w)
      c += __popc(~no_in[w]);
    cnt[k] = c;
  }
}

// Head records of the vertices with in-edges only, in id order: a thread
// per window writes its live vertices' records from the window's base on.
__global__ void __launch_bounds__(kBlock)
k_list_heads_live(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                  const uint32_t* out_degree, const uint32_t* no_in, const uint32_t* base,
                  uint32_t windows, uint4* head) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < windows; k += stride) {
    uint32_t r = base[k];
    for (uint64_t w = (uint64_t)k * 32; w < (uint64_t)k * 32 + 32 && w < words; ++w) {
      uint32_t bits = ~no_in[w];
      while (bits) {
        const uint32_t b = __ffs(bits) - 1;
        bits &= bits - 1;
        const uint32_t v = (uint32_t)(w * 32 + b);
        if (v >= n) continue;
        const uint64_t o = offsets[v];
        const uint32_t d = (uint32_t)(offsets[v + 1] - o);
        uint4 rec;
        rec.x = d > 0 ? ngh[o] : 0xFFFFFFFFu;
        rec.y = d > 1 ? ngh[o + 1] : 0xFFFFFFFFu;
        if (out_degree) {
          rec.z = d;
          rec.w = out_degree[v];
        } else {
          rec.z = d > 2 ? ngh[o + 2] : 0xFFFFFFFFu;
          rec.w = d;
        }
        head[r++] = rec;
      }
    }
  }
}

// Indexed head records (scaph_list_heads_indexed): a thread per 1024-id
// window finds the page of the window's first vertex once, then walks the
// window's vertices (all of them, or those with in-edges when no_in is set)
// with the page advancing as ids pass its end.
__global__ void __launch_bounds__(kBlock)
k_list_heads_idx(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                 const uint32_t* out_degree, const scaph_page* pages, uint32_t num_pages,
                 const uint32_t* no_in, const uint32_t* base, uint32_t windows, uint4* head) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < windows; k += stride) {
    const uint32_t v0 = k * 1024u;
    uint32_t lo = 0, hi = num_pages;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pages[mid].left <= v0) lo = mid; else hi = mid;
    }
    uint32_t p = lo;
    uint64_t page_e0 = offsets[pages[p].left];
    uint32_t r = base ? base[k] : v0;
    for (uint64_t w = (uint64_t)k * 32; w < (uint64_t)k * 32 + 32 && w < words; ++w) {
      uint32_t bits = no_in ? ~no_in[w] : ~0u;
      while (bits) {
        const uint32_t b = __ffs(bits) - 1;
        bits &= bits - 1;
        const uint32_t v = (uint32_t)(w * 32 + b);
        if (v >= n) continue;
        while (p + 1 < num_pages && pages[p + 1].left <= v) {
          ++p;
          page_e0 = offsets[pages[p].left];
        }
        const uint64_t o = offsets[v];
        const uint32_t d = (uint32_t)(offsets[v + 1] - o);
        uint4 rec;
        rec.x = d > 0 ? ngh[o] : 0xFFFFFFFFu;
        rec.y = d > 1 ? ngh[o + 1] : 0xFFFFFFFFu;
        rec.z = ((uint32_t)(o - page_e0) << 8) | (d < 255u ? d : 255u);
        rec.w = out_degree[v];
        head[r++] = rec;
      }
    }
  }
}

// Power-law digraph edges: edge e belongs to the last vertex v with
// offsets[v] <= e (offsets = exclusive prefix of the out-degrees, n + 1
// entries); its destination is the multiply-high of SplitMix64 word e of the
// salted seed (the host and oracle generators use the same stream).
constexpr uint64_t kZipfDstSalt = 0x7A6970662D647374ull;  // "zipf-dst"

__global__ void __launch_bounds__(kBlock)
k_zipf_keys(uint64_t* keys, const uint64_t* offsets, uint32_t n, uint64_t e_begin, uint64_t count,
            uint64_t seed) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t e = e_begin + i;
    uint32_t lo = 0, hi = n;  // offsets[lo] <= e < offsets[hi]
    while (hi - lo > 1) {
      const uint32_t mid = lo + (hi - lo) / 2;
      if (offsets[mid] <= e) lo = mid; else hi = mid;
    }
    const uint64_t w = scaph_spec::word(seed ^ kZipfDstSalt, e);
    const uint32_t t = (uint32_t)__umul64hi(w, (uint64_t)n);
    keys[i] = ((uint64_t)lo << 32) | t;
  }
}

__global__ void k_offsets_tail(const uint32_t* deg, uint64_t* offsets, uint32_t n) {
  offsets[n] = offsets[n - 1] + deg[n - 1];
}

// (u<<32|v) -> packed sort key.  Out-edges: (u << bits) | v.  In-edges:
// (v << (bits+1)) | (flag << bits) | low where a hub source u (hub_idx[u]
// != ~0) gets flag 0 and low = its hub index, any other source flag 1 and
// low = u — so each in-list sorts hubs first (by hub index), then the rest
// by id.  Self-loops all become the largest key, dropped after dedup.
__global__ void __launch_bounds__(kBlock)
k_pack(uint64_t* keys, uint64_t count, uint32_t bits, int transpose, uint32_t n,
       const uint32_t* hub_idx) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t sentinel_t = ((uint64_t)(n - 1) << (bits + 1)) | (1ull << bits) | (n - 1);
  const uint64_t sentinel_o = ((uint64_t)(n - 1) << bits) | (n - 1);
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t k = keys[i];
    const uint32_t u = (uint32_t)(k >> 32), v = (uint32_t)k;
    if (!transpose) {
      keys[i] = u == v ? sentinel_o : (((uint64_t)u << bits) | v);
    } else if (u == v) {
      keys[i] = sentinel_t;
    } else {
      const uint32_t h = hub_idx ? hub_idx[u] : 0xFFFFFFFFu;
      const uint64_t low = h != 0xFFFFFFFFu ? (uint64_t)h : ((1ull << bits) | u);
      keys[i] = ((uint64_t)v << (bits + 1)) | low;
    }
  }
}

// Sorted unique keys -> neighbour entries + run starts of each vertex.
// `abits` is the shift of the owning vertex, `bits` the width of the id
// field; in-edge keys (tagged) carry the hub flag at bit `bits`.
__global__ void __launch_bounds__(kBlock)
k_run_starts(const uint64_t* uniq, uint64_t count, uint32_t abits, uint32_t bits, int tagged,
             uint32_t* ngh, uint64_t* start) {
  const uint64_t mask = (1ull << bits) - 1;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint64_t k = uniq[i];
    const uint32_t a = (uint32_t)(k >> abits);
    uint32_t e = (uint32_t)(k & mask);
    if (tagged && !((k >> bits) & 1ull)) e |= 0x80000000u;  // hub index
    ngh[i] = e;
    if (i == 0 || (uint32_t)(uniq[i - 1] >> abits) != a) start[a] = i;
  }
}

__global__ void __launch_bounds__(kBlock)
k_run_ends(const uint64_t* uniq, uint64_t count, uint32_t abits, const uint64_t* start,
           uint32_t* degree) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) {
    const uint32_t a = (uint32_t)(uniq[i] >> abits);
    if (i + 1 == count || (uint32_t)(uniq[i + 1] >> abits) != a)
      degree[a] = (uint32_t)(i + 1 - start[a]);
  }
}

__global__ void k_set_u64(uint64_t* p, uint64_t v) { *p = v; }

// Nodes of every page: Node{vtx, idx (page-relative), len}.
__global__ void __launch_bounds__(kBlock)
k_page_nodes(const uint64_t* offsets, const scaph_page* pages, const uint64_t* page_ptr,
             uint32_t num_pages, uint32_t n) {
  const uint32_t stride = gridDim.x * blockDim.x;
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += stride) {
    uint32_t lo = 0, hi = num_pages;
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) >> 1;
      if (pages[mid].left <= v) lo = mid; else hi = mid;
    }
    const uint32_t left = pages[lo].left;
    if (v > pages[lo].right) continue;
    uint32_t* buf = reinterpret_cast<uint32_t*>(page_ptr[lo]);
    const uint32_t i = v - left;
    buf[3ull * i] = v;
    buf[3ull * i + 1] = (uint32_t)(offsets[v] - offsets[left]);
    buf[3ull * i + 2] = (uint32_t)(offsets[v + 1] - offsets[v]);
  }
}

// Neighbour arrays: blockIdx.y = page.
__global__ void __launch_bounds__(kBlock)
k_page_edges(const uint64_t* offsets, const uint32_t* ngh, const scaph_page* pages,
             const uint64_t* page_ptr) {
  const scaph_page pg = pages[blockIdx.y];
  const uint64_t e0 = offsets[pg.left], e1 = offsets[(uint64_t)pg.right + 1];
  uint32_t* dst = reinterpret_cast<uint32_t*>(page_ptr[blockIdx.y]) + 3ull * pg.nodenum;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < e1 - e0; i += stride)
    dst[i] = ngh[e0 + i];
}

inline uint32_t bits_for(uint32_t n) {
  uint32_t b = 1;
  while (b < 32 && (1ull << b) < n) ++b;
  return b;
}

}  // namespace

extern "C" {

int scaph_gen_rmat_keys(uint64_t* keys, uint64_t e_begin, uint64_t e_end, uint32_t scale,
                        uint32_t thr_a, uint32_t thr_ab, uint32_t thr_abc, uint64_t seed,
                        void* stream) {
  if (!keys || scale == 0 || scale > 31 || e_end < e_begin) return SCAPH_STATUS_INVALID;
  if (e_end == e_begin) return 0;
  const scaph_spec::PermKey pk = scaph_spec::make_perm_key(seed, scale);
  const uint64_t count = e_end - e_begin;
  hipLaunchKernelGGL(k_rmat_keys, dim3(grid_for(count, kBlock * 8, 65536)), dim3(kBlock), 0,
                     as_stream(stream), keys, e_begin, count, scale, thr_a, thr_ab, thr_abc, seed,
                     pk);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_gen_er_keys(uint64_t* keys, uint64_t e_begin, uint64_t e_end, uint32_t n,
                      uint64_t seed, void* stream) {
  if (!keys || n == 0 || e_end < e_begin) return SCAPH_STATUS_INVALID;
  if (e_end == e_begin) return 0;
  const uint64_t count = e_end - e_begin;
  hipLaunchKernelGGL(k_er_keys, dim3(grid_for(count, kBlock * 8, 65536)), dim3(kBlock), 0,
                     as_stream(stream), keys, e_begin, count, n, seed);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_list_heads_from_csr(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                              const uint32_t* out_degree, uint32_t* head, void* stream) {
  if (!offsets || !head || n == 0) return SCAPH_STATUS_INVALID;
  hipLaunchKernelGGL(k_list_heads, dim3(grid_for(n, kBlock, 16384)), dim3(kBlock), 0,
                     as_stream(stream), offsets, ngh, n, out_degree,
                     reinterpret_cast<uint4*>(head));
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_live_window_base(const uint32_t* no_in_edges, uint32_t n, uint32_t* base,
                           void* stream) {
  if (!no_in_edges || !base || n == 0) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  const uint32_t windows = (uint32_t)((words + 31) / 32);
  uint32_t* cnt = nullptr;
  SCAPH_TRY(hipMalloc(&cnt, (size_t)windows * 4 + 16));
  hipLaunchKernelGGL(k_window_live, dim3(grid_for(windows, kBlock, 4096)), dim3(kBlock), 0, st,
                     no_in_edges, words, windows, cnt);
  int rc = (int)hipGetLastError();
  size_t bytes = 0;
  void* tmp = nullptr;
  if (!rc) rc = (int)rocprim::exclusive_scan(nullptr, bytes, cnt, base, 0u, (size_t)windows,
                                             rocprim::plus<uint32_t>(), st);
  if (!rc) rc = (int)hipMalloc(&tmp, bytes + 16);
  if (!rc) rc = (int)rocprim::exclusive_scan(tmp, bytes, cnt, base, 0u, (size_t)windows,
                                             rocprim::plus<uint32_t>(), st);
  if (!rc) rc = (int)hipStreamSynchronize(st);
  const int rf = (int)hipFree(cnt);
  const int rt = tmp ? (int)hipFree(tmp) : 0;
  return rc ? rc : (rf ? rf : rt);
}

int scaph_list_heads_live(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                          const uint32_t* out_degree, const uint32_t* no_in_edges,
                          const uint32_t* base, uint32_t* head, void* stream) {
  if (!offsets || !no_in_edges || !base || !head || n == 0) return SCAPH_STATUS_INVALID;
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  const uint32_t windows = (uint32_t)((words + 31) / 32);
  hipLaunchKernelGGL(k_list_heads_live, dim3(grid_for(windows, kBlock, 4096)), dim3(kBlock), 0,
                     as_stream(stream), offsets, ngh, n, out_degree, no_in_edges, base, windows,
                     reinterpret_cast<uint4*>(head));
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_list_heads_indexed(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                             const uint32_t* out_degree, const scaph_page* pages,
                             uint32_t num_pages, const uint32_t* no_in_edges, const uint32_t* base,
                             uint32_t* head, void* stream) {
  if (!offsets || !out_degree || !pages || num_pages == 0 || !head || n == 0 ||
      (!no_in_edges) != (!base))
    return SCAPH_STATUS_INVALID;
  const uint64_t words = (((uint64_t)n + 63) / 64) * 2;
  const uint32_t windows = (uint32_t)((words + 31) / 32);
  hipLaunchKernelGGL(k_list_heads_idx, dim3(grid_for(windows, kBlock, 4096)), dim3(kBlock), 0,
                     as_stream(stream), offsets, ngh, n, out_degree, pages, num_pages,
                     no_in_edges, base, windows, reinterpret_cast<uint4*>(head));
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_degrees_to_offsets(const uint32_t* degree, uint64_t* offsets, uint32_t n,
                             void* stream) {
  if (!degree || !offsets || n == 0) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  size_t bytes = 0;
  SCAPH_TRY(rocprim::exclusive_scan(nullptr, bytes, degree, offsets, (uint64_t)0, (size_t)n,
                                    rocprim::plus<uint64_t>(), st));
  void* tmp = nullptr;
  SCAPH_TRY(hipMalloc(&tmp, bytes + 16));
  int rc = (int)rocprim::exclusive_scan(tmp, bytes, degree, offsets, (uint64_t)0, (size_t)n,
                                        rocprim::plus<uint64_t>(), st);
  if (!rc) {
    hipLaunchKernelGGL(k_offsets_tail, dim3(1), dim3(1), 0, st, degree, offsets, n);
    rc = (int)hipGetLastError();
  }
  if (!rc) rc = (int)hipStreamSynchronize(st);
  const int rf = (int)hipFree(tmp);
  return rc ? rc : rf;
}

int scaph_gen_zipf_keys(uint64_t* keys, const uint64_t* offsets, uint32_t n, uint64_t e_begin,
                        uint64_t e_end, uint64_t seed, void* stream) {
  if (!keys || !offsets || n == 0 || e_end < e_begin) return SCAPH_STATUS_INVALID;
  if (e_end == e_begin) return 0;
  const uint64_t count = e_end - e_begin;
  hipLaunchKernelGGL(k_zipf_keys, dim3(grid_for(count, kBlock * 8, 65536)), dim3(kBlock), 0,
                     as_stream(stream), keys, offsets, n, e_begin, count, seed);
  SCAPH_LAUNCH_CHECK();
  return 0;
}

int scaph_csr_from_keys(uint64_t* keys, uint64_t* keys_tmp, uint64_t num_keys, uint32_t n,
                        int transpose, const uint32_t* hub_idx, uint64_t* offsets, uint32_t* ngh,
                        uint32_t* degree, uint64_t* num_edges, void* stream) {
  if (!keys || !keys_tmp || !offsets || !ngh || !degree || !num_edges || n == 0)
    return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  const uint32_t bits = bits_for(n);
  if (transpose && n > 0x7FFFFFFFu) return SCAPH_STATUS_INVALID;  // top bit tags hub entries
  const uint32_t abits = transpose ? bits + 1 : bits;
  const uint32_t key_bits = abits + bits;
  const int tagged = transpose ? 1 : 0;
  SCAPH_TRY(hipMemsetAsync(degree, 0, (size_t)n * 4, st));
  if (num_keys == 0) {
    SCAPH_TRY(hipMemsetAsync(offsets, 0, ((size_t)n + 1) * 8, st));
    SCAPH_TRY(hipStreamSynchronize(st));
    *num_edges = 0;
    return 0;
  }
  hipLaunchKernelGGL(k_pack, dim3(grid_for(num_keys, kBlock * 8, 65536)), dim3(kBlock), 0, st,
                     keys, num_keys, bits, transpose, n, transpose ? hub_idx : nullptr);
  SCAPH_LAUNCH_CHECK();

  // sort (ping-pong between keys and keys_tmp)
  rocprim::double_buffer<uint64_t> db(keys, keys_tmp);
  size_t sort_bytes = 0;
  SCAPH_TRY(rocprim::radix_sort_keys(nullptr, sort_bytes, db, (size_t)num_keys, 0u, key_bits, st));
  uint64_t* sorted_count = nullptr;
  void* tmp = nullptr;
  size_t uniq_bytes = 0;
  SCAPH_TRY(rocprim::unique(nullptr, uniq_bytes, db.current(), db.alternate(), sorted_count,
                            (size_t)num_keys, rocprim::equal_to<uint64_t>(), st));
  size_t scan_bytes = 0;
  SCAPH_TRY(rocprim::exclusive_scan(nullptr, scan_bytes, degree, offsets, (uint64_t)0, (size_t)n,
                                    rocprim::plus<uint64_t>(), st));
  size_t tmp_bytes = sort_bytes;
  if (uniq_bytes > tmp_bytes) tmp_bytes = uniq_bytes;
  if (scan_bytes > tmp_bytes) tmp_bytes = scan_bytes;
  tmp_bytes = (tmp_bytes + 255) & ~size_t(255);
  SCAPH_TRY(hipMalloc(&tmp, tmp_bytes + 256));
  uint64_t* d_count = reinterpret_cast<uint64_t*>(static_cast<char*>(tmp) + tmp_bytes);
  int rc = 0;
  do {
    size_t b = sort_bytes;
    if ((rc = (int)rocprim::radix_sort_keys(tmp, b, db, (size_t)num_keys, 0u, key_bits, st)))
      break;
    uint64_t* in = db.current();
    uint64_t* out = db.alternate();
    b = uniq_bytes;
    if ((rc = (int)rocprim::unique(tmp, b, in, out, d_count, (size_t)num_keys,
                                   rocprim::equal_to<uint64_t>(), st)))
      break;
    uint64_t h_count = 0, last = 0;
    if ((rc = (int)hipMemcpyAsync(&h_count, d_count, 8, hipMemcpyDeviceToHost, st))) break;
    if ((rc = (int)hipStreamSynchronize(st))) break;
    if (h_count > 0) {
      if ((rc = (int)hipMemcpyAsync(&last, out + (h_count - 1), 8, hipMemcpyDeviceToHost, st)))
        break;
      if ((rc = (int)hipStreamSynchronize(st))) break;
      const uint64_t loop_key =
          transpose ? (((uint64_t)(n - 1) << (bits + 1)) | (1ull << bits) | (uint64_t)(n - 1))
                    : (((uint64_t)(n - 1) << bits) | (uint64_t)(n - 1));
      if (last == loop_key) --h_count;
    }
    // `in` is free now: reuse it for the run starts (n u64 <= num_keys u64?)
    uint64_t* start = in;
    if ((size_t)n > num_keys) {
      // tiny graph with more vertices than raw edges: need a separate buffer
      if ((rc = (int)hipMalloc(reinterpret_cast<void**>(&start), (size_t)n * 8))) break;
    }
    if (h_count > 0) {
      hipLaunchKernelGGL(k_run_starts, dim3(grid_for(h_count, kBlock * 8, 65536)), dim3(kBlock),
                         0, st, out, h_count, abits, bits, tagged, ngh, start);
      if ((rc = (int)hipGetLastError())) break;
      hipLaunchKernelGGL(k_run_ends, dim3(grid_for(h_count, kBlock * 8, 65536)), dim3(kBlock), 0,
                         st, out, h_count, abits, start, degree);
      if ((rc = (int)hipGetLastError())) break;
    }
    b = scan_bytes;
    if ((rc = (int)rocprim::exclusive_scan(tmp, b, degree, offsets, (uint64_t)0, (size_t)n,
                                           rocprim::plus<uint64_t>(), st)))
      break;
    hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, st, offsets + n, h_count);
    if ((rc = (int)hipGetLastError())) break;
    if ((rc = (int)hipStreamSynchronize(st))) break;
    if (start != in) (void)hipFree(start);
    *num_edges = h_count;
  } while (0);
  (void)hipFree(tmp);
  return rc;
}

int scaph_pages_from_csr(const uint64_t* offsets, const uint32_t* ngh, const scaph_page* pages,
                         const uint64_t* page_ptr, uint32_t num_pages, uint32_t n, void* stream) {
  if (!offsets || !pages || !page_ptr) return SCAPH_STATUS_INVALID;
  if (num_pages == 0 || n == 0) return 0;
  if (num_pages > 65535) return SCAPH_STATUS_INVALID;
  hipStream_t st = as_stream(stream);
  hipLaunchKernelGGL(k_page_nodes, dim3(grid_for(n, kBlock * 4, 16384)), dim3(kBlock), 0, st,
                     offsets, pages, page_ptr, num_pages, n);
  SCAPH_LAUNCH_CHECK();
  if (ngh) {
    hipLaunchKernelGGL(k_page_edges, dim3(64, num_pages), dim3(kBlock), 0, st, offsets, ngh, pages,
                       page_ptr);
    SCAPH_LAUNCH_CHECK();
  }
  return 0;
}

}  // extern "C"
