/*
 * oracle.c — CPU restatement of the Scaph BFS hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by the cpu_baseline leg of
 * bench.py, and by nothing else.  The product (scaph_amd/) never links,
 * imports or calls it.
 *
 * Functions, the reference code they restate, and what pins them
 * (tests/test_oracle.py):
 *   oracle_bfs            app/bfs.cu:67-117 (check()), relax of
 *                         app/bfs_kernel.cuh:6-10 — pinned by scipy
 *                         shortest paths, closed forms (path, star, grid,
 *                         binary tree, cycle), brute-force relaxation and the
 *                         depth-63 cap on a long path.
 *   oracle_bfs_mt         the same distances, level-synchronous on POSIX
 *                         threads (full-size checks) — pinned by equality
 *                         with oracle_bfs, including the depth-63 cap.
 *   oracle_wcc            app/wcc.cu:60-85 (check()): value[v] = v, then
 *                         sweeps of the out-edges in vertex order lowering
 *                         value[ng] to value[vid] until nothing changes —
 *                         pinned by scipy connected components on symmetric
 *                         graphs and by brute-force reachability (least id
 *                         among the vertices reaching v) on small digraphs.
 *   oracle_sssp           app/sssp.cu / app/sssp_kernel.cuh (the relax
 *                         dst = min(dst, src + w), value 0x7FFFFFFF for
 *                         unreached): Dijkstra with a binary heap — pinned by
 *                         scipy's dijkstra and, with unit weights, by
 *                         oracle_bfs.
 *   oracle_page_cuts,     tools/bin2pagecsr.cpp:106-163, 197-252 — pinned
 *   oracle_fill_page      byte-for-byte against the reference tool compiled
 *                         from its own source (oracle/_ref) and the
 *                         committed golden hashes.
 *   oracle_page_counts    src/gpu_engine.cuh:397-456 — pinned by a numpy
 *                         closed form (sum of stat and stat*(deg+3)).
 *   oracle_compress_page  src/compress.h:18-40 — pinned by the closed form
 *                         (compacted words = sum over active (len) + 3*count,
 *                         each active list equal to its source list).
 *   oracle_max_zero_run,  src/tools.h:10-85 — pinned by a brute-force
 *   oracle_fit_zero_run,  definition over all 2^16 low masks and random
 *   oracle_set_run,       32-bit masks.
 *   oracle_clear_run
 *   oracle_pool_size      src/app.cuh:281-286, 323-324 — closed form.
 *   oracle_rmat_csr,      the project's own generator definition (the
 *   oracle_er_csr         reference ships none): SplitMix64 words, integer
 *                         quadrant thresholds, keyed bijective relabelling.
 *                         Pinned by quadrant frequencies, bijectivity of the
 *                         relabelling, and agreement with the product's host
 *                         and device generators.
 *   oracle_zipf_degree,   the project's power-law digraph: out-degree
 *   oracle_zipf_csr       min(cap, floor(x_min * (1 - u)^(-1/(alpha - 1))))
 *                         with u from SplitMix64 word v (53 bits), uniform
 *                         destinations (multiply-high of word e of the
 *                         salted seed).  Pinned by the Pareto tail
 *                         P(d >= k) = (x_min / k)^(alpha - 1), the capped
 *                         mean in closed form, and agreement with the
 *                         product's host and device generators.
 *   oracle_mlp_*          fp32 restatement of the 16-64-64-2 MLP; pinned
 *                         by finite differences and a numpy matmul form.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORACLE_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------ BFS */

/* FIFO level BFS (app/bfs.cu:67-117): cur/next vertex lists, a vertex v is
 * relaxed from u when value[u] + 1 < value[v].  `inf` is the initial
 * value: 63 reproduces MAXDEPTH (app/bfs.cu:18), 0xFFFFFFFF gives exact
 * depths. */
/* app/wcc.cu:60-85: label sweeps until a sweep changes nothing. */
ORACLE_API int oracle_wcc(uint32_t n, const uint64_t* offsets, const uint32_t* ngh,
                          uint32_t* value) {
  for (uint32_t v = 0; v < n; ++v) value[v] = v;
  int changed = 1;
  while (changed) {
    changed = 0;
    for (uint32_t vid = 0; vid < n; ++vid)
      for (uint64_t k = offsets[vid]; k < offsets[(uint64_t)vid + 1]; ++k) {
        const uint32_t ng = ngh[k];
        if (ng < n && value[vid] < value[ng]) {
          value[ng] = value[vid];
          changed = 1;
        }
      }
  }
  return 0;
}

/* Shortest paths from src over u32 weights (Dijkstra, binary heap of
 * (distance, vertex) pairs with lazy deletion); 0x7FFFFFFF for unreached
 * (MAXPATH, app/sssp.cu:11).  Returns 3 if a distance reaches 2^31. */
ORACLE_API int oracle_sssp(uint32_t n, const uint64_t* offsets, const uint32_t* ngh,
                           const uint32_t* wgh, uint32_t src, uint32_t* value) {
  if (src >= n) return 1;
  const uint64_t m = offsets[n];
  uint64_t cap = m + 2, size = 0;
  uint64_t* hd = (uint64_t*)malloc(sizeof(uint64_t) * cap);
  uint32_t* hv = (uint32_t*)malloc(sizeof(uint32_t) * cap);
  uint64_t* dist = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
  if (!hd || !hv || !dist) {
    free(hd);
    free(hv);
    free(dist);
    return 2;
  }
  for (uint32_t v = 0; v < n; ++v) dist[v] = UINT64_MAX;
  dist[src] = 0;
  hd[0] = 0;
  hv[0] = src;
  size = 1;
  int rc = 0;
  while (size) {
    const uint64_t d = hd[0];
    const uint32_t u = hv[0];
    /* pop: move the last element to the root and sift down */
    --size;
    hd[0] = hd[size];
    hv[0] = hv[size];
    for (uint64_t i = 0;;) {
      uint64_t l = 2 * i + 1, r = l + 1, s = i;
      if (l < size && hd[l] < hd[s]) s = l;
      if (r < size && hd[r] < hd[s]) s = r;
      if (s == i) break;
      const uint64_t td = hd[i];
      const uint32_t tv = hv[i];
      hd[i] = hd[s];
      hv[i] = hv[s];
      hd[s] = td;
      hv[s] = tv;
      i = s;
    }
    if (d != dist[u]) continue; /* stale entry */
    for (uint64_t k = offsets[u]; k < offsets[(uint64_t)u + 1]; ++k) {
      const uint32_t v = ngh[k];
      if (v >= n) continue;
      const uint64_t nd = d + wgh[k];
      if (nd < dist[v]) {
        dist[v] = nd;
        if (size == cap) { /* lazy entries beyond m + 2: grow */
          cap *= 2;
          uint64_t* nhd = (uint64_t*)realloc(hd, sizeof(uint64_t) * cap);
          uint32_t* nhv = (uint32_t*)realloc(hv, sizeof(uint32_t) * cap);
          if (!nhd || !nhv) {
            free(nhd ? nhd : hd);
            free(nhv ? nhv : hv);
            free(dist);
            return 2;
          }
          hd = nhd;
          hv = nhv;
        }
        uint64_t i = size++;
        hd[i] = nd;
        hv[i] = v;
        while (i > 0) { /* sift up */
          const uint64_t p = (i - 1) / 2;
          if (hd[p] <= hd[i]) break;
          const uint64_t td = hd[i];
          const uint32_t tv = hv[i];
          hd[i] = hd[p];
          hv[i] = hv[p];
          hd[p] = td;
          hv[p] = tv;
          i = p;
        }
      }
    }
  }
  for (uint32_t v = 0; v < n; ++v) {
    if (dist[v] == UINT64_MAX) {
      value[v] = 0x7FFFFFFFu;
    } else {
      if (dist[v] >= 0x7FFFFFFFull) rc = 3;
      value[v] = (uint32_t)dist[v];
    }
  }
  free(hd);
  free(hv);
  free(dist);
  return rc;
}

ORACLE_API int oracle_bfs(uint32_t n, const uint64_t* offsets, const uint32_t* ngh, uint32_t src,
                          uint32_t inf, uint32_t* value) {
  if (src >= n) return 1;
  uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  if (!cur || !next) {
    free(cur);
    free(next);
    return 2;
  }
  for (uint32_t v = 0; v < n; ++v) value[v] = inf;
  size_t ncur = 0, nnext = 0;
  cur[ncur++] = src;
  value[src] = 0;
  while (ncur) {
    for (size_t i = 0; i < ncur; ++i) {
      const uint32_t vid = cur[i];
      for (uint64_t k = offsets[vid]; k < offsets[(uint64_t)vid + 1]; ++k) {
        const uint32_t ng = ngh[k];
        if (value[vid] + 1 < value[ng]) {
          value[ng] = value[vid] + 1;
          next[nnext++] = ng; /* each vertex enters next once: its value only drops once */
        }
      }
    }
    uint32_t* t = cur;
    cur = next;
    next = t;
    ncur = nnext;
    nnext = 0;
  }
  free(cur);
  free(next);
  return 0;
}

/* The same distances as oracle_bfs, level-synchronous on `threads` POSIX
 * threads so that full-size graphs (billions of edges) check in seconds:
 * at level L every frontier vertex u offers L + 1 to its out-neighbours and
 * a neighbour still at `inf` is claimed by compare-and-swap when
 * L + 1 < inf (the relax of app/bfs_kernel.cuh:6-10 under the cap).  BFS
 * distances are unique, so the claim order cannot change the result.  It
 * is pinned by equality with oracle_bfs (tests/test_oracle.py). */
typedef struct {
  const uint64_t* off;
  const uint32_t* ngh;
  uint32_t* value;
  const uint32_t* cur;
  uint64_t ncur;
  uint32_t* next;
  uint64_t* nnext;  /* shared append position of next */
  uint64_t* take;   /* shared position in cur (chunks of 64 vertices) */
  uint32_t lvl, inf;
} bfs_mt_arg;

static void* bfs_mt_worker(void* p) {
  bfs_mt_arg* a = (bfs_mt_arg*)p;
  uint32_t buf[4096];
  uint32_t nb = 0;
  const uint32_t nl = a->lvl + 1;
  for (;;) {
    const uint64_t i0 = __atomic_fetch_add(a->take, 64, __ATOMIC_RELAXED);
    if (i0 >= a->ncur) break;
    const uint64_t i1 = i0 + 64 < a->ncur ? i0 + 64 : a->ncur;
    for (uint64_t i = i0; i < i1; ++i) {
      const uint32_t u = a->cur[i];
      for (uint64_t k = a->off[u]; k < a->off[(uint64_t)u + 1]; ++k) {
        const uint32_t v = a->ngh[k];
        uint32_t old = __atomic_load_n(&a->value[v], __ATOMIC_RELAXED);
        if (old != a->inf) continue;
        if (!__atomic_compare_exchange_n(&a->value[v], &old, nl, 0, __ATOMIC_RELAXED,
                                         __ATOMIC_RELAXED))
          continue;
        buf[nb++] = v;
        if (nb == 4096) {
          const uint64_t at = __atomic_fetch_add(a->nnext, nb, __ATOMIC_RELAXED);
          memcpy(a->next + at, buf, sizeof(uint32_t) * nb);
          nb = 0;
        }
      }
    }
  }
  if (nb) {
    const uint64_t at = __atomic_fetch_add(a->nnext, nb, __ATOMIC_RELAXED);
    memcpy(a->next + at, buf, sizeof(uint32_t) * nb);
  }
  return NULL;
}

ORACLE_API int oracle_bfs_mt(uint32_t n, const uint64_t* offsets, const uint32_t* ngh,
                             uint32_t src, uint32_t inf, uint32_t* value, uint32_t threads) {
  if (src >= n) return 1;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  uint32_t* cur = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* next = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * threads);
  if (!cur || !next || !th) {
    free(cur);
    free(next);
    free(th);
    return 2;
  }
  for (uint32_t v = 0; v < n; ++v) value[v] = inf;
  value[src] = 0;
  cur[0] = src;
  uint64_t ncur = 1;
  int rc = 0;
  for (uint32_t lvl = 0; ncur && lvl + 1 < inf; ++lvl) {
    uint64_t nnext = 0, take = 0;
    bfs_mt_arg a = {offsets, ngh, value, cur, ncur, next, &nnext, &take, lvl, inf};
    uint32_t started = 0;
    for (; started < threads; ++started)
      if (pthread_create(&th[started], NULL, bfs_mt_worker, &a) != 0) break;
    if (started == 0) bfs_mt_worker(&a);
    for (uint32_t t = 0; t < started; ++t) pthread_join(th[t], NULL);
    uint32_t* t = cur;
    cur = next;
    next = t;
    ncur = nnext;
  }
  free(cur);
  free(next);
  free(th);
  return rc;
}

/* --------------------------------------------------------- page layout */

/* Ranges past the last vertex never fit (the reference would read past the
 * offsets array there). */
static int fits(const uint64_t* pos, uint32_t n, uint32_t left, uint64_t r, uint64_t P, int w) {
  if (r >= n) return 0;
  uint64_t edges = pos[r + 1] - pos[left];
  return edges * (w ? 2u : 1u) + (r - left + 1) * 3u < P;
}

/* Page cut of tools/bin2pagecsr.cpp:206-208 with bsearch() (:106-163):
 * right starts at left + pagesize (clamped to the last vertex); if the
 * whole range fits it is kept, otherwise the bisection looks for the r
 * where fits(r) holds and fits(r+1) does not. */
ORACLE_API int oracle_page_cuts(const uint64_t* pos, uint32_t n, uint64_t P, int weighted,
                                uint32_t* pages /* 4 words per page */, uint32_t max_pages,
                                uint32_t* num_pages) {
  uint32_t left = 0, np = 0;
  while (left < n) {
    uint64_t right = (uint64_t)left + P;
    if (right > (uint64_t)n - 1) right = n - 1;
    if (!fits(pos, n, left, left, P, weighted)) return 2;
    if (!fits(pos, n, left, right, P, weighted)) {
      int64_t bleft = left, bright = (int64_t)right;
      int64_t bmid = ((int64_t)left + (int64_t)right) / 2;
      for (;;) {
        int fmid = fits(pos, n, left, (uint64_t)bmid, P, weighted);
        int fmidr = fits(pos, n, left, (uint64_t)bmid + 1, P, weighted);
        if (fmid && fmidr) {
          bleft = bmid + 1;
          bmid = (bleft + bright) / 2;
        } else if (!fmid && !fmidr) {
          bright = bmid - 1;
          bmid = (bleft + bright) / 2;
        } else if (fmid && !fmidr) {
          right = (uint64_t)bmid;
          break;
        } else {
          return 3; /* "surprise": non-monotone offsets */
        }
      }
    }
    if (pages) {
      if (np >= max_pages) return 4;
      pages[4 * np + 0] = left;
      pages[4 * np + 1] = (uint32_t)right;
      pages[4 * np + 2] = (uint32_t)(right - left + 1);
      pages[4 * np + 3] = (uint32_t)(pos[right + 1] - pos[left]);
    }
    ++np;
    left = (uint32_t)right + 1;
  }
  *num_pages = np;
  return 0;
}

/* Content of one page file before zero fill (tools/bin2pagecsr.cpp:221-242):
 * Node{vtx, idx - idx[left], len} per vertex, then the neighbours (or
 * {ngr, wgh} pairs). */
ORACLE_API void oracle_fill_page(const uint64_t* pos, const uint32_t* ngh, const uint32_t* wgh,
                                 uint32_t left, uint32_t right, int weighted, uint32_t* out) {
  const uint32_t nodenum = right - left + 1;
  for (uint32_t id = left; id <= right; ++id) {
    const uint32_t off = id - left;
    out[3 * (uint64_t)off + 0] = id;
    out[3 * (uint64_t)off + 1] = (uint32_t)(pos[id] - pos[left]);
    out[3 * (uint64_t)off + 2] = (uint32_t)(pos[(uint64_t)id + 1] - pos[id]);
  }
  uint32_t* e = out + 3 * (uint64_t)nodenum;
  const uint64_t edgenum = pos[(uint64_t)right + 1] - pos[left];
  for (uint64_t k = 0; k < edgenum; ++k) {
    if (weighted) {
      e[2 * k] = ngh[pos[left] + k];
      e[2 * k + 1] = wgh ? wgh[pos[left] + k] : 0u;
    } else {
      e[k] = ngh[pos[left] + k];
    }
  }
}

/* ------------------------------------------------------- page counts */

/* get_page_info + get_total_info (src/gpu_engine.cuh:397-456). */
ORACLE_API void oracle_page_counts(const uint32_t* pages, uint32_t np, const uint8_t* stat,
                                   const uint32_t* degree, int weighted, uint32_t* nodes,
                                   uint32_t* datas, uint64_t* totals) {
  uint64_t tn = 0, td = 0;
  for (uint32_t p = 0; p < np; ++p) {
    const uint32_t left = pages[4 * p], right = pages[4 * p + 1];
    uint32_t cn = 0, cd = 0;
    for (uint32_t id = left; id <= right; ++id) {
      cn += stat[id];
      cd += stat[id] * (weighted ? (degree[id] * 2 + 3) : (degree[id] + 3));
    }
    nodes[p] = cn;
    datas[p] = cd;
    tn += cn;
    td += cd;
  }
  totals[0] = tn;
  totals[1] = td;
}

/* compress_stat (src/compress.h:18-40): the active Nodes of an unweighted
 * page, re-indexed, followed by their adjacency.  Returns the node count. */
ORACLE_API uint32_t oracle_compress_page(const uint32_t* src, uint32_t* dst, const uint8_t* stat,
                                         uint32_t left, uint32_t srcnodenum,
                                         uint32_t dstnodenum) {
  const uint32_t* esrc = src + 3 * (uint64_t)srcnodenum;
  uint32_t* edst = dst + 3 * (uint64_t)dstnodenum;
  uint64_t offset = 0;
  uint32_t counter = 0;
  for (uint32_t i = 0; i < srcnodenum; ++i) {
    const uint32_t node = i + left;
    if (stat[node]) {
      const uint32_t idx = src[3 * (uint64_t)i + 1], len = src[3 * (uint64_t)i + 2];
      dst[3 * (uint64_t)counter + 0] = node;
      dst[3 * (uint64_t)counter + 1] = (uint32_t)offset;
      dst[3 * (uint64_t)counter + 2] = len;
      memcpy(edst + offset, esrc + idx, (size_t)len * 4);
      offset += len;
      ++counter;
    }
  }
  return counter;
}

/* ----------------------------------------------------- chunk masks */

/* get_max_continus_zero (src/tools.h:10-27): longest run of zero bits. */
ORACLE_API int oracle_max_zero_run(int mark) {
  int local = 0, best = 0;
  unsigned m = (unsigned)mark;
  for (int i = 0; i < 32; ++i) {
    if (!(m & 1u)) {
      ++local;
    } else {
      if (local > best) best = local;
      local = 0;
    }
    m >>= 1;
  }
  if (local > best) best = local;
  return best;
}

/* get_fit_continus_zero (src/tools.h:29-64): start of the shortest zero run
 * of length >= required (first such run on ties; the scan stops early at a
 * run of exactly `required`), -1 if none. */
ORACLE_API int oracle_fit_zero_run(int mark, int required) {
  int local = 0, localoff = 0;
  int fit = oracle_max_zero_run(mark);
  int fitoff = -1;
  unsigned m = (unsigned)mark;
  if (fit < required) return -1;
  for (int i = 0; i < 32; ++i) {
    if (!(m & 1u)) {
      ++local;
    } else {
      if (local >= required && local <= fit) {
        fit = local;
        fitoff = localoff;
      }
      if (local == required) break;
      local = 0;
      localoff = i + 1;
    }
    m >>= 1;
  }
  if (local >= required && local <= fit) {
    fit = local;
    fitoff = localoff;
  }
  return fitoff;
}

/* set_continus_one / set_continus_zero (src/tools.h:66-85). */
ORACLE_API int oracle_set_run(int mark, int off, int number) {
  unsigned m = (unsigned)mark;
  for (int i = 0; i < number; ++i)
    if (off + i < 32) m |= 1u << (off + i);
  return (int)m;
}

ORACLE_API int oracle_clear_run(int mark, int off, int number) {
  unsigned m = (unsigned)mark;
  for (int i = 0; i < number; ++i)
    if (off + i < 32) m &= ~(1u << (off + i));
  return (int)m;
}

/* Device page pool size (src/app.cuh:281-286, 299, 323-324); -1 when the
 * vertex arrays alone exceed the budget ("vertex data out of gpu memory"). */
ORACLE_API int64_t oracle_pool_size(int64_t max_memory_mib, uint32_t n, uint32_t list_size,
                                    uint64_t pagesize_words) {
  int64_t m = max_memory_mib * 1024 * 1024;
  m -= 4ll * n;
  m -= 1ll * n;
  m -= 4ll * n;
  m -= 4ll * list_size;
  m -= 4ll * list_size;
  m -= 16ll * list_size;
  if (m <= (int64_t)pagesize_words * 4) return -1;
  int64_t pool = m / ((int64_t)pagesize_words * 4);
  return pool > list_size ? list_size : pool;
}

/* -------------------------------------------------------- generators */

static uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* i-th output of SplitMix64 started at state `key`. */
ORACLE_API uint64_t oracle_splitmix64(uint64_t key, uint64_t i) {
  return mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull);
}

typedef struct {
  uint64_t mul[4], add[4], mask;
  uint32_t shift;
} perm_t;

static perm_t perm_make(uint64_t seed, uint32_t bits) {
  perm_t k;
  uint64_t pk = mix64(seed ^ 0x7065726D7574650Aull);
  k.mask = bits >= 64 ? ~0ull : ((1ull << bits) - 1);
  k.shift = (bits + 1) / 2;
  if (k.shift == 0) k.shift = 1;
  for (int r = 0; r < 4; ++r) {
    k.mul[r] = (oracle_splitmix64(pk, 2 * r) | 1ull) & k.mask;
    if (k.mul[r] == 0) k.mul[r] = 1;
    k.add[r] = oracle_splitmix64(pk, 2 * r + 1) & k.mask;
  }
  return k;
}

static uint64_t perm_apply(const perm_t* k, uint64_t x, uint64_t n) {
  do {
    for (int r = 0; r < 4; ++r) {
      x = (x * k->mul[r] + k->add[r]) & k->mask;
      x ^= x >> k->shift;
    }
  } while (x >= n);
  return x;
}

ORACLE_API uint64_t oracle_perm(uint64_t seed, uint32_t bits, uint64_t x, uint64_t n) {
  perm_t k = perm_make(seed, bits);
  return perm_apply(&k, x, n);
}

/* Raw R-MAT edge e (before relabelling): draw l (0-based) is the low half
 * (even l) or high half (odd l) of word e*ceil(S/2) + l/2; it selects the
 * quadrant of bit S-1-l: r < A: (0,0); r < AB: (0,1); r < ABC: (1,0);
 * else (1,1). */
ORACLE_API void oracle_rmat_edge(uint64_t seed, uint64_t e, uint32_t scale, uint32_t A, uint32_t AB,
                                 uint32_t ABC, uint32_t* u, uint32_t* v) {
  const uint64_t words = (scale + 1) / 2;
  uint32_t uu = 0, vv = 0;
  for (uint32_t l = 0; l < scale; ++l) {
    const uint64_t w = oracle_splitmix64(seed, e * words + l / 2);
    const uint32_t r = (l % 2 == 0) ? (uint32_t)(w & 0xFFFFFFFFull) : (uint32_t)(w >> 32);
    const uint32_t bit = 1u << (scale - 1 - l);
    int row, col;
    if (r < A) {
      row = 0;
      col = 0;
    } else if (r < AB) {
      row = 0;
      col = 1;
    } else if (r < ABC) {
      row = 1;
      col = 0;
    } else {
      row = 1;
      col = 1;
    }
    if (row) uu |= bit;
    if (col) vv |= bit;
  }
  *u = uu;
  *v = vv;
}

static int cmp_u64(const void* a, const void* b) {
  const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
  return x < y ? -1 : (x > y ? 1 : 0);
}

/* keys (u<<32|v) -> sorted, deduplicated, self-loop-free CSR. */
static int keys_to_csr(uint64_t* keys, uint64_t m, uint32_t n, uint64_t* offsets, uint32_t* ngh,
                       uint64_t cap, uint64_t* num_edges) {
  qsort(keys, (size_t)m, sizeof(uint64_t), cmp_u64);
  memset(offsets, 0, sizeof(uint64_t) * ((uint64_t)n + 1));
  uint64_t E = 0;
  for (uint64_t i = 0; i < m; ++i) {
    if (i > 0 && keys[i] == keys[i - 1]) continue;
    const uint32_t u = (uint32_t)(keys[i] >> 32), v = (uint32_t)keys[i];
    if (u == v) continue;
    if (E >= cap) return 3;
    ngh[E++] = v;
    offsets[(uint64_t)u + 1]++;
  }
  for (uint32_t v = 0; v < n; ++v) offsets[v + 1] += offsets[v];
  *num_edges = E;
  return 0;
}

ORACLE_API int oracle_rmat_csr(uint32_t scale, uint32_t ef, uint32_t A, uint32_t AB, uint32_t ABC,
                               uint64_t seed, uint64_t* offsets, uint32_t* ngh, uint64_t cap,
                               uint64_t* num_edges) {
  if (scale == 0 || scale > 31) return 1;
  const uint64_t n = 1ull << scale, m = (uint64_t)ef * n;
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
  if (!keys) return 2;
  perm_t k = perm_make(seed, scale);
  for (uint64_t e = 0; e < m; ++e) {
    uint32_t u, v;
    oracle_rmat_edge(seed, e, scale, A, AB, ABC, &u, &v);
    u = (uint32_t)perm_apply(&k, u, n);
    v = (uint32_t)perm_apply(&k, v, n);
    keys[e] = ((uint64_t)u << 32) | v;
  }
  int rc = keys_to_csr(keys, m, (uint32_t)n, offsets, ngh, cap, num_edges);
  free(keys);
  return rc;
}

/* Directed G(n, m): endpoints floor(w * n / 2^64) of words 2e and 2e+1. */
ORACLE_API int oracle_er_csr(uint32_t n, uint64_t m, uint64_t seed, uint64_t* offsets,
                             uint32_t* ngh, uint64_t cap, uint64_t* num_edges) {
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
  if (!keys) return 2;
  for (uint64_t e = 0; e < m; ++e) {
    const uint64_t wu = oracle_splitmix64(seed, 2 * e), wv = oracle_splitmix64(seed, 2 * e + 1);
    const uint32_t u = (uint32_t)(((unsigned __int128)wu * n) >> 64);
    const uint32_t v = (uint32_t)(((unsigned __int128)wv * n) >> 64);
    keys[e] = ((uint64_t)u << 32) | v;
  }
  int rc = keys_to_csr(keys, m, n, offsets, ngh, cap, num_edges);
  free(keys);
  return rc;
}

/* Power-law out-degrees (continuous Pareto with density ~ x^-alpha for
 * x >= x_min, floored and capped) and uniform destinations. */
#define ZIPF_DST_SALT 0x7A6970662D647374ull /* "zipf-dst" */

ORACLE_API uint32_t oracle_zipf_degree(uint64_t seed, uint32_t v, double x_min, double alpha,
                                       uint32_t cap) {
  const double u = (double)(oracle_splitmix64(seed, v) >> 11) * (1.0 / 9007199254740992.0);
  const double x = x_min * pow(1.0 - u, -1.0 / (alpha - 1.0));
  return x >= (double)cap ? cap : (uint32_t)x;
}

ORACLE_API int oracle_zipf_csr(uint32_t n, double x_min, double alpha, uint32_t cap, uint64_t seed,
                               uint64_t* offsets, uint32_t* ngh, uint64_t ngh_cap,
                               uint64_t* num_edges) {
  uint64_t m = 0;
  for (uint32_t v = 0; v < n; ++v) m += oracle_zipf_degree(seed, v, x_min, alpha, cap);
  uint64_t* keys = (uint64_t*)malloc(sizeof(uint64_t) * (m ? m : 1));
  if (!keys) return 2;
  uint64_t e = 0;
  for (uint32_t v = 0; v < n; ++v) {
    const uint32_t d = oracle_zipf_degree(seed, v, x_min, alpha, cap);
    for (uint32_t j = 0; j < d; ++j, ++e) {
      const uint64_t w = oracle_splitmix64(seed ^ ZIPF_DST_SALT, e);
      const uint32_t t = (uint32_t)(((unsigned __int128)w * n) >> 64);
      keys[e] = ((uint64_t)v << 32) | t;
    }
  }
  int rc = keys_to_csr(keys, m, n, offsets, ngh, ngh_cap, num_edges);
  free(keys);
  return rc;
}

/* --------------------------------------------------------------- MLP */

/* fp32 forward of the 16-64-64-2 ReLU MLP on a batch; parameters in the
 * layout W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2]. */
ORACLE_API void oracle_mlp_forward(const float* p, const float* x, uint32_t batch, float* y,
                                   float* h1, float* h2) {
  const float *W1 = p, *b1 = W1 + 64 * 16, *W2 = b1 + 64, *b2 = W2 + 64 * 64, *W3 = b2 + 64,
              *b3 = W3 + 2 * 64;
  for (uint32_t s = 0; s < batch; ++s) {
    const float* xs = x + 16 * (uint64_t)s;
    float* a1 = h1 + 64 * (uint64_t)s;
    float* a2 = h2 + 64 * (uint64_t)s;
    for (int j = 0; j < 64; ++j) {
      float a = b1[j];
      for (int i = 0; i < 16; ++i) a += W1[j * 16 + i] * xs[i];
      a1[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < 64; ++j) {
      float a = b2[j];
      for (int i = 0; i < 64; ++i) a += W2[j * 64 + i] * a1[i];
      a2[j] = a > 0 ? a : 0;
    }
    for (int j = 0; j < 2; ++j) {
      float a = b3[j];
      for (int i = 0; i < 64; ++i) a += W3[j * 64 + i] * a2[i];
      y[2 * (uint64_t)s + j] = a;
    }
  }
}

/* Loss = mean over the batch of
 *   softplus(z0) - t0*z0            (logistic loss, t0 in {0,1})
 *   + w1 * (z1 - t1)^2              (squared error)
 * and its gradient w.r.t. every parameter (same layout as p). */
ORACLE_API float oracle_mlp_loss_grad(const float* p, const float* x, const float* t,
                                      uint32_t batch, float w1, float* grad) {
  const float *W1 = p, *b1 = W1 + 64 * 16, *W2 = b1 + 64, *b2 = W2 + 64 * 64, *W3 = b2 + 64;
  float *gW1 = grad, *gb1 = gW1 + 64 * 16, *gW2 = gb1 + 64, *gb2 = gW2 + 64 * 64, *gW3 = gb2 + 64,
        *gb3 = gW3 + 2 * 64;
  const int np = 64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2;
  memset(grad, 0, sizeof(float) * np);
  double loss = 0;
  float h1[64], h2[64], y[2];
  for (uint32_t s = 0; s < batch; ++s) {
    oracle_mlp_forward(p, x + 16 * (uint64_t)s, 1, y, h1, h2);
    const float z0 = y[0], z1 = y[1], t0 = t[2 * s], t1 = t[2 * s + 1];
    const double sp = z0 > 0 ? z0 + log1p(exp(-z0)) : log1p(exp(z0));
    loss += sp - t0 * z0 + w1 * (z1 - t1) * (z1 - t1);
    const float sig = (float)(1.0 / (1.0 + exp(-(double)z0)));
    const float dz[2] = {(sig - t0) / batch, 2.f * w1 * (z1 - t1) / batch};
    float d2[64], d1[64];
    for (int i = 0; i < 64; ++i) d2[i] = 0;
    for (int j = 0; j < 2; ++j) {
      gb3[j] += dz[j];
      for (int i = 0; i < 64; ++i) {
        gW3[j * 64 + i] += dz[j] * h2[i];
        d2[i] += W3[j * 64 + i] * dz[j];
      }
    }
    for (int i = 0; i < 64; ++i) {
      d2[i] = h2[i] > 0 ? d2[i] : 0;
      d1[i] = 0;
    }
    for (int j = 0; j < 64; ++j) {
      gb2[j] += d2[j];
      for (int i = 0; i < 64; ++i) {
        gW2[j * 64 + i] += d2[j] * h1[i];
        d1[i] += W2[j * 64 + i] * d2[j];
      }
    }
    for (int j = 0; j < 64; ++j) {
      const float d = h1[j] > 0 ? d1[j] : 0;
      gb1[j] += d;
      for (int i = 0; i < 16; ++i) gW1[j * 16 + i] += d * x[16 * (uint64_t)s + i];
    }
  }
  (void)b1;
  (void)b2;
  return (float)(loss / batch);
}
