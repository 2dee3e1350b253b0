/*
 * scaph_engine.h — C-ABI of the host BFS engine (libscaph_host.so).
 *
 * This is the binding surface for callers that are not C++ (ctypes, cgo,
 * JNI ...).  It exposes the reference's App lifecycle (src/app.cuh:120-135):
 * construct from a command line (App(argc, argv)), load_graph(), run() one
 * BFS, read the levels (h_value, src/app.cuh:262-264) and check().
 * Every call returns 0 on success or a nonzero status; when `err` is
 * non-NULL a message is written there (truncated to errlen).  Nothing
 * exits the process.
 */
#ifndef SCAPH_ENGINE_H
#define SCAPH_ENGINE_H

#include <stddef.h>
#include <stdint.h>

#include "scaph_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct scaph_engine scaph_engine;

typedef struct scaph_bfs_stats {
  uint32_t src;
  uint32_t depth;             /* levels run */
  double ms;                  /* t_BFS: init .. levels final on device */
  uint64_t n_reached;         /* vertices reached (n_c) */
  uint64_t m_reached;         /* sum of their out-degrees (m_c) */
  uint64_t page_execs;        /* paged mode: page kernel launches */
  uint64_t page_execs_cached; /* ... of pages already resident */
  uint64_t pages_full;        /* full-page copies */
  uint64_t pages_compacted;   /* compacted (cold) page copies */
  uint64_t bytes_h2d;         /* host-to-device page bytes */
  double h2d_ms;              /* copy-stream busy time */
  uint64_t pages_prefetched;  /* whole-page copies made ahead for the next level */
  uint64_t pages_direct;      /* cold pages pushed straight from pinned host memory */
} scaph_bfs_stats;

typedef struct scaph_level_record {
  uint32_t level;
  int32_t dir; /* 0 push, 1 pull */
  uint64_t n_f, m_f, m_u;
  uint64_t n_next, m_next, scanned, examined;
  float ms;        /* step time including host turnaround */
  float kernel_ms; /* expand kernels of the step alone (events around them) */
  float page_hist[12];
  /* paged mode (zero when resident): pages with active vertices, those
   * already in the pool at the level's start, whole-page and compacted
   * copies, pages copied ahead for the next level, cold pages pushed from
   * pinned host memory, host-to-device bytes and the copy stream's busy
   * time in this level */
  uint32_t pages_active, pages_cached, pages_full, pages_compacted, pages_prefetched, pages_direct;
  uint64_t bytes_h2d;
  float h2d_ms;
  float logit; /* learned schedule: the model's pull logit behind this level's direction */
  uint32_t sparse; /* 1: a pull over the list of the few unvisited vertices left */
  uint32_t pad;
} scaph_level_record;

/* App(argc, argv): flags in gflags syntax (argv[0] is the program name). */
int scaph_engine_create(int argc, const char* const* argv, scaph_engine** out, char* err,
                        size_t errlen);
void scaph_engine_destroy(scaph_engine* e);
/* load_graph(): -graphfile, else -gen. */
int scaph_engine_load(scaph_engine* e, char* err, size_t errlen);
/* Load a flat CSR from host memory (offsets u64[n+1], ngh u32[offsets[n]]). */
int scaph_engine_load_csr(scaph_engine* e, uint32_t n, const uint64_t* offsets,
                          const uint32_t* ngh, int keep_host, char* err, size_t errlen);
/* Build an R-MAT graph directly in device memory. */
int scaph_engine_gen_rmat_device(scaph_engine* e, uint32_t scale, uint32_t edge_factor, double a,
                                 double b, double c, uint64_t seed, int keep_host, char* err,
                                 size_t errlen);
/* Build an Erdos-Renyi G(n, m) graph directly in device memory. */
int scaph_engine_gen_er_device(scaph_engine* e, uint32_t n, uint64_t m, uint64_t seed,
                               int keep_host, char* err, size_t errlen);
/* Power-law digraph: out-degree of v = min(cap, floor(x_min * (1 -
 * u_v)^(-1/(alpha - 1)))) with u_v from SplitMix64 word v of seed (53 bits),
 * uniform destinations, self-loops and duplicates removed.  Built on the
 * device, then loaded like a host CSR: beyond -max_memory it runs out of
 * core.  No reference counterpart (the reference's graphs come from files). */
int scaph_engine_gen_zipf(scaph_engine* e, uint32_t n, double x_min, double alpha, uint32_t cap,
                          uint64_t seed, char* err, size_t errlen);
int scaph_engine_info(scaph_engine* e, uint32_t* n, uint64_t* m, uint32_t* num_pages,
                      int* resident, int* has_pull);
/* Page slots of the loaded graph's device pool (scaph_host_plan_pool of its
 * -max_memory; equal to num_pages when every page is resident) and the
 * device bytes the engine holds for it. */
int scaph_engine_pool(scaph_engine* e, int64_t* slots, uint64_t* device_bytes);
/* run(): one BFS from src.  schedule: NULL keeps the -schedule flag, else
 * "fixed" | "push" | "pull" | "learned". */
int scaph_engine_bfs(scaph_engine* e, uint32_t src, scaph_bfs_stats* out, char* err,
                     size_t errlen);
int scaph_engine_set_schedule(scaph_engine* e, const char* name, double alpha, double beta,
                              char* err, size_t errlen);
/* Paged mode with -page_predict=exact: the u32 levels (n of them) of an
 * earlier BFS from the source about to run, so that the page cache knows
 * every page's future activity (the bound for any page-activity
 * predictor; the levels computed are unaffected).  No reference
 * counterpart. */
int scaph_engine_set_page_levels(scaph_engine* e, const uint32_t* levels, uint32_t n, char* err,
                                 size_t errlen);
/* The reference's other App plugins on the resident graph (app/wcc.cu,
 * app/sssp.cu): values until no update, read with scaph_engine_levels.
 * WCC: value[v] = the least vertex id reaching v (v itself included), as
 * app/wcc.cu's label push computes on directed edges.  SSSP: shortest-path
 * length from src over the edge weights, 0x7FFFFFFF (MAXPATH) if
 * unreached; weights must keep distances below 2^31.  stats: depth =
 * rounds, m_reached = edges relaxed over all rounds, ms = wall time. */
int scaph_engine_wcc(scaph_engine* e, scaph_bfs_stats* out, char* err, size_t errlen);
int scaph_engine_sssp(scaph_engine* e, uint32_t src, scaph_bfs_stats* out, char* err,
                      size_t errlen);
/* Edge weights in CSR edge order (_csr.wgh, src/graph.h:56-60); count = m.
 * Loading a CSR that carries weights sets them already. */
int scaph_engine_set_weights(scaph_engine* e, const uint32_t* wgh, uint64_t count, char* err,
                             size_t errlen);
/* Learned schedule weights: SCAPH_MLP_NPARAM bf16 values + feature scale. */
int scaph_engine_set_model(scaph_engine* e, const uint16_t* params_bf16, const float* feat_mean,
                           const float* feat_scale, float prefetch_gain, char* err,
                           size_t errlen);
int scaph_engine_level_records(scaph_engine* e, scaph_level_record* out, uint32_t max,
                               uint32_t* count);

/* Training traces for the learned schedule: a BFS from src in which every
 * level runs both directions from the same state (min of `reps` timings
 * each) and continues with the faster one.  One record per level: the raw
 * features of the frontier entering it (the decision kernel's), both step
 * times, and the fraction of pages active at the next level. */
typedef struct scaph_explore_record {
  uint64_t n_f, m_f, n_next;
  uint32_t level;
  float t_push_ms, t_pull_ms, pages_active_next;
  float x[16];
} scaph_explore_record;
int scaph_engine_explore(scaph_engine* e, uint32_t src, int reps, scaph_explore_record* out,
                         uint32_t max, uint32_t* count, char* err, size_t errlen);
int scaph_engine_levels(scaph_engine* e, uint32_t* host_out);

/* Multi-source BFS (no reference counterpart): up to 256 queries over the
 * loaded resident graph in one traversal (scaph_ms_* kernels).  A NULL tune
 * uses the last autotuned choice, else the defaults.  Levels of each query
 * equal a single-source BFS from its source. */
typedef struct scaph_ms_tune {
  uint32_t push_div, push_mdiv; /* push while n_active*push_div < n and m_active*push_mdiv < m */
  uint32_t pull_block;          /* 256, 512 or 1024 */
  uint32_t use_graph;           /* replay hipGraph-captured level segments */
  uint32_t segment;             /* levels per captured segment */
  uint32_t pull_grid;           /* cap on pull blocks (0 = one per block of vertices) */
  uint32_t persistent;          /* every level in one persistent cooperative launch */
} scaph_ms_tune;
typedef struct scaph_ms_stats {
  double ms;                    /* device time: init and every level */
  uint32_t depth, queries;
  unsigned long long pairs;     /* (vertex, query) pairs reached */
} scaph_ms_stats;
int scaph_engine_msbfs(scaph_engine* e, const uint32_t* sources, uint32_t count,
                       const scaph_ms_tune* tune, scaph_ms_stats* out, char* err, size_t errlen);
/* uint32 levels of query `query` of the last multi-source run. */
int scaph_engine_msbfs_levels(scaph_engine* e, uint32_t query, uint32_t* host_out, char* err,
                              size_t errlen);
/* Per-level counters of the last run: records 0..depth. */
int scaph_engine_msbfs_counters(scaph_engine* e, scaph_ms_counters* out, uint32_t max,
                                uint32_t* count);
/* Times each candidate tuning on these sources and keeps the fastest for
 * later NULL-tune runs; `log` (may be NULL) receives up to log_max
 * (tune, ms) pairs. */
int scaph_engine_msbfs_autotune(scaph_engine* e, const uint32_t* sources, uint32_t count,
                                scaph_ms_tune* best, double* best_ms, scaph_ms_tune* log_tunes,
                                double* log_ms, uint32_t log_max, uint32_t* log_count,
                                char* err, size_t errlen);
int scaph_engine_validate(scaph_engine* e, uint32_t src, uint64_t* out2);
int scaph_engine_out_degrees(scaph_engine* e, uint32_t* host_out);
/* Copy of the host CSR kept at load (offsets n+1, ngh m); 1 if none kept. */
int scaph_engine_host_csr(scaph_engine* e, uint64_t* offsets, uint32_t* ngh);
/* check(): CPU BFS from src compared with the device levels; *errors gets
 * the number of mismatching vertices (reference-capped and exact forms). */
int scaph_engine_check(scaph_engine* e, uint32_t src, uint64_t* errors, char* err, size_t errlen);
/* Device-side graph handles (for tests that drive kernels directly). */
int scaph_engine_graph_desc(scaph_engine* e, int transposed, void* desc_out /* scaph_graph_desc* */);
int scaph_engine_device_degree(scaph_engine* e, const uint32_t** degree);

/* Host-side tools, no device needed. */
int scaph_host_page_cuts(const uint64_t* offsets, uint32_t n, uint64_t pagesize_words,
                         int weighted, void* pages_out /* scaph_page*, may be NULL */,
                         uint32_t max_pages, uint32_t* num_pages);
int scaph_host_gen_rmat(uint32_t scale, uint32_t edge_factor, double a, double b, double c,
                        uint64_t seed, uint64_t* offsets /* n+1 */, uint32_t* ngh /* cap */,
                        uint64_t ngh_cap, uint64_t* num_edges);
int scaph_host_gen_er(uint32_t n, uint64_t m, uint64_t seed, uint64_t* offsets, uint32_t* ngh,
                      uint64_t ngh_cap, uint64_t* num_edges);
int scaph_host_write_graph(const char* base, uint32_t n, const uint64_t* offsets,
                           const uint32_t* ngh, uint64_t pagesize_words, int idx64,
                           int write_wgh, char* err, size_t errlen);
int scaph_host_cpu_bfs(uint32_t n, const uint64_t* offsets, const uint32_t* ngh, uint32_t src,
                       uint32_t inf, uint32_t* levels);
/* Device page-pool slots for -max_memory=max_memory_mib (the budget of
 * App::load_graph, src/app.cuh:281-286, 323-324): the budget less 9 bytes
 * per vertex and 24 per page, in whole pagesize_words slots, at most one per
 * page.  *slots = -1 when the vertex arrays leave no more than one slot,
 * where the reference prints "vertex data out of gpu memory. exit!"
 * (src/app.cuh:310-313).  The engine sizes its pool with this function. */
int scaph_host_plan_pool(int64_t max_memory_mib, uint32_t n, uint32_t num_pages,
                         uint64_t pagesize_words, int64_t* slots);

/* Batch of small graphs (no reference counterpart): `count` R-MAT graphs
 * generated on the device (graph g: scale and edge factor uniform in the
 * given ranges, its own seed and source, from SplitMix64 words of `seed`),
 * all queried by one persistent launch. */
typedef struct scaph_batch scaph_batch;
typedef struct scaph_batch_tune {
  uint32_t block;       /* threads per workgroup: 256, 512, 1024 */
  uint32_t wgs_per_cu;  /* persistent workgroups per compute unit */
  uint32_t use_graph;   /* replay a captured hipGraph */
} scaph_batch_tune;
int scaph_batch_create_rmat(uint32_t count, uint32_t min_scale, uint32_t max_scale,
                            uint32_t min_ef, uint32_t max_ef, uint64_t seed, int device,
                            scaph_batch** out, char* err, size_t errlen);
void scaph_batch_destroy(scaph_batch* b);
int scaph_batch_info(scaph_batch* b, uint32_t* count, unsigned long long* vertices,
                     unsigned long long* edges);
int scaph_batch_graph_info(scaph_batch* b, uint32_t g, scaph_batch_entry* out);
/* One batch of queries (every graph from its source); NULL tune = the
 * autotuned choice (or the defaults).  *ms = device time. */
int scaph_batch_run(scaph_batch* b, const scaph_batch_tune* tune, double* ms, char* err,
                    size_t errlen);
int scaph_batch_levels(scaph_batch* b, uint32_t g, uint32_t* host_out /* n_g */);
/* Graph g's CSR with local ids; ngh may be NULL to get the edge count. */
int scaph_batch_csr(scaph_batch* b, uint32_t g, uint64_t* offsets /* n_g + 1 */, uint32_t* ngh,
                    unsigned long long cap, unsigned long long* num_edges);
int scaph_batch_autotune(scaph_batch* b, scaph_batch_tune* best, double* best_ms,
                         scaph_batch_tune* log_tunes, double* log_ms, uint32_t log_max,
                         uint32_t* log_count, char* err, size_t errlen);

/* Page-slot chunk masks (src/tools.h:10-85). */
int scaph_host_max_zero_run(uint32_t mask);
int scaph_host_fit_zero_run(uint32_t mask, int required);
uint32_t scaph_host_set_run(uint32_t mask, int off, int number);
uint32_t scaph_host_clear_run(uint32_t mask, int off, int number);

#ifdef __cplusplus
}
#endif
#endif
