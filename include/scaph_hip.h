/*
 * scaph_hip.h — C-ABI of the gfx950 BFS kernel library (libscaph_hip.so).
 *
 * Everything here is plain C: device buffers are passed as raw pointers,
 * sizes as integers, streams/events as opaque handles.  Every entry point
 * returns an int status (0 = success, otherwise the hipError_t value) and
 * never exits the process; scaph_error_string() turns a status into text.
 * The caller owns every buffer; nothing is allocated implicitly except where
 * an entry point says so (the graph builders allocate their own temporary
 * sort storage and free it before returning).
 *
 * Reference interfaces replaced (paths relative to the Scaph sources):
 *   - cudaMalloc / cudaMallocHost / cudaMemcpyAsync / cudaStreamQuery call
 *     sites: src/app.cuh:288-339, 191-197, 371-393, 421, 447; src/graph.h:100
 *   - bfs_init / setsrc:           app/bfs.cu:20-37, 58-65
 *   - gpu_push_csr:                src/gpu_engine.cuh:41-169
 *   - cond / update (writeMin):    app/bfs_kernel.cuh:6-16, src/gpu_tools.cuh:53-64
 *   - get_page_info/get_total_info src/gpu_engine.cuh:397-456
 *   - compress_stat (cold pages):  src/compress.h:18-56
 * Entry points without a reference counterpart (pull expand, bitmap
 * conversions, generators, MLP, batched BFS) say so in their comment.
 */
#ifndef SCAPH_HIP_H
#define SCAPH_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCAPH_UNREACHED 0xFFFFFFFFu /* level of a vertex no path reaches */
#define SCAPH_STATUS_NOT_READY 600  /* == hipErrorNotReady (stream/event busy) */
#define SCAPH_STATUS_INVALID 1      /* == hipErrorInvalidValue */
#define SCAPH_STATUS_SCRATCH_TOO_SMALL 0x5C500001 /* input beyond a no-alloc entry point's scratch */

/* On-disk / in-memory page record: src/page.h:22-29 (Page<uint32>). */
typedef struct scaph_page {
  uint32_t left;    /* first vertex of the page */
  uint32_t right;   /* last vertex of the page (inclusive) */
  uint32_t nodenum; /* right - left + 1 */
  uint32_t edgenum; /* out-edges stored in the page */
} scaph_page;

/* A paged CSR resident (wholly or partly) in device memory.
 * Page p's buffer starts at page_ptr[p] and holds, as in the files written
 * by tools/bin2pagecsr.cpp:221-242, Node{vtx,idx,len}[nodenum] (3 u32 words
 * each, idx relative to the page's edge array) followed by the u32 neighbour
 * IDs.  page_ptr and pages are DEVICE arrays of num_pages entries. */
typedef struct scaph_graph_desc {
  const uint64_t* page_ptr;  /* device: u32* address of each page buffer */
  const scaph_page* pages;   /* device: page list */
  uint32_t num_pages;
  uint32_t num_vertices;
} scaph_graph_desc;

typedef struct scaph_device_props {
  char name[256];
  char arch[64];
  int compute_units;
  int xcds;                 /* 0 when the runtime does not report it */
  int wave_size;
  size_t lds_per_block;
  size_t total_mem;
  size_t l2_bytes;
  int clock_khz;
  int mem_clock_khz;
  int mem_bus_width;
} scaph_device_props;

/* ---------------------------------------------------------------- runtime */
const char* scaph_error_string(int status);
int scaph_init(int device);
int scaph_device_count(int* count);
int scaph_get_device_props(int device, scaph_device_props* out);
/* out[b] = the XCC (XCD) that ran workgroup b of a `blocks`-workgroup
 * launch, read from HW_REG_XCC_ID.  The sliced push expand assumes the
 * round-robin dispatch this shows (workgroup b on XCD (s + b) mod 8 for a
 * start s that varies between launches); the assumption affects speed
 * only. */
int scaph_xcc_ids(uint32_t* out, uint32_t blocks, void* stream);
/* out[2b] = XCC and out[2b+1] = HW_REG_HW_ID (CU id bits 11:8, shader
 * array bit 12, shader engine bits 15:13) of the wave running workgroup b;
 * shows which CUs a stream's launches reach. */
int scaph_hw_ids(uint32_t* out, uint32_t blocks, void* stream);
int scaph_stream_create(void** stream);
/* A stream whose kernels run on all but `reserved_cus` CUs (0 = all; must
 * be below the device's CU count).  Replaces the reference's trick kernel
 * (src/app.cuh:210-217, src/gpu_engine.cuh:464-475), which pins
 * FLAGS_tricknum SMs with a spinning block so that BFS runs on fewer; a CU
 * mask gives the same reduced machine without a spinning kernel. */
int scaph_stream_create_cu_limited(void** stream, uint32_t reserved_cus);
int scaph_stream_destroy(void* stream);
int scaph_stream_query(void* stream); /* 0 = idle, SCAPH_STATUS_NOT_READY = busy */
int scaph_stream_sync(void* stream);
int scaph_stream_wait_event(void* stream, void* event);
int scaph_event_create(void** event);
int scaph_event_destroy(void* event);
int scaph_event_record(void* event, void* stream);
int scaph_event_query(void* event);
int scaph_event_sync(void* event);
int scaph_event_elapsed_ms(void* start, void* stop, float* ms);
int scaph_device_sync(void);
int scaph_mem_info(size_t* free_bytes, size_t* total_bytes);

/* ----------------------------------------------------------------- memory */
int scaph_dev_alloc(void** ptr, size_t bytes);
int scaph_dev_free(void* ptr);
int scaph_host_pinned_alloc(void** ptr, size_t bytes);
int scaph_host_pinned_free(void* ptr);
/* The address at which kernels read memory from scaph_host_pinned_alloc
 * (each access crosses PCIe).  No reference counterpart: the reference only
 * copies pages (src/app.cuh:371-393). */
int scaph_host_device_pointer(const void* host_ptr, void** device_ptr);
int scaph_copy_h2d_async(void* dst, const void* src, size_t bytes, void* stream);
int scaph_copy_d2h_async(void* dst, const void* src, size_t bytes, void* stream);
int scaph_copy_d2d_async(void* dst, const void* src, size_t bytes, void* stream);
/* Fill n u32 words with value (a kernel, safe to capture in a hipGraph). */
int scaph_fill_u32(uint32_t* dst, uint32_t value, size_t n, void* stream);
/* *out = the sum of src[0..n) (u64), every word read once.  With
 * scaph_fill_u32 it gives kernels of known read and write bytes, against
 * which rocprof's FETCH_SIZE / WRITE_SIZE are scaled.  No reference
 * counterpart. */
int scaph_checksum_u32(const uint32_t* src, size_t n, unsigned long long* out, void* stream);
/* Host memory that kernels write directly (pinned, mapped, coherent):
 * *host_ptr for the CPU, *device_ptr for kernels; free with
 * scaph_host_pinned_free(host_ptr). */
int scaph_host_mapped_alloc(void** host_ptr, void** device_ptr, size_t bytes);
/* Level hand-off without a copy engine: copies `words` u32 of device
 * memory into mapped host memory, clears the first `clear_words` of the
 * source, then writes `seq` to the mapped flag (system-scope release), so a
 * host spinning on the flag reads complete data.  Replaces the per-level
 * counter copy of the reference's iter_init (src/app.cuh:188-207). */
int scaph_publish_words(uint32_t* src, uint32_t words, uint32_t clear_words, uint32_t* dst_mapped,
                        uint32_t* flag_mapped, uint32_t seq, void* stream);

/* ------------------------------------------------------------- BFS state */
/* Counters written by the expand kernels, one record per level. */
typedef struct scaph_level_counters {
  unsigned long long n_next;   /* vertices first reached in this level */
  unsigned long long m_next;   /* sum of their out-degrees */
  unsigned long long q_len;    /* entries appended to the next queue */
  unsigned long long scanned;  /* edges examined (push: all; pull: until hit) */
  unsigned long long examined; /* pull: unvisited vertices whose in-list was read */
  unsigned long long reserved;
} scaph_level_counters;

/* Level init (app/bfs.cu:20-37): level[v] = SCAPH_UNREACHED for v < n,
 * visited = static_visited (bitmap, may be NULL for all-zero) then the
 * source bit set, level[src] = 0, queue[0] = src.  Bitmaps hold
 * ceil(n/64)*2 u32 words.  counters (one record) is zeroed. */
int scaph_bfs_init(uint32_t* level, uint32_t* visited, const uint32_t* static_visited,
                   uint32_t n, uint32_t src, uint32_t* queue,
                   scaph_level_counters* counters, void* stream);

/* Bitmap of vertices with in-degree 0 (they are never reached except as
 * the source, so the pull expand skips them).  in_degree: device u32[n].
 * Also sets the padding bits past n.  No reference counterpart. */
int scaph_bitmap_from_zero_degree(uint32_t* bitmap, const uint32_t* in_degree, uint32_t n,
                                  void* stream);

/* Workspace bytes the push expand needs for a frontier of up to
 * max_entries vertices (or a page of max_entries Node records) whose
 * out-edges total at most max_edges. */
size_t scaph_push_workspace_bytes(uint32_t max_entries, unsigned long long max_edges);

/* Launch shape of the push expand (NULL or 0 fields = built-in choice). */
typedef struct scaph_push_tuning {
  uint32_t grid_blocks; /* cap on expand blocks (one tile per block per pass: 2048 edges, 1024 on
                           XCD-sliced levels) */
  uint32_t edge_rows;   /* 1: a wave reads 256 consecutive edges per step (coalesced rows) */
  uint32_t small_edges; /* frontiers with at most this many out-edges run in one launch
                           (0: 8192; 1: effectively never) */
  uint32_t scan_single; /* at most this many prep blocks are scanned by one 1024-thread block;
                           more take three coalesced launches (0: 1024) */
  uint32_t xcd_slices;  /* levels with at least 1024 x this many out-edges expand XCD-sliced:
                           tiles inside one out-list go to the XCD owning their targets'
                           eighth of the vertex ids (0: off) */
  uint32_t coop_edges;  /* frontiers of at most 256 vertices with at most this many out-edges
                           (and more than small_edges, or more than 64 per vertex) run in one
                           launch with their edges spread over the grid (0: 2^20; 1: never) */
  uint32_t reserved[2];
} scaph_push_tuning;
/* Top-down (push) expand of one level from a frontier queue — the resident
 * counterpart of gpu_push_csr (src/gpu_engine.cuh:41-169) with the BFS
 * relax of app/bfs_kernel.cuh:6-16 turned into a level-synchronous claim:
 * a vertex v reached by an edge for the first time gets
 * level[v] = cur_level + 1, its visited bit, and a slot in q_out.
 * n_in is the frontier length, m_in the sum of its out-degrees (known from
 * the previous level's counters; used only to size the launch).
 * degree: device out-degree u32[n] (src/app.cuh:302, d_degree). */
int scaph_bfs_push(const scaph_graph_desc* out_graph, const uint32_t* degree,
                   uint32_t* level, uint32_t* visited, const uint32_t* q_in, uint32_t n_in,
                   unsigned long long m_in, uint32_t* q_out, scaph_level_counters* counters,
                   uint32_t cur_level, void* workspace, size_t workspace_bytes,
                   const scaph_push_tuning* tune, void* stream);

/* scaph_bfs_push without a workspace argument, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * workspace is a fixed device array of scaph_bfs_push_noalloc_max_bytes()
 * (64 MiB) that each device the library runs on holds until the process
 * ends; the call itself allocates and frees nothing.  A frontier whose
 * workspace (scaph_push_workspace_bytes(n_in, m_in)) exceeds it is refused
 * with SCAPH_STATUS_SCRATCH_TOO_SMALL before anything is launched.  Every
 * call on one device shares that array: calls that overlap on one device
 * (other streams or host threads) race and must be serialised by the
 * caller.  Outputs and semantics are those of scaph_bfs_push. */
size_t scaph_bfs_push_noalloc_max_bytes(void);
int scaph_bfs_push_noalloc(const scaph_graph_desc* out_graph, const uint32_t* degree,
                           uint32_t* level, uint32_t* visited, const uint32_t* q_in, uint32_t n_in,
                           unsigned long long m_in, uint32_t* q_out,
                           scaph_level_counters* counters, uint32_t cur_level,
                           const scaph_push_tuning* tune, void* stream);

/* Bottom-up (pull) expand of one level: every unvisited vertex scans its
 * in-neighbours (transposed paged CSR) for one in the frontier bitmap.
 * In-list entries with the top bit set are hub indices into hub_list
 * (num_hubs <= 262144; see scaph_csr_from_keys): their frontier bits are
 * gathered once per level and probed from LDS (num_hubs / 8 bytes per block,
 * so the hub count sets how many pull blocks stay resident on a CU).  Writes level/visited for
 * the vertices found and the full next-frontier bitmap (all words, so it
 * needs no clearing).  In-lists still unmatched after 512 entries are
 * finished by a second, grid-wide pass.  max_edges bounds the graph's
 * edges (workspace: scaph_pull_workspace_bytes).  heads (optional,
 * scaph_pull_heads) holds per vertex its first in-list entries and its
 * in-degree (scaph_list_heads_from_csr / scaph_list_heads_live): vertices
 * found there skip their page.  No reference counterpart (the reference
 * engine is push-only). */
size_t scaph_pull_workspace_bytes(unsigned long long max_edges);
/* The pull expand's head records (optional): per vertex its first in-list
 * entries, its in-degree and (head_outdeg) its out-degree, 16 bytes.  With
 * no_in_edges and live_base set, records exist only for vertices with
 * in-edges, in id order (vertex v's record is number live_base[v / 1024]
 * plus the vertices with in-edges in [1024 * (v / 1024), v)): vertices
 * without in-edges are never examined (they start visited), and leaving
 * them out halves the record bytes a pull level reads on R-MAT graphs.
 * Otherwise record v is at index v. */
typedef struct scaph_pull_heads {
  const uint32_t* head;         /* u32[4 * records] */
  const uint32_t* no_in_edges;  /* bitmap of the vertices without in-edges (padding bits set) */
  const uint32_t* live_base;    /* per 1024-id window: vertices with in-edges before it */
} scaph_pull_heads;
/* Launch shape of the pull expand (NULL or 0 fields = built-in choice). */
typedef struct scaph_pull_tuning {
  uint32_t grid_blocks; /* cap on blocks (each owns block_waves x 1024-vertex windows per pass) */
  uint32_t head_outdeg; /* in_head records are {e0, e1, in-degree, out-degree} */
  uint32_t no_prefetch; /* do not request the next batch's head records early */
  uint32_t head_indexed; /* heads from scaph_list_heads_indexed (implies head_outdeg) */
  uint32_t block_waves; /* waves per pull block: 4 or 8 (0: 8); the grid is what stays resident */
  uint32_t waves_per_simd; /* 8: registers capped so 8 waves share a SIMD (0: compiler's choice) */
  uint32_t two_phase;   /* with head records: a window's vertices probed by their heads first,
                           the undecided ones by their lists after (batches of like chains);
                           2: the head probes pipelined, a batch's words requested while the
                           batch before it is answered (0, 1 or 2) */
  uint32_t summary_shift; /* 0, or 5..31: scaph_bfs_pull first builds a frontier summary (one
                             bit per 2^shift vertices, set when any of them is in front_in;
                             ceil(n / 2^shift) <= 2^22) and tests it before each non-hub
                             in-neighbour's frontier word.  Pays when the frontier is sparse:
                             most probes then end at the summary.  scaph_bfs_pull_sparse
                             ignores it. */
} scaph_pull_tuning;
int scaph_bfs_pull(const scaph_graph_desc* in_graph, const scaph_pull_heads* heads,
                   const uint32_t* hub_list, uint32_t num_hubs,
                   const uint32_t* degree, uint32_t* level, uint32_t* visited,
                   const uint32_t* front_in, uint32_t* front_out, scaph_level_counters* counters,
                   uint32_t cur_level, unsigned long long max_edges, void* workspace,
                   size_t workspace_bytes, const scaph_pull_tuning* tune, void* stream);
/* scaph_bfs_pull for a level with few unvisited vertices left: each wave
 * gathers the unvisited vertices of its 1024-vertex windows into LDS and
 * probes them 64 at a time, instead of probing window by window with one
 * or two lanes busy.  Same arguments, results, counters and workspace as
 * scaph_bfs_pull; faster when unvisited vertices are a small share of the
 * graph, slower otherwise.  No reference counterpart. */
int scaph_bfs_pull_sparse(const scaph_graph_desc* in_graph, const scaph_pull_heads* heads,
                          const uint32_t* hub_list, uint32_t num_hubs, const uint32_t* degree,
                          uint32_t* level, uint32_t* visited, const uint32_t* front_in,
                          uint32_t* front_out, scaph_level_counters* counters, uint32_t cur_level,
                          unsigned long long max_edges, void* workspace, size_t workspace_bytes,
                          const scaph_pull_tuning* tune, void* stream);

/* Frontier format conversions (no reference counterpart). */
int scaph_frontier_queue_to_bitmap(const uint32_t* queue, uint32_t n_queue, uint32_t* bitmap,
                                   uint32_t n, void* stream);
/* The frontier a push level produced, as a bitmap: visited & ~snapshot,
 * where snapshot is the visited bitmap copied before the push (16-byte
 * aligned buffers).  Costs three bitmap passes instead of one atomic per
 * frontier vertex. */
int scaph_frontier_from_visited(const uint32_t* visited, const uint32_t* snapshot,
                                uint32_t* bitmap, uint32_t n, void* stream);
/* bitmap (16-byte aligned) -> the list of its set bits below n, appended
 * at counters->q_len (which it advances); in vertex order within each run
 * of 262 144 vertices. */
int scaph_frontier_bitmap_to_queue(const uint32_t* bitmap, uint32_t n, uint32_t* queue,
                                   scaph_level_counters* counters, void* stream);

/* Per-page active counts — get_page_info + get_total_info
 * (src/gpu_engine.cuh:397-456): for each page p,
 *   nodes[p] = #active v in [left,right],
 *   datas[p] = sum over active v of (deg(v) + 3)   (2*deg+3 if weighted),
 * totals[0] = sum nodes, totals[1] = sum datas.  "Active" is a bit in the
 * frontier bitmap (the reference's bool stat[] array). pages: device. */
int scaph_page_counts(const scaph_page* pages, uint32_t num_pages, const uint32_t* front,
                      const uint32_t* degree, int weighted, uint32_t* nodes, uint32_t* datas,
                      unsigned long long* totals, void* stream);

/* Per-page count of vertices still reachable by a later level: out[p] =
 * #v in [left, right] set in live (NULL = every vertex) and clear in the
 * visited bitmap.  Feeds the page-activity prediction of the paged
 * scheduler (a page whose count is 0 is never active again).  No reference
 * counterpart. */
int scaph_page_unvisited(const scaph_page* pages, uint32_t num_pages, const uint32_t* visited,
                         const uint32_t* live, uint32_t* out, void* stream);

/* What scaph_page_counts already says about one page push (optional).
 * first_vertex: Node i of a full page is vertex first_vertex + i (the
 * page's left); SCAPH_PAGE_COMPACTED for a compacted buffer, whose Nodes
 * are all active.  active_edges: datas[p] - 3 * nodes[p] for the same
 * frontier, i.e. the exact out-edge count of the active Nodes — it sizes
 * the launch, and an under-count truncates the level. */
#define SCAPH_PAGE_COMPACTED 0xFFFFFFFFu
typedef struct scaph_page_hint {
  uint32_t first_vertex;
  uint32_t active_nodes;
  unsigned long long active_edges;
} scaph_page_hint;

/* Push expand of ONE page buffer (full page, or a compacted cold page as
 * written by compress_stat, src/compress.h:18-40) — gpu_push_csr's unit of
 * work (src/app.cuh:438).  A Node is active when the frontier bitmap has
 * its vtx bit.  Newly reached vertices get level/visited and a bit in
 * front_out (atomic OR; front_out must be cleared before the level).
 * edgenum bounds the page's edges (sizes the launch), n = vertex count.
 * With a hint, a full page is read through its frontier bits (only active
 * Nodes are loaded), the launch is sized by active_edges, and a page with
 * at most tune->small_edges (default 8192) active out-edges runs as one
 * launch.  tune: NULL for the built-in launch shape.
 * Counters accumulate across calls: clear them at the start of a level. */
int scaph_bfs_push_page(const uint32_t* page_buf, uint32_t nodenum, unsigned long long edgenum,
                        uint32_t n, const uint32_t* degree, uint32_t* level, uint32_t* visited,
                        const uint32_t* front_in, uint32_t* front_out,
                        scaph_level_counters* counters, uint32_t cur_level, void* workspace,
                        size_t workspace_bytes, const scaph_page_hint* hint,
                        const scaph_push_tuning* tune, void* stream);

/* Push expand of a GROUP of resident full pages in one launch sequence —
 * several of gpu_push_batch's per-page gpu_push_csr launches
 * (src/app.cuh:414-459) fused.  runs (device memory) lists the pages in
 * any order: run j is the buffer at runs[j].buf (Nodes, then edges), Node
 * i is vertex first_vertex + i, and seg0 = the sum of
 * ceil(nodenum / SCAPH_RUN_SEGMENT) over the runs before j; num_segs is that
 * sum over all runs.  Each segment is read through its frontier bits, so
 * only active Nodes are loaded.  active_nodes and active_edges are the
 * runs' nodes[p] and datas[p] - 3 * nodes[p] sums of scaph_page_counts for
 * the same frontier: they size the launch and the workspace; an
 * under-count yields a truncated or empty (wrong) level, never an access
 * outside the workspace.  Groups with at most tune->small_edges (default
 * 8192) active out-edges run as one launch.  Output as scaph_bfs_push_page.
 * Workspace: scaph_push_runs_workspace_bytes(num_segs, active_nodes,
 * active_edges). */
#define SCAPH_RUN_SEGMENT 8192u
typedef struct scaph_page_run {
  unsigned long long buf;
  uint32_t first_vertex;
  uint32_t nodenum;
  uint32_t seg0;
  uint32_t reserved;
} scaph_page_run;
size_t scaph_push_runs_workspace_bytes(uint32_t num_segs, unsigned long long active_nodes,
                                       unsigned long long active_edges);
int scaph_bfs_push_runs(const scaph_page_run* runs, uint32_t num_runs, uint32_t num_segs,
                        unsigned long long active_nodes, unsigned long long active_edges,
                        uint32_t n, const uint32_t* degree, uint32_t* level, uint32_t* visited,
                        const uint32_t* front_in, uint32_t* front_out,
                        scaph_level_counters* counters, uint32_t cur_level, void* workspace,
                        size_t workspace_bytes, const scaph_push_tuning* tune, void* stream);

/* ----------------------------------------------------- other App plugins */
/* Label-correcting push rounds for the reference's WCC and SSSP apps on the
 * same paged graph (app/wcc.cu:14-24 with app/wcc_kernel.cuh: value[v] = v,
 * every vertex active, dst = min(dst, src); app/sssp.cu:13-31 with
 * app/sssp_kernel.cuh: value = 0x7fffffff except value[src] = 0,
 * dst = min(dst, src + w)).  The fixpoint is the minimum over all paths,
 * independent of the update order.  Sums are u32 as in the reference's
 * value_t: shortest distances must stay below 0x7FFFFFFF (MAXPATH, which
 * marks unreached vertices) and value + weight below 2^32. */
#define SCAPH_RELAX_WCC 0
#define SCAPH_RELAX_SSSP 1
#define SCAPH_RELAX_INF 0x7FFFFFFFu /* unreached distance (MAXPATH, app/sssp.cu:11) */
/* value[v] = v (WCC), or SCAPH_RELAX_INF with value[src] = 0 (SSSP). */
int scaph_relax_init(uint32_t* value, uint32_t n, int mode, uint32_t src, void* stream);
/* Workspace for the chunks of long out-lists of one round (max_edges =
 * the graph's edges). */
size_t scaph_relax_workspace_bytes(unsigned long long max_edges);
/* One round over the active vertices queue[0..n_queue): every out-edge
 * (u, v) offers value[u] (+ its weight when `weights` is given) to v; a value
 * lowered marks v in front_out (atomic OR; front_out must be clear) and
 * counts in counters->n_next (newly marked vertices) and ->scanned.
 * weights (optional): a device array laid out like the page pool whose
 * base is pool_base — the weight of the out-edge at address a is at
 * weights + (a - pool_base).  counters->reserved becomes 1 if the chunk
 * workspace overflowed (the round is then incomplete). */
int scaph_relax_round(const scaph_graph_desc* graph, const uint32_t* queue, uint32_t n_queue,
                      const uint32_t* weights, const uint32_t* pool_base, uint32_t* value,
                      uint32_t* front_out, scaph_level_counters* counters, void* workspace,
                      size_t workspace_bytes, void* stream);
/* Near/far split of a round's active list (the bucketed rounds of SSSP):
 * vertices of queue[0..n_queue) whose value is below `bound` are appended
 * to near_queue (this round's list); the others are set in front_out (they
 * stay active) and their least value is kept.  split (3 u64, device):
 * [0] near count, [1] far count, [2] least far value (~0 if none).  The
 * values reached are the same for any bound: only the number of times a
 * vertex is relaxed changes (label-correcting rounds, app/sssp.cu). */
int scaph_relax_split(const uint32_t* queue, uint32_t n_queue, const uint32_t* value,
                      uint32_t bound, uint32_t n, uint32_t* near_queue, uint32_t* front_out,
                      unsigned long long* split, void* stream);
/* The same round over one page buffer in device memory (paged mode; the
 * reference's per-page gpu_push_csr launch, src/gpu_engine.cuh:41-169):
 * the active Nodes of [first_vertex, first_vertex + nodenum) are those set
 * in front_in; with first_vertex == SCAPH_PAGE_COMPACTED the buffer holds
 * only active Nodes (the compacted layout), each naming its vertex.  weights/pool_base as above, with pool_base the base of the
 * page slot pool holding buf.  Workspace from scaph_relax_workspace_bytes
 * (page edges). */
int scaph_relax_page(const uint32_t* buf, uint32_t nodenum, uint32_t first_vertex,
                     const uint32_t* front_in, uint32_t n, const uint32_t* weights,
                     const uint32_t* pool_base, uint32_t* value, uint32_t* front_out,
                     scaph_level_counters* counters, void* workspace, size_t workspace_bytes,
                     void* stream);

/* Level-array check kernel: counts edges (u,v) violating
 * level[v] <= level[u] + 1 and reached vertices without a parent at
 * level-1 (BFS invariants; in_graph entries decoded through hub_list).
 * out[0] = violations of the edge rule,
 * out[1] = reached non-source vertices lacking a parent. */
int scaph_bfs_validate(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                       const uint32_t* hub_list, const uint32_t* level, uint32_t src,
                       unsigned long long* out, void* stream);

/* ------------------------------------------------------------ generators */
/* R-MAT edge generator (no reference counterpart; the reference ships no
 * generator).  Edge e draws `scale` quadrant choices from SplitMix64 words
 * keyed by seed; quadrant thresholds are u32 (thr_a, thr_ab, thr_abc).
 * Vertex IDs are permuted by the keyed bijection scaph_perm_key.
 * Writes packed keys (u << 32 | v) for e in [e_begin, e_end); self-loops
 * become the key (n-1, n-1) so they collapse to one key after dedup. */
int scaph_gen_rmat_keys(uint64_t* keys, uint64_t e_begin, uint64_t e_end, uint32_t scale,
                        uint32_t thr_a, uint32_t thr_ab, uint32_t thr_abc, uint64_t seed,
                        void* stream);

/* Erdos-Renyi G(n, m) edge keys for e in [e_begin, e_end): endpoints are
 * the high halves of SplitMix64 words 2e and 2e+1 times n.  No reference
 * counterpart. */
int scaph_gen_er_keys(uint64_t* keys, uint64_t e_begin, uint64_t e_end, uint32_t n,
                      uint64_t seed, void* stream);

/* head[4v..4v+3] = the first three entries of v's list (0xFFFFFFFF past
 * its end) and its length, from a flat device CSR; with out_degree (u32[n])
 * the record is {first two entries, length, out_degree[v]} instead
 * (scaph_pull_tuning.head_outdeg). */
int scaph_list_heads_from_csr(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                              const uint32_t* out_degree, uint32_t* head, void* stream);
/* base[k] = vertices with in-edges among ids < 1024 k, for every 1024-id
 * window k of the no_in_edges bitmap (scaph_bitmap_from_zero_degree);
 * allocates its own scan storage and returns after the stream drains. */
int scaph_live_window_base(const uint32_t* no_in_edges, uint32_t n, uint32_t* base,
                           void* stream);
/* The head records of scaph_list_heads_from_csr for the vertices with
 * in-edges only, in id order, numbered from base (scaph_live_window_base);
 * head must hold 4 u32 per such vertex. */
int scaph_list_heads_live(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                          const uint32_t* out_degree, const uint32_t* no_in_edges,
                          const uint32_t* base, uint32_t* head, void* stream);
/* Indexed head records {first two entries, (offset << 8) | min(length, 255),
 * out_degree[v]}: offset is v's Node idx, the word offset of its list in its
 * page's edge array (pages: the page table of the same CSR, page_cuts; every
 * page's edges must number below 2^24).  A pull that finds no parent in the
 * first two entries then reads the list without the Node record; a length
 * field of 255 means 255 or more (the Node is read).  With no_in_edges and
 * base (scaph_live_window_base) only vertices with in-edges get records, as
 * in scaph_list_heads_live; with both NULL record v is at index v. */
int scaph_list_heads_indexed(const uint64_t* offsets, const uint32_t* ngh, uint32_t n,
                             const uint32_t* out_degree, const scaph_page* pages,
                             uint32_t num_pages, const uint32_t* no_in_edges, const uint32_t* base,
                             uint32_t* head, void* stream);
/* offsets[0..n] = exclusive prefix of degree[0..n) (offsets[n] = total);
 * allocates its own scan storage and returns after the stream drains. */
int scaph_degrees_to_offsets(const uint32_t* degree, uint64_t* offsets, uint32_t n, void* stream);
/* Power-law digraph edges [e_begin, e_end) as keys (u << 32 | v): the
 * source is the vertex whose range [offsets[u], offsets[u+1]) holds e (the
 * out-degrees come from the caller), the destination floor(w * n / 2^64)
 * for SplitMix64 word e of (seed ^ "zipf-dst").  No reference counterpart. */
int scaph_gen_zipf_keys(uint64_t* keys, const uint64_t* offsets, uint32_t n, uint64_t e_begin,
                        uint64_t e_end, uint64_t seed, void* stream);
/* Sorts keys, removes duplicates and the self-loop key, and writes a flat
 * CSR: offsets u64[n+1], neighbours u32[*num_edges], out-degree u32[n].
 * transpose != 0 builds in-edges (the key halves swapped); then, when
 * hub_idx (device u32[n], 0xFFFFFFFF = not a hub) is given, each in-list
 * holds its hub sources first, as entries 0x80000000 | hub index, then the
 * other sources by id (the layout scaph_bfs_pull reads).  keys is
 * clobbered; keys_tmp must hold num_keys u64; ngh must hold num_keys u32.
 * Allocates and frees its own sort scratch. */
int scaph_csr_from_keys(uint64_t* keys, uint64_t* keys_tmp, uint64_t num_keys, uint32_t n,
                        int transpose, const uint32_t* hub_idx, uint64_t* offsets, uint32_t* ngh,
                        uint32_t* degree, uint64_t* num_edges, void* stream);

/* Writes the pages of a flat device CSR into page buffers, in the byte
 * layout of tools/bin2pagecsr.cpp:221-242 (Node records then neighbours).
 * pages/page_ptr: device arrays (page cuts computed by the host). Words of
 * each buffer past the used part are NOT touched. */
int scaph_pages_from_csr(const uint64_t* offsets, const uint32_t* ngh, const scaph_page* pages,
                         const uint64_t* page_ptr, uint32_t num_pages, uint32_t n, void* stream);

/* ------------------------------------------------------------------- MLP */
/* 16 -> 64 -> 64 -> 2 ReLU MLP of the learned direction/prefetch schedule
 * (no reference counterpart).  Weights are bf16 (u16 bit patterns),
 * accumulation is fp32.  Parameter layout:
 *   W1[64][16], b1[64], W2[64][64], b2[64], W3[2][64], b3[2].
 * x: batch x 16 features; y: batch x 2 outputs (0: pull-is-faster logit,
 * 1: next-level page-activity threshold); act (optional, batch x 128) keeps
 * the hidden activations for backward. */
#define SCAPH_MLP_IN 16
#define SCAPH_MLP_H 64
#define SCAPH_MLP_OUT 2
#define SCAPH_MLP_NPARAM (64 * 16 + 64 + 64 * 64 + 64 + 2 * 64 + 2)
int scaph_mlp_forward(const uint16_t* params_bf16, const float* x, uint32_t batch, float* y,
                      float* act, void* stream);
/* Batch-mean loss softplus(y0) - t0*y0 + w1*(y1 - t1)^2 with labels
 * t = label[2*s..]; grad (fp32, SCAPH_MLP_NPARAM) is overwritten with its
 * gradient and loss[0] with the loss. */
int scaph_mlp_backward(const uint16_t* params_bf16, const float* x, const float* act,
                       const float* y, const float* label, uint32_t batch, float w1, float* grad,
                       float* loss, void* stream);
/* momentum = mu*momentum + grad; master -= lr*momentum;
 * params_bf16 = round-to-nearest-even(master). */
int scaph_mlp_sgd(float* master, float* momentum, const float* grad, uint32_t nparam, float lr,
                  float mu, uint16_t* params_bf16, void* stream);
/* Per-page active counts of a frontier into page_cnt[num_pages]: from a
 * bitmap, or (bitmap NULL) from a vertex list whose length is read on the
 * device from *queue_len (e.g. the level counters' q_len). */
int scaph_frontier_page_counts(const uint32_t* bitmap, const uint32_t* queue,
                               const unsigned long long* queue_len, const scaph_page* pages,
                               uint32_t num_pages, uint32_t n, uint32_t* page_cnt, void* stream);
/* One decision on the device: the level's counters (next frontier size and
 * edges), the unexplored edges before the level (m_u; the kernel subtracts
 * the counters' m_next) and the per-page counts become the 16 features
 * (normalised by norm = mean[16], scale[16]); out (32 floats) receives
 * [0] the pull logit, [1] the threshold, [2] the number of active pages,
 * [3] padding, [4..15] the 12-bin histogram, [16..31] the raw features. */
int scaph_mlp_decide(const uint16_t* params_bf16, const float* norm, const uint32_t* page_cnt,
                     const scaph_page* pages, uint32_t num_pages,
                     const scaph_level_counters* counters, unsigned long long m_u,
                     uint32_t level, float* out, void* stream);

/* The decision of a model whose W1 columns for the 12 histogram features
 * are zero (it reads n_f, m_f, m_u and the level only), fused with
 * scaph_publish_words: one launch per level computes the features from
 * counters (= the counters words) and m_u, runs the forward pass, and
 * publishes `words` words with the 32-float record of scaph_mlp_decide at
 * word decide_word (histogram and page count zero), clearing the first
 * clear_words.  Requires decide_word >= the counters' size in words,
 * decide_word + 32 <= words and clear_words <= decide_word. */
int scaph_mlp_decide_publish(const uint16_t* params_bf16, const float* norm, uint32_t* counters,
                             uint32_t decide_word, unsigned long long m_u, uint32_t level,
                             uint32_t words, uint32_t clear_words, uint32_t* dst_mapped,
                             uint32_t* flag_mapped, uint32_t seq, void* stream);

/* scaph_frontier_page_counts and scaph_mlp_decide as the engine uses them
 * once per level (two launches, no host work between): the frontier is a
 * bitmap, or (bitmap NULL) a list whose length is counters->q_len.
 * page_cnt (num_pages u32) is workspace that must be zero before the first
 * call; the decision kernel leaves it zero again.  Calls sharing that
 * workspace must not overlap.  out as for scaph_mlp_decide. */
int scaph_mlp_decide_frontier(const uint16_t* params_bf16, const float* norm,
                              const uint32_t* bitmap, const uint32_t* queue,
                              const scaph_page* pages, uint32_t num_pages, uint32_t n,
                              const scaph_level_counters* counters, unsigned long long m_u,
                              uint32_t level, uint32_t* page_cnt, float* out, void* stream);

/* ------------------------------------------------- multi-source bitset BFS */
/* Up to 256 BFS queries over one graph at once (no reference counterpart):
 * each vertex carries 4 x 64-bit words, one bit per query, for the vertices
 * it has been reached from ("seen") and, per level, for the queries whose
 * frontier it is in at that level.  The per-level frontier words are kept
 * (front[L]), so the level of vertex v in query q is the L whose front[L][v]
 * holds bit q; the decode entry points turn that into uint32 levels.
 * Every kernel of a level reads the level's counters on the device to pick
 * push or pull and to stop once the frontier is empty, so a sequence of
 * levels can be captured into a hipGraph and replayed without host work. */
#define SCAPH_MS_WORDS 4
#define SCAPH_MS_MAX_QUERIES 256
typedef struct scaph_ms_counters {
  unsigned long long n_active;  /* vertices in the level's frontier (any query) */
  unsigned long long bits;      /* (vertex, query) pairs first reached at the level */
  unsigned long long m_active;  /* out-edges of the frontier's vertices */
  unsigned long long q_len;     /* push: frontier list length */
  unsigned long long heavy_len; /* push: long out-lists cut into chunks */
  unsigned long long overflow;  /* non-zero: a workspace bound was exceeded */
  unsigned long long pad[2];
} scaph_ms_counters; /* 64 bytes */
typedef struct scaph_ms_params {
  uint32_t num_queries;       /* 1..256 */
  uint32_t push_div;          /* push while n_active * push_div < n ... */
  uint32_t push_mdiv;         /* ... and m_active * push_mdiv < m */
  uint32_t pull_block;        /* threads per pull block: 256, 512 or 1024 */
  uint32_t pull_grid;         /* cap on pull blocks (0: one per block of vertices, at most 16384) */
  uint32_t num_hubs;          /* entries of hub_list (<= SCAPH_MS_MAX_HUBS; 0: none) */
  unsigned long long m;       /* edges of the graph */
} scaph_ms_params;
#define SCAPH_MS_MAX_HUBS 262144u
typedef struct scaph_ms_state {
  uint64_t* seen;             /* n x 4 words */
  uint64_t* front;            /* max_levels x n x 4 words */
  uint64_t* act;              /* max_levels x act_words: frontier membership bits */
  scaph_ms_counters* ctr;     /* max_levels + 1 records */
  uint32_t* queue;            /* n entries */
  uint64_t* heavy;            /* heavy_cap chunk items */
  uint64_t* hub_front;        /* SCAPH_MS_MAX_HUBS x 4 words: the hubs' frontier words of the
                                 level being pulled (hub i = hub_list[i]; set where hub_bits is) */
  uint32_t* hub_bits;         /* SCAPH_MS_MAX_HUBS bits: hub i in that level's frontier */
  uint32_t n, max_levels;
  unsigned long long act_words; /* ceil(n / 64) */
  unsigned long long heavy_cap;
} scaph_ms_state;
/* Device bytes of the state for n vertices, m edges and max_levels levels
 * (the caller carves the buffers; see scaph_ms_state_carve). */
size_t scaph_ms_state_bytes(uint32_t n, unsigned long long m, uint32_t max_levels);
/* Fill *st's pointers from one allocation of scaph_ms_state_bytes bytes. */
int scaph_ms_state_carve(void* base, uint32_t n, unsigned long long m, uint32_t max_levels,
                         scaph_ms_state* st);
/* Level 0: clears the state and sets query q's bit at sources[q]. */
int scaph_ms_init(const scaph_ms_state* st, const uint32_t* sources, uint32_t count,
                  const uint32_t* degree, void* stream);
/* One level (level -> level+1): push (out_graph) or pull (in_graph with
 * hub-tagged in-lists, as scaph_bfs_pull) chosen on the device.  A pull
 * level first gathers the frontier words of the prm->num_hubs hubs into
 * st->hub_front (their membership bits into st->hub_bits, staged in LDS),
 * so a hub-tagged in-edge is one LDS bit test and one load from a 32-byte
 * x num_hubs array instead of three dependent loads. */
int scaph_ms_level(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                   const uint32_t* hub_list, const uint32_t* degree, const scaph_ms_state* st,
                   uint32_t level, const scaph_ms_params* prm, void* stream);
/* Pull blocks of pull_block threads that stay resident on one CU for an
 * in-graph of in_pages pages and num_hubs hubs (a pull_grid of this times
 * the CU count keeps every launched block resident). */
int scaph_ms_pull_blocks_per_cu(uint32_t pull_block, uint32_t in_pages, uint32_t num_hubs,
                                uint32_t* out);
/* Every level from first_level on, in one cooperative launch of resident
 * 256-thread blocks (the same phases as scaph_ms_level, separated by grid
 * barriers; prm->pull_block and pull_grid are not used).  Stops when a
 * level's frontier is empty or at level max_levels - 1, whose counters then
 * show a frontier still active.  No reference counterpart. */
int scaph_ms_run_persistent(const scaph_graph_desc* out_graph, const scaph_graph_desc* in_graph,
                            const uint32_t* hub_list, const uint32_t* degree,
                            const scaph_ms_state* st, uint32_t first_level,
                            const scaph_ms_params* prm, void* stream);
/* uint32 levels of one query (SCAPH_UNREACHED where not reached) from the
 * first `depth` levels. */
int scaph_ms_levels_u32(const scaph_ms_state* st, uint32_t depth, uint32_t query, uint32_t* out,
                        void* stream);
/* Levels of all queries as bytes, out[v * 256 + q] (0xFF = not reached);
 * needs depth < 255. */
int scaph_ms_levels_u8(const scaph_ms_state* st, uint32_t depth, uint8_t* out, void* stream);

/* ------------------------------------------------ batch of small graphs */
/* Many independent small graphs (n <= 65536 each) stored as one disjoint-
 * union CSR and traversed by one persistent launch, one graph per
 * workgroup at a time (no reference counterpart). */
typedef struct scaph_batch_entry {
  unsigned long long vbase;   /* first global vertex id of the graph */
  unsigned long long ebase;   /* first raw edge index (generator stream) */
  unsigned long long seed;    /* the graph's R-MAT seed */
  uint32_t n;                 /* vertices (<= 65536) */
  uint32_t scale;             /* n = 2^scale for generated graphs */
  uint32_t src;               /* BFS source (local id) */
  uint32_t edge_factor;
} scaph_batch_entry;
/* Device workspace of scaph_batch_gen_rmat_keys for `count` graphs. */
size_t scaph_batch_perm_ws_bytes(uint32_t count);
/* Raw R-MAT edge keys (global u << 32 | global v) of every graph of the
 * batch, graph g's edges at [ebase_g, ebase_g + edge_factor_g * 2^scale_g):
 * each graph is the same edge stream as scaph_gen_rmat_keys with its seed.
 * Build the union CSR from them with scaph_csr_from_keys. */
int scaph_batch_gen_rmat_keys(uint64_t* keys, const scaph_batch_entry* graphs, uint32_t count,
                              unsigned long long total_edges, uint32_t thr_a, uint32_t thr_ab,
                              uint32_t thr_abc, void* perm_ws, void* stream);
/* BFS of every graph from its source: levels (global vertex order) get the
 * uint32 levels (SCAPH_UNREACHED where not reached).  order lists graph
 * indices in the order workgroups take them; `work` is one u32 of scratch,
 * qbuf grid * qcap u32 (qcap >= the largest n).  block: 256, 512 or 1024
 * threads; grid: persistent workgroups.  Capturable (no memset). */
int scaph_batch_bfs(const uint64_t* offsets, const uint32_t* ngh, const scaph_batch_entry* graphs,
                    const uint32_t* order, uint32_t count, uint32_t* levels, uint32_t* work,
                    uint32_t* qbuf, uint32_t qcap, uint32_t block, uint32_t grid, void* stream);

/* ------------------------------------------------------------ hipGraph */
/* Capture everything launched on `stream` between begin and end into an
 * executable graph (opaque handle); launch replays it on `stream`. */
int scaph_graph_capture_begin(void* stream);
int scaph_graph_capture_end(void* stream, void** exec);
int scaph_graph_launch(void* exec, void* stream);
int scaph_graph_destroy(void* exec);

#ifdef __cplusplus
}
#endif
#endif /* SCAPH_HIP_H */
