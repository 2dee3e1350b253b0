// graph_io.h — the reference's on-disk graph layout and the host-side CSR.
//
// Files for a base path <g> (src/graph.h:24-108, tools/bin2pagecsr.cpp:96-252):
//   <g>_csr.idx   row offsets, V+1 entries (u32, or u64 — detected by size)
//   <g>_csr.ngh   u32 neighbour IDs, E entries
//   <g>_csr.deg   u32 out-degree, V entries
//   <g>_csr.wgh   u32 weights, E entries (the page tool reads it even for
//                 unweighted graphs, tools/bin2pagecsr.cpp:281-287)
//   <g>graph.info {u32 pagesize, u32 nodenum, u64 edgenum} (16 bytes)
//   <g>page.info  Page{u32 left,right,nodenum,edgenum} per page
//   <g>page<N>    exactly pagesize u32 words: Node{vtx,idx,len}[nodenum]
//                 (idx relative to the page's first edge), the neighbours
//                 (or {ngr,wgh} pairs when weighted), then zero fill.
#ifndef SCAPH_HOST_GRAPH_IO_H
#define SCAPH_HOST_GRAPH_IO_H

#include <cstdint>
#include <string>
#include <vector>

#include "scaph_hip.h"

namespace scaph {

// src/page.h:31-37 with vertex_t = index_t = uint32.
struct Node {
  uint32_t vtx;
  uint32_t idx;
  uint32_t len;
};
static_assert(sizeof(Node) == 12, "Node record is 3 u32 words");

// src/page.h:14-20 as written by tools/bin2pagecsr.cpp:211-216.
struct GraphInfo {
  uint32_t pagesize;
  uint32_t nodenum;
  uint64_t edgenum;
};
static_assert(sizeof(GraphInfo) == 16, "graph.info is 16 bytes");
static_assert(sizeof(scaph_page) == 16, "Page record is 4 u32 words");

// A flat CSR held on the host.
struct HostCsr {
  uint32_t n = 0;
  std::vector<uint64_t> offsets;  // n+1
  std::vector<uint32_t> ngh;      // E
  std::vector<uint32_t> wgh;      // E or empty
  uint64_t num_edges() const { return ngh.size(); }
  uint32_t degree(uint32_t v) const { return (uint32_t)(offsets[v + 1] - offsets[v]); }
};

// Page cut (tools/bin2pagecsr.cpp:106-163, 206-208): starting at `left`, the
// largest right <= min(left + P, n - 1) with
//   (off[right+1] - off[left]) * (weighted ? 2 : 1) + 3 * (right - left + 1) < P.
// Returns false when even the single vertex `left` does not fit.
bool page_cut(const uint64_t* off, uint32_t n, uint32_t left, uint64_t pagesize_words,
              bool weighted, uint32_t* right);
bool page_cuts(const HostCsr& g, uint64_t pagesize_words, bool weighted,
               std::vector<scaph_page>* pages, std::string* err);
// Same over a raw offsets array.
bool page_cuts_raw(const uint64_t* off, uint32_t n, uint64_t pagesize_words, bool weighted,
                   std::vector<scaph_page>* pages, std::string* err);

// Words one page occupies before zero fill.
inline uint64_t page_used_words(const scaph_page& p, bool weighted) {
  return 3ull * p.nodenum + (weighted ? 2ull : 1ull) * p.edgenum;
}

// Fills buf (>= page_used_words words) with page p of g (Node records then
// neighbours), the content tools/bin2pagecsr.cpp:221-242 writes.
void fill_page(const HostCsr& g, const scaph_page& p, bool weighted, uint32_t* buf);

// Writers / readers of the files above.
bool write_csr_files(const std::string& base, const HostCsr& g, bool idx64, bool write_wgh,
                     std::string* err);
bool read_csr_files(const std::string& base, HostCsr* g, std::string* err);
// pagegencsr equivalent (tools/bin2pagecsr.cpp:197-252): writes page<N>,
// page.info and graph.info.  pagesize_words = MB * 2^20.
bool write_page_files(const std::string& base, const HostCsr& g, uint64_t pagesize_words,
                      bool weighted, std::vector<scaph_page>* pages_out, std::string* err);
bool read_page_info(const std::string& base, std::vector<scaph_page>* pages, std::string* err);
bool read_graph_info(const std::string& base, GraphInfo* info, std::string* err);
bool read_page_file(const std::string& base, int index, uint64_t pagesize_words, uint32_t* buf,
                    std::string* err);

// Text exporters for CPU baselines, in the formats of tools/bin2ligra.cpp,
// tools/bin2ligraweight.cpp, tools/bin2totem.cpp and
// tools/bin2totemweight.cpp:
//   <base>.ligra (or <base>.ligra.weight): "AdjacencyGraph" (or
//     "WeightedAdjacencyGraph"), n, m, then the n row offsets, the m
//     neighbours (and the m weights), one number per line;
//   <base>.totem (or <base>weight.totem): "#Nodes: n", "#Edges: m",
//     "#Undirected", then one "u\tv" (or "u\tv\tw") line per edge.
bool write_ligra(const std::string& base, const HostCsr& g, bool weighted, std::string* err);
bool write_totem(const std::string& base, const HostCsr& g, bool weighted, std::string* err);

// In-edge CSR (counting sort; neighbour lists come out sorted by source).
HostCsr transpose(const HostCsr& g);

// Order-free 64-bit digest of the edge multiset {(u, v)} of g, or of
// {(v, u)} when `transposed` (a sum of mixed per-edge words): equal for a
// graph and its transpose read as transposed, so a stored transpose can be
// checked against its graph in two sequential passes.
uint64_t edge_multiset_digest(const HostCsr& g, bool transposed);

// The transposed-page generator for the pull step: <base>_in_csr.{idx,ngh,
// deg} (u64 offsets) and <base>_in page files of the in-edge graph, cut with
// the same page size.
bool write_transposed_files(const std::string& base, const HostCsr& g, uint64_t pagesize_words,
                            std::vector<scaph_page>* pages_out, std::string* err);

// Hubs: the (at most) max_hubs vertices of largest out-degree (> 0), by
// degree descending then id.  Entry i of the result has hub index i.
std::vector<uint32_t> select_hubs(const std::vector<uint32_t>& out_degree, uint32_t max_hubs);
// Re-encodes an in-edge CSR for the pull expand: in each list the hub
// sources come first, by hub index, as 0x80000000 | index; the others
// follow by id.
void encode_hub_in_lists(HostCsr* in, const std::vector<uint32_t>& hubs);

}  // namespace scaph

#endif
