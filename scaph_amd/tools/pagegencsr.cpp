// pagegencsr — offline page generator with the command line and output of
// tools/bin2pagecsr.cpp:254-312:  pagegencsr <base> <pagesize MiB words> w|u [t]
// Reads <base>_csr.{ngh,deg,idx} (and _csr.wgh when weighted) and writes
// <base>page<N>, <base>page.info and <base>graph.info.  With a fourth
// argument `t` it also writes the in-edge graph for the pull step:
// <base>_in_csr.{idx,ngh,deg} and its pages <base>_inpage<N>, which the
// engine loads instead of transposing.
#include <cstdio>
#include <cstdlib>
#include <string>

#include "../host/graph_io.h"

int main(int argc, char** argv) {
  if (argc < 4) {
    std::printf("usage : ./pagegencsr filename pagesize(MB Vertex) w/u(weighted) [t]\n");
    return 0;
  }
  const std::string base = argv[1];
  const uint64_t P = (uint64_t)std::atol(argv[2]) * 1024 * 1024;
  const bool weighted = argv[3][0] == 'w';
  scaph::HostCsr g;
  std::string err;
  if (!scaph::read_csr_files(base, &g, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  if (weighted && g.wgh.size() != g.ngh.size()) {
    std::fprintf(stderr, "weighted pages need %s_csr.wgh\n", base.c_str());
    return 1;
  }
  std::vector<scaph_page> pages;
  if (!scaph::write_page_files(base, g, P, weighted, &pages, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  if (argc > 4 && argv[4][0] == 't') {
    std::vector<scaph_page> in_pages;
    if (!scaph::write_transposed_files(base, g, P, &in_pages, &err)) {
      std::fprintf(stderr, "%s\n", err.c_str());
      return 1;
    }
    std::printf("in-edge pages: %zu\n", in_pages.size());
  }
  for (const auto& p : pages)
    std::printf("left: %u right: %u\nedgenum: %u\nnodenum: %u\n\n", p.left, p.right, p.edgenum,
                p.nodenum);
  std::printf("node: %u\nedge: %llu\n", g.n, (unsigned long long)g.num_edges());
  return 0;
}
