// bin2totem — exports a flat CSR graph (<base>_csr.{idx,ngh,deg[,wgh]}) as text
// for a multi-core CPU framework, in the format of tools/bin2totem.cpp, tools/bin2totemweight.cpp.
//   bin2totem <base>          unweighted
//   bin2totem <base> w        with the _csr.wgh weights (also when the binary is
//                          invoked as bin2totemweight, the reference's name)
// u32 and u64 row offsets are both read (the reference reads u32 only).
#include <cstdio>
#include <cstring>
#include <string>

#include "../host/graph_io.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::printf("usage: bin2totem <base> [w]\n");
    return 1;
  }
  const std::string self = argv[0];
  const bool weighted = (argc > 2 && argv[2][0] == 'w') ||
                        (self.size() >= 6 && self.compare(self.size() - 6, 6, "weight") == 0);
  scaph::HostCsr g;
  std::string err;
  if (!scaph::read_csr_files(argv[1], &g, &err) || !scaph::write_totem(argv[1], g, weighted, &err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  return 0;
}
