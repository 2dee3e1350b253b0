// graphgen — writes a synthetic graph in the reference's flat CSR layout
// (<base>_csr.{idx,ngh,deg[,wgh]}) and, with -pagesize, its page files.
//   graphgen -gen rmat:scale=20,ef=16,seed=1[,wmax=255] -out /data/g [-idx64] [-wgh]
//            [-pagesize 10]
// With -wgh the weights file holds the spec's wmax= weights, or zeros.
#include <cstdio>
#include <string>
#include <vector>

#include "../host/flags.h"
#include "../host/gen.h"

int main(int argc, char** argv) {
  std::string spec = "rmat:scale=16,ef=16,seed=1", out;
  bool idx64 = true, wgh = false;
  int32_t pagesize = 0;
  scaph::FlagRegistry& r = scaph::FlagRegistry::global();
  r.add("gen", scaph::FlagType::kString, &spec, "generator spec");
  r.add("out", scaph::FlagType::kString, &out, "output base path");
  r.add("idx64", scaph::FlagType::kBool, &idx64, "u64 row offsets (else u32)");
  r.add("wgh", scaph::FlagType::kBool, &wgh, "also write _csr.wgh (wmax= weights, else zeros)");
  r.add("pagesize", scaph::FlagType::kInt32, &pagesize, "also write pages of this many Mi words");
  std::string err;
  if (!scaph::parse_command_line_flags(&argc, &argv, &err) || out.empty()) {
    std::fprintf(stderr, "%s\nusage: graphgen -gen <spec> -out <base>\n", err.c_str());
    return 1;
  }
  scaph::HostCsr g;
  if (!scaph::generate_from_spec(spec, &g, &err) || !scaph::write_csr_files(out, g, idx64, wgh, &err) ||
      (pagesize > 0 &&
       !scaph::write_page_files(out, g, (uint64_t)pagesize << 20, false, nullptr, &err))) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  std::printf("vertices %u edges %llu\n", g.n, (unsigned long long)g.num_edges());
  return 0;
}
