// sssp — command-line driver of the SSSP app (app/sssp.cu:63-79): loads
// -graphfile with its _csr.wgh (or -gen with wmax=), runs -rounds searches
// from -src and prints the cost; -check compares against Dijkstra on the
// CPU; -levels_out writes the distances.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "../host/app.h"

int main(int argc, char** argv) {
  using namespace scaph;
  try {
    sssp<uint32_t, uint32_t, uint32_t> mysssp(argc, argv);
    mysssp.load_graph();
    const int rounds = mysssp.cfg.rounds > 0 ? mysssp.cfg.rounds : 1;
    double total = 0;
    for (int i = 0; i < rounds; ++i) {
      mysssp.run();
      total += mysssp.last.ms / 1e3;
    }
    std::cout << "sssp cost " << total / rounds << " seconds " << mysssp.last.depth
              << " rounds, " << mysssp.last.m_reached << " edges relaxed" << std::endl;
    if (!mysssp.cfg.levels_out.empty()) {
      std::vector<uint32_t> v(mysssp.vert_count);
      mysssp.engine->levels_to_host(v.data());
      std::ofstream(mysssp.cfg.levels_out, std::ios::binary)
          .write(reinterpret_cast<const char*>(v.data()), v.size() * 4);
    }
    if (mysssp.cfg.check) {
      if (!mysssp.check()) return EXIT_FAILURE;
      std::cout << "check passed!" << std::endl;
    }
    return 0;
  } catch (const Error& e) {
    std::printf("%s\n", e.what());
    return EXIT_FAILURE;
  }
}
