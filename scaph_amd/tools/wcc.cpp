// wcc — command-line driver of the WCC app (app/wcc.cu:99-116): loads
// -graphfile (or -gen), runs -rounds label propagations, prints the cost;
// -check compares against the CPU label sweeps; -levels_out writes the labels.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>

#include "../host/app.h"

int main(int argc, char** argv) {
  using namespace scaph;
  try {
    wcc<uint32_t, uint32_t, uint32_t> mywcc(argc, argv);
    mywcc.load_graph();
    const int rounds = mywcc.cfg.rounds > 0 ? mywcc.cfg.rounds : 1;
    double total = 0;
    for (int i = 0; i < rounds; ++i) {
      mywcc.run();
      total += mywcc.last.ms / 1e3;
    }
    std::cout << "wcc cost " << total / rounds << " seconds " << mywcc.last.depth
              << " rounds, " << mywcc.last.m_reached << " edges relaxed" << std::endl;
    if (!mywcc.cfg.levels_out.empty()) {
      std::vector<uint32_t> v(mywcc.vert_count);
      mywcc.engine->levels_to_host(v.data());
      std::ofstream(mywcc.cfg.levels_out, std::ios::binary)
          .write(reinterpret_cast<const char*>(v.data()), v.size() * 4);
    }
    if (mywcc.cfg.check) {
      if (!mywcc.check()) return EXIT_FAILURE;
      std::cout << "check passed!" << std::endl;
    }
    return 0;
  } catch (const Error& e) {
    std::printf("%s\n", e.what());
    return EXIT_FAILURE;
  }
}
