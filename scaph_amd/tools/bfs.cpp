// bfs — command-line driver of the BFS app (app/bfs.cu:119-216).
// Prints the reference's banners; runs -rounds timed traversals from -src;
// -check compares against the CPU BFS; -levels_out writes the levels.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <fstream>

#include "../host/app.h"

int main(int argc, char** argv) {
  using namespace scaph;
  try {
    std::printf("|=============================Execute===============================|\n");
    bfs<uint32_t, uint32_t, uint32_t> mybfs(argc, argv);
    mybfs.load_graph();
    const int rounds = mybfs.cfg.rounds > 0 ? mybfs.cfg.rounds : 1;
    std::ofstream trace;
    if (!mybfs.cfg.trace.empty()) trace.open(mybfs.cfg.trace);
    const auto t1 = std::chrono::steady_clock::now();
    for (int i = 0; i < rounds; ++i) {
      mybfs.run();
      if (trace) {
        for (const auto& r : mybfs.last.levels) {
          trace << "{\"round\":" << i << ",\"level\":" << r.level << ",\"dir\":" << r.dir
                << ",\"n_f\":" << r.n_f << ",\"m_f\":" << r.m_f << ",\"m_u\":" << r.m_u
                << ",\"n_next\":" << r.n_next << ",\"m_next\":" << r.m_next
                << ",\"scanned\":" << r.scanned << ",\"ms\":" << r.ms << "}\n";
        }
      }
    }
    const auto t2 = std::chrono::steady_clock::now();
    const double avg = std::chrono::duration<double>(t2 - t1).count() / rounds;
    const double label = mybfs.cfg.threshold * 12.5;  // 32 chunks = 100%, printed as k/8
    std::printf("|========================Results(alpha=%.2f%%)======================|\n", label);
    std::printf("|      Avgrage(%d Rounds) BFS costs     : %.8f seconds       |\n", rounds, avg);
    std::printf("|      Avgrage(%d Rounds) Throughput    : %.8f GTEPS          |\n", rounds,
                Total_throughput / rounds);
    std::printf("|===================================================================|\n");
    if (!mybfs.cfg.levels_out.empty()) {
      std::vector<uint32_t> lv(mybfs.vert_count);
      mybfs.engine->levels_to_host(lv.data());
      if (mybfs.cfg.compat63)
        for (auto& x : lv) x = compat63(x);
      std::ofstream out(mybfs.cfg.levels_out, std::ios::binary);
      out.write(reinterpret_cast<const char*>(lv.data()), lv.size() * 4);
    }
    if (mybfs.cfg.check) {
      if (!mybfs.check()) return EXIT_FAILURE;
      std::printf("check passed!\n");
    }
    return 0;
  } catch (const Error& e) {
    std::printf("%s\n", e.what());
    return EXIT_FAILURE;
  }
}
