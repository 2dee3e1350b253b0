// engine_paged.cpp — out-of-core (paged) BFS: a bounded pool of device page
// slots caches host pages; per level the host classifies every active page
// as cached / full copy / compacted copy and streams the missing ones from
// pinned memory (src/pagehandle.h, src/compress.h, src/app.cuh:395-468).
#include "paged.h"

namespace scaph {

void Engine::setup_paged() {
  throw Error(SCAPH_STATUS_INVALID, "out-of-core paged mode is not available in this build");
}

RunStats Engine::bfs_paged(uint32_t) {
  throw Error(SCAPH_STATUS_INVALID, "out-of-core paged mode is not available in this build");
}

}  // namespace scaph
