// paged.h — state of the out-of-core (paged) BFS mode.
#ifndef SCAPH_HOST_PAGED_H
#define SCAPH_HOST_PAGED_H

#include "engine.h"

namespace scaph {

struct Engine::PagedState {};

}  // namespace scaph

#endif
