// engine_internal.h — engine state shared by the engine's source files.
#ifndef SCAPH_HOST_ENGINE_INTERNAL_H
#define SCAPH_HOST_ENGINE_INTERNAL_H

#include "engine.h"
#include "paged.h"

namespace scaph {

// Device copy of the learned schedule: bf16 parameters, normalisation
// (mean[16], scale[16]), per-page counts and the decision record.
struct Engine::DeviceModel {
  DevBuf params;    // SCAPH_MLP_NPARAM u16
  DevBuf norm;      // 32 f32
  DevBuf page_cnt;  // u32 per page
  DevBuf out;       // 32 f32 (logit, threshold, pages active, pad, hist[12], feat[16])
  DevBuf ccnt;      // a counters record describing the frontier entering a level
  PinnedBuf hout;   // host landing zone for out
  PinnedBuf hccnt;
  bool ready = false;
  // the model's W1 columns for the 12 histogram features are zero: the
  // decision needs no frontier pass and rides on the level hand-off
  bool scalar_only = false;
};

}  // namespace scaph

#endif
