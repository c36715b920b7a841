// Shared declarations of the persistent visit kernel (BFS and s–t).
#pragma once

#include <stdint.h>

#include "wave.hpp"

namespace fcd {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int32_t kColorBit = 1 << 30;        // colour of an s–t state word (0 = s side, 1 = t side)
constexpr int32_t kLevelMask = kColorBit - 1;
constexpr uint64_t kHubEdgeBits = 40;         // hub counter: (entries << 40) | edges
constexpr uint64_t kHubEdgeMask = (1ull << kHubEdgeBits) - 1;

enum VisitMode : int32_t { kModeBfs = 0, kModePlainSt = 1, kModeSchedSt = 2 };

// Strategy of one expansion step. Top-down steps carry a virtual-warp width.
enum Strategy : int32_t { kTopDown = 0, kBottomUp = 1 };

// Counters of one frontier (one colour, one step). Lives in a ring of 3 so a
// slot can be cleared two steps before it is appended to.
struct Ring {
  uint32_t n;      // normal-queue entries
  uint32_t pad;
  uint64_t hub;    // (hub entries << 40) | hub edges
  uint64_t mf;     // sum of degrees of all entries
};

struct VisitCtl {
  GridBar bar;
  Ring ring[2][3];
  unsigned long long best;   // (candidate distance << 32) | meet vertex; ~0 = none
  uint32_t err;              // bit 0: queue slot out of range, bit 1: bad vertex id read
  uint32_t pad;
  // Accumulators carried across launches when a visit is resumed step by step.
  uint64_t seen_edges;
  int64_t visited_v;
  int64_t visited_e;
};

constexpr int32_t FC_TIMEOUT_STATUS = 8;    // fc_status FC_ERR_TIMEOUT
constexpr int32_t FC_CAPACITY_STATUS = 6;   // fc_status FC_ERR_CAPACITY

// Written by block 0 at the end of a launch into pinned host memory.
struct VisitOut {
  int32_t status;            // 0 ok, 8 barrier timeout, 6 capacity/corruption guard
  int32_t levels;            // steps executed
  int32_t depth;             // s–t: steps until the meet (ST-CON DEPTH)
  int32_t dist;              // s–t distance, -1 if not connected
  int32_t meet;
  int32_t topdown;
  int32_t bottomup;
  int32_t max_frontier;
  int64_t visited_vertices;
  int64_t visited_edges;
  int64_t expanded_vertices;
  int64_t expanded_edges;
  int32_t lvl_end[2];        // per-colour step count at exit (resume point)
  int32_t err;
  int32_t pad;
};

// Per-step choice of the learned or heuristic scheduler (mode kModeSchedSt):
// side 0/1, strategy, and top-down width.
struct StepChoice {
  int32_t side;
  int32_t strategy;
  int32_t vw;
};

// 12 frontier features fed to the scheduler MLP.
constexpr int kFeatures = 12;
constexpr int kHidden = 32;
constexpr int kActions = 6;
constexpr int kMlpParams = kFeatures * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden * kActions + kActions;

template <typename OffT>
struct VisitParams {
  const OffT* off;
  const int32_t* col;
  int32_t V;
  int32_t hub_threshold;
  int64_t E;
  int32_t* state;
  int32_t* parent;          // nullptr: not recorded
  uint32_t* visited;        // ceil(V/64)*2 words
  uint32_t* fbm[2][3];      // frontier bitmaps, per colour, ring of 3
  int32_t* q[2][2];         // normal queues, per colour and parity
  int32_t* hq[2][2];        // hub-queue vertices
  int64_t* hb[2][2];        // hub-queue edge bases (ascending with slot)
  uint32_t qcap;            // normal-queue capacity
  uint32_t hcap;            // hub-queue capacity
  VisitCtl* ctl;
  VisitOut* out;            // pinned host memory
  int32_t src;
  int32_t dst;              // -1 for BFS
  int32_t mode;
  int32_t policy;           // scheduled mode: 1 heuristic, 2 learned
  int32_t vw_fixed;         // 0: width chosen from frontier degree
  int32_t force_strategy;   // -1: scheduler decides; else forced (tracing)
  int32_t force_side;       // -1: scheduler decides
  int32_t max_steps;        // stop after this many steps (tracing); 0 = no limit
  uint64_t spin_limit;      // barrier wait bound in 100 MHz ticks
  const float* mlp;         // kMlpParams FP32 weights, or nullptr
  // resume support for step-by-step tracing: when resume != 0 the reset
  // phase is skipped and the step counters continue from lvl0/lvl1.
  int32_t resume;
  int32_t lvl0;
  int32_t lvl1;
  int32_t trace_slot;       // index into trace_out, or -1
  float* trace_feat;        // kFeatures per traced step
};

}  // namespace fcd
