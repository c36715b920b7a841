// Shared declarations of the visit engine (BFS and s–t): the control block of
// the device-side phase machine and the launch parameters.
#pragma once

#include <math.h>
#include <stdint.h>

#include "wave.hpp"

namespace fcd {

constexpr int kBlock = 256;
constexpr int kTicketGroups = 16;  // two-level launch ticket: block b counts in group b % 16
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int32_t kColorBit = 1 << 30;        // colour of an s–t state word (0 = s side, 1 = t side)
constexpr int32_t kLevelMask = kColorBit - 1;
constexpr uint64_t kHubEdgeBits = 40;         // hub counter: (entries << 40) | edges
constexpr uint64_t kHubEdgeMask = (1ull << kHubEdgeBits) - 1;

enum VisitMode : int32_t { kModeBfs = 0, kModePlainSt = 1 };
enum Strategy : int32_t { kTopDown = 0, kBottomUp = 1 };
enum Phase : int32_t { kPhaseIdle = 0, kPhaseReset = 1, kPhaseLevel = 2 };

constexpr int32_t FC_TIMEOUT_STATUS = 8;    // fc_status FC_ERR_TIMEOUT
constexpr int32_t FC_CAPACITY_STATUS = 6;   // fc_status FC_ERR_CAPACITY

// Counters of one frontier (one colour, one step). Ring of 3 per colour.
struct Ring {
  uint32_t n;      // normal-queue entries
  uint32_t pad;
  uint64_t hub;    // (hub entries << 40) | hub edges
  uint64_t mf;     // sum of degrees of all entries
};

// What the next level launch does, per colour.
struct Decision {
  int32_t ex[2];      // expand this colour
  int32_t strat[2];   // kTopDown / kBottomUp
  int32_t vw[2];      // top-down virtual-warp width
};

// Result of one query (device array, copied back by the host).
struct VisitOut {
  int32_t status;            // 0 ok, 6 capacity/corruption guard
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
  int32_t lvl_end[2];
  int32_t err;
  int32_t launches;          // working launches of this query (its reset included)
};

// Control block of the phase machine. Written by the finalising block of one
// launch, read by every block of the next (kernel boundaries order them).
struct VisitCtl {
  int32_t phase;
  int32_t qi;               // current query
  int32_t nq;               // number of queries
  int32_t mode;
  int32_t launches;         // working launches of the current query
  int32_t vw_fixed;         // 0: width from frontier degree
  int32_t record_parents;
  int32_t pad1;
  int32_t src, dst;
  int32_t lvl[2];
  int32_t steps, td, bu, maxf;
  int32_t bu_mode;
  int32_t pad0;
  Decision dec;
  Ring ring[2][3];
  unsigned long long best;  // (candidate distance << 32) | meet vertex; ~0 = none
  uint32_t ticket;
  uint32_t err;             // bit 0: queue slot out of range, bit 1: bad vertex id read
  uint32_t gticket[kTicketGroups * 32];  // group tickets, one 128-byte line each
  uint64_t seen_edges;      // sum of degrees of every vertex labelled so far
  int64_t visited_v, visited_e, expv, expe;
};

constexpr int kFeatures = 12;
constexpr int kHidden = 32;
constexpr int kActions = 9;  // sides {s, t, both} x {top-down, top-down width 64, bottom-up}
constexpr int kMlpParams = kFeatures * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden * kActions + kActions;

template <typename OffT>
struct VisitParams {
  const OffT* off;
  const int32_t* col;
  int32_t V;
  int32_t hub_threshold;
  int64_t E;
  int32_t* state;
  int32_t* parent;
  uint32_t* visited;        // ceil(V/64)*2 words
  uint32_t* fbm[2][3];      // frontier bitmaps, per colour, ring of 3
  int32_t* q[2][2];         // normal queues, per colour and parity
  int32_t* hq[2][2];        // hub-queue vertices
  int64_t* hb[2][2];        // hub-queue edge bases (ascending with slot)
  uint32_t qcap;
  uint32_t hcap;
  VisitCtl* ctl;
  const int32_t* qs;        // query sources (device), nq entries
  const int32_t* qt;        // query targets
  VisitOut* out;            // results (device), nq entries
  int32_t lazy_state;       // BFS leaves unreached state/parent unwritten (renumbered graphs:
                            // the un-permutation takes reached-ness from the visited bitmap)
  int32_t v_active;         // vertices [0, v_active) have degree > 0 (the rest are isolated)
};

}  // namespace fcd
