// Shared declarations of the visit engine (BFS and s–t): the control block of
// the device-side phase machine and the launch parameters.
#pragma once

#include <math.h>
#include <stdint.h>

#include "wave.hpp"

namespace fcd {

constexpr int kBlock = 256;
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

// Counters of one frontier (one colour, one level): appended to with device
// atomics by the launch that builds the frontier, read by the next.
struct Ring {
  uint32_t n;      // normal-queue entries
  uint32_t pad;
  uint64_t hub;    // (hub entries << 40) | hub edges
  uint64_t mf;     // sum of degrees of all entries
};

struct VisitCtl;

// Result of one query (device array, copied back by the host).
struct VisitOut {
  int32_t status;            // 0 ok, 6 capacity/corruption guard
  int32_t levels;            // expansion steps executed
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
  int32_t err;
  int32_t launches;          // working launches of this query
};

// Machine state between two launches. Launches alternate a parity p (a
// kernel argument): every block of a launch reads state st[p], and block 0
// writes the state of the next launch into st[p ^ 1], which no block of the
// same launch reads. All blocks derive the same step from st[p] and the
// counters below, so no block waits for another and no launch ends with a
// vote for a finalising block.
struct VisitState {
  int32_t phase;
  int32_t qi;               // current query
  int32_t src, dst;         // its roots (device ids)
  int32_t nsrc, ndst;       // the next query's roots (-1: none)
  int32_t lvl;              // frontier level this launch expands (both colours advance together)
  uint32_t epoch;           // s–t: tag of the current query in the state words
  int32_t steps, td, bu, maxf;
  int32_t launches;
  int32_t bu_mode;          // BFS: direction-optimising switch state
  int32_t pad[2];
  uint64_t seen_edges;      // sum of degrees of the vertices labelled before this level
  int64_t visited_v, visited_e, expv, expe;
  int64_t pad2[3];
};
static_assert(sizeof(VisitState) == 128, "VisitState layout");

// Pointers into device memory. In the device pass they carry the global
// address space: the kernel reads them from its LDS copy of the control
// block, and without it the compiler cannot tell them from LDS pointers. It
// then emits flat instructions, each of which also counts as an LDS access
// (every wait for one waits for the LDS too), and reloads the pointers from
// LDS after each flat store, since such a store might have written there.
#if defined(__HIP_DEVICE_COMPILE__)
#define FC_GLOBAL __attribute__((address_space(1)))
#else
#define FC_GLOBAL
#endif

// Host code stores a device pointer into such a field with an explicit cast.
template <typename F, typename T>
inline void set_param(F& field, T* v) {
  field = (F)v;
}

// Launch parameters: the graph and workspace pointers. They travel inside
// the control block (the kernel's only pointer argument), so that every
// block gets them with the same 1 KB load as the machine state instead of
// with the scalar loads of a by-value argument, which the compiler issues
// lazily and waits for one by one.
template <typename OffT>
struct VisitParams {
  FC_GLOBAL const OffT* off;
  FC_GLOBAL const int32_t* col;
  int32_t V;
  int32_t hub_threshold;
  int64_t E;
  FC_GLOBAL int32_t* state;           // BFS: levels
  FC_GLOBAL unsigned long long* sst;  // s–t: (epoch << 32) | colour | level; another epoch = unvisited
  FC_GLOBAL int32_t* parent;
  FC_GLOBAL uint32_t* visited;        // ceil(V/64)*2 words
  FC_GLOBAL uint32_t* fbm[2][3];      // frontier bitmaps, per colour, ring of 3
  FC_GLOBAL int32_t* q[2][2];         // normal queues, per colour and parity
  FC_GLOBAL int32_t* hq[2][2];        // hub-queue vertices
  FC_GLOBAL int64_t* hb[2][2];        // hub-queue edge bases (ascending with slot)
  uint32_t qcap;
  uint32_t hcap;
  FC_GLOBAL VisitCtl* ctl;
  FC_GLOBAL const int32_t* qs;        // query sources (device), nq entries
  FC_GLOBAL const int32_t* qt;        // query targets
  FC_GLOBAL VisitOut* out;            // results (device), nq entries
  int32_t lazy_state;       // BFS leaves unreached state/parent unwritten (renumbered graphs:
                            // the un-permutation takes reached-ness from the visited bitmap)
  int32_t v_active;         // vertices [0, v_active) have degree > 0 (the rest are isolated)
};

// Control block: 1 KB, loaded whole by every block at the start of a launch
// (one 16-byte load per lane of one wave).
struct VisitCtl {
  VisitState st[2];                  // by launch parity
  Ring ring[2][2][3];                // [query parity][colour][level % 3]
  unsigned long long best[2];        // by query parity: (distance << 32) | meet vertex; ~0 = none
  uint32_t err;                      // bit 0: queue slot out of range, bit 1: bad vertex id read,
                                     // bit 2: an s–t claim did not settle
  int32_t nq;                        // number of queries
  int32_t mode;
  int32_t vw_fixed;                  // 0: width from frontier degree
  int32_t record_parents;
  int32_t pad[31];
  alignas(16) unsigned char params[320];  // VisitParams<OffT> of the run
};
static_assert(sizeof(VisitCtl) == 1024, "VisitCtl layout");
static_assert(sizeof(VisitParams<int64_t>) <= 320 && sizeof(VisitParams<int32_t>) <= 320, "params fit the block");

constexpr int kFeatures = 12;
constexpr int kHidden = 32;
constexpr int kActions = 9;  // sides {s, t, both} x {top-down, top-down width 64, bottom-up}
constexpr int kMlpParams = kFeatures * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden * kActions + kActions;

// s–t state words.
__host__ __device__ inline unsigned long long st_word(uint32_t epoch, int c, int32_t lvl) {
  return ((unsigned long long)epoch << 32) | (uint32_t)((c ? kColorBit : 0) | lvl);
}

}  // namespace fcd
