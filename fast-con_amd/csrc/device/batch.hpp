// Batched bit-parallel s–t engine: declarations shared by the kernels and
// the host driver.
#pragma once

#include <stdint.h>

#include "visit.hpp"

namespace fcd {

enum BatchPhase : int32_t {
  kBIdle = 0,
  kBInit = 1,      // seed the wave's 64 queries
  kBExpand = 2,    // top-down: frontier vertices OR their masks into neighbours
  kBCommit = 3,    // new masks -> visited, meets, next frontier lists
  kBBottomUp = 4,  // bottom-up step (expand + commit in one pass)
  kBClearF = 5,    // clear the consumed frontier masks after a bottom-up step
  kBCleanup = 6,   // zero every mask the wave touched, write results
  kBFilter = 7,    // drop resolved queries from the current frontier lists
  kBPause = 8,     // tracing: stopped at a step decision (launches are no-ops)
  kBDecide = 9,    // tracing: take the pending decision (forced or by policy)
  kBBridge = 10    // test the edges between the two sides' current frontiers
};

enum BatchSides : int32_t { kSideS = 1, kSideT = 2, kSideBoth = 3 };

struct BatchDecision {
  int32_t sides;     // kSideS / kSideT / kSideBoth
  int32_t strat;     // kTopDown / kBottomUp
  int32_t vw;        // top-down virtual-warp width
  int32_t action;    // scheduler action: sides {s, t, both}[a % 3], a < 3 top-down, else bottom-up
};

constexpr int kPhaseLog = 1024;

// Per-launch reductions go to kRedSlots copies (block b adds into slot
// b % kRedSlots) so that the blocks of a launch do not all hit one address;
// the finalising wave sums the slots and clears them.
constexpr int kRedSlots = 16;
constexpr int kRedWords = 16;  // words per slot (one 128-byte line)
enum BatchRed : int32_t {
  kROdd = 0, kREven, kRAnyS, kRAnyT,  // OR
  kRMfS, kRMfT, kRActV, kRActE, kRTouched,
  kRVisS, kRVisT,  // (vertex, query) pairs newly visited per side
  kRCount  // sum
};

struct BatchCtl {
  int32_t phase;
  int32_t wave;            // current wave index
  int32_t nwaves;
  int32_t nq;              // total queries
  int32_t policy;
  int32_t vw_fixed;
  int32_t tile;            // edges per hub chunk
  int32_t lvl[2];          // levels done by each side in this wave
  int32_t steps;
  int32_t cur[2];          // parity of the current frontier list/mask per side
  int32_t bu_sides;        // sides of the last bottom-up step (for kBClearF)
  int32_t filt;            // sides (bit x) compacted by kBFilter
  int32_t trace;           // pause at step decisions (every trace_every-th one)
  int32_t trace_every;
  int32_t decisions;       // step decisions taken so far
  int32_t force_action;    // kBDecide: action to take, -1 = the policy's
  int32_t stop_wave;       // go idle when the wave index reaches this
  int32_t pause_filter;    // the paused decision may filter first
  int32_t heur_action;     // the heuristic's action at the pause
  int32_t learn_margin_milli;  // learned policy: keep the heuristic's action unless the
                               // network predicts a log-cost lower by this (x1e-3)
  float f_branch;          // feature 4: log2(1 + sum d^2 / sum d) / 16 of the graph
  int64_t vis_pairs[2];    // (vertex, query) pairs visited per side in this wave
  int32_t live0;           // live queries at the start of the wave
  int32_t waves_done;      // waves this machine has finished
  int64_t t_decide;        // device clock at the last kBDecide
  int64_t t_wave_end;      // device clock at the last wave cleanup
  unsigned long long lact[2];  // live queries when side x's current lists were built
  BatchDecision dec;
  unsigned long long act;  // unresolved queries of the wave
  uint32_t fl_n[2][2];     // normal frontier list sizes [side][parity]
  unsigned long long hub[2][2];   // hub lists: (entries << 40) | edges
  unsigned long long fl_mf[2][2]; // degree sums of the frontier (normal + hub)
  unsigned long long filt_min_v;  // a filter runs when it drops at least this many list entries
  unsigned long long filt_min_e;  //   or this many frontier edges
  unsigned long long both_max_e;  // heuristic: both sides while their frontiers have at most this many edges
  unsigned long long bridge_min_e;  // a bridge step runs before a step whose cheaper side has this many live edges
  int32_t bridge_lv;       // (lvl[0] + 1) << 16 | (lvl[1] + 1) of this wave's last bridge step (0: none)
  int32_t bridges;         // bridge steps (whole batch)
  int64_t bridged;         // queries answered by bridge steps (whole batch)
  uint32_t ticket;
  uint32_t err;
  // statistics (whole batch)
  int64_t levels_total;
  int64_t active_vertices;  // sum over steps of frontier vertices processed
  int64_t active_edges;     // sum over steps of their degrees
  int64_t touched_total;
  int64_t filters;          // filter launches
  int64_t phases;           // launches that did work
  float feat[12];
  int8_t plog[kPhaseLog];   // phase of each working launch (diagnostics)
  int64_t ptime[kPhaseLog];  // wall clock at its finalisation
  int64_t pwork[kPhaseLog];  // frontier edges (expand, bottom-up) or touched vertices (commit) of it
};

template <typename OffT>
struct BatchParams {
  const OffT* off;
  const int32_t* col;
  int32_t V;
  int32_t hub_threshold;
  int64_t E;
  // Per-side records of two 64-bit words, side-major: rec[2V*x + 2u] =
  // side-x visited mask of u and rec[2V*x + 2u + 1] = side-x masks ORed in
  // during the current top-down step. One 16-byte load gives a probe both
  // words it needs, and a line holds eight vertices' records of one side.
  unsigned long long* rec;     // [4*V]
  unsigned long long* F[2][2]; // frontier masks [side][parity], V entries each
  uint8_t* tm;                 // touched marks: one byte per 32 vertices whose N words may be non-zero
  unsigned long long* red;     // [kRedSlots][kRedWords] launch reductions
  uint32_t* db;                // dirty bitmap: visited words of u non-zero this wave
  int32_t* fl[2][2];           // frontier lists [side][parity], V entries each
  int32_t* hl[2][2];           // hub frontier lists (degree >= hub_threshold)
  int64_t* hb[2][2];           // their edge bases (ascending with slot)
  uint32_t hcap;
  BatchCtl* ctl;
  const int32_t* qs;           // queries (device), caller ids
  const int32_t* qt;
  const int32_t* perm;         // caller id -> device id (null: the same)
  int32_t* dist;               // results (device), nq entries
  uint32_t* wave_ticket;       // next unclaimed wave, shared by concurrent machines (null: this
                               // machine takes waves wave, wave + 1, ... up to stop_wave)
  const float* mlp;
  int32_t v_active;            // vertices [0, v_active) have degree > 0 (the rest are isolated)
  int32_t sorted_lists;        // every adjacency list is in ascending id order (bridge steps search them)
};

template <typename OffT>
hipError_t launch_batch_phase(const BatchParams<OffT>& p, int blocks, hipStream_t s);
template <typename OffT>
int batch_occupancy();

}  // namespace fcd
