// engine.hpp -- data structures shared by the host launch code and the gfx950
// kernels of the SOSP / combined-graph engine.
//
// One persistent cooperative launch runs a whole update: Step 1 (pack, roots,
// pointer-jumping subtree invalidation, pull) and Step 2 (near-far push) for
// up to MOSP_MAX_OBJ objectives at once, optionally followed by the combined
// graph build and its from-scratch search (Steps 2-3 of the MOSP update).
// The semantics follow the reference's engine (src/sospUpdateGpu.cu:211-519,
// src/combinedGraphGpu.cu:55-121); the schedule is this engine's own: all
// objectives advance in lock-step with ONE grid barrier per near-far
// iteration (rotating counter slots instead of the reference's two barriers),
// and the far-pile minimum is tracked at append time instead of by an extra
// reduction pass.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mosp {

using u64 = unsigned long long;

constexpr u64 PACKED_INF = ~0ULL;
constexpr long long DISTANCE_INF = 0x7fffffffffffffffLL / 4;  // csrGraph.cuh:185
constexpr int MOSP_MAX_OBJ = 32;     // K <= 32 (csrGraph.cu:736-739)
#ifndef MOSP_BLOCK
#define MOSP_BLOCK 1024
#endif
constexpr int BLOCK = MOSP_BLOCK;    // threads per workgroup (wave64 multiple)
// LDS queue entries per list and workgroup (the most a launch can use).
#define MOSP_QUEUE_CAP_MAX 4096
#ifndef MOSP_MIN_BLOCKS
#define MOSP_MIN_BLOCKS 1            // co-resident workgroups per CU the code is built for
#endif
// Vertices of the largest graph mosp_sosp_update_noalloc takes: its scratch
// (36 bytes per vertex) is a static device array of that size.
#ifndef MOSP_NOALLOC_MAX_VERTICES
#define MOSP_NOALLOC_MAX_VERTICES (1 << 21)
#endif
// Graphs with a vertex of more than HUB_DEGREE out- or in-edges run the
// kernel variant that hands such vertices to whole waves.
constexpr int HUB_DEGREE = 32;
// Hub variant: frontier vertices of more than HUB_GRID out-edges are listed
// for the next near-far iteration, in which every workgroup of their
// objective relaxes an even share of their edges (at most HUB_LIST of them
// per iteration and objective; the others are relaxed by one workgroup).
constexpr int HUB_GRID = 4096;
constexpr int HUB_LIST = 512;

// Phases of one launch.
enum : int {
  PH_UPDATE = 1,      // Step 1 + Step 2 on objs[0..nobj)
  PH_SCRATCH = 2,     // from-scratch search on objs[0] (Step 3 engine)
  PH_COMBINED = 4,    // combined build from cParents + from-scratch search
};

// Phase stamps recorded by the persistent kernel (Ctl::clk index).
enum : int {
  PHASE_PACK = 0, PHASE_JUMP, PHASE_INVALIDATE, PHASE_PULL, PHASE_NEARFAR,
  PHASE_UNPACK, PHASE_COMB_COUNT, PHASE_COMB_SCAN, PHASE_COMB_FILL,
  PHASE_COMB_SEARCH, PHASE_COMB_UNPACK, PHASE_COUNT
};

// Scratch of one objective (a slice of the workspace).
struct ObjState {
  u64 *packed;     // (distance << b | parent) words, n
  int *stamp;      // generation a vertex was last listed in, n
  int *inFar;      // vertex is in the far pile, n (left all-zero)
  int *ancestor;   // pointer-jumping ancestors, n
  int *near[2];    // near frontiers (ping-pong), n each
  int *far[2];     // far piles (ping-pong), n each
  int *hubs;       // grid-relaxed hubs of the next iteration (ping-pong), 2 * HUB_LIST
  int *pool;       // asynchronous near-far: ASYNC_SEGS shared pool segments of
                   // LaunchArgs::segCap entries each, all -1 between launches
                   // (null: the workspace has no pools)
};

// One objective of a launch (device-resident descriptor).
struct ObjDesc {
  const int *outRow, *outCol, *outW;  // out-CSR + this objective's weights
  const int *inRow, *inCol, *inW;     // in-CSR + weights (update only)
  // The same edges as {column, weight} records (one 8-byte load per edge),
  // when the graph holds them (device graphs, the combined graph); null:
  // the separate arrays above are read.
  const int2 *outEdge, *inEdge;
  // The out-edges as {column, weight, rowBegin(column), rowEnd(column)}
  // (asynchronous near-far: a vertex listed by a relaxation carries its own
  // edge range), or null.
  const int4 *outEdge4;
  const int *chFrom, *chTo;           // deleted / weight-increased edges
  int nChanged;
  const int *heads;                   // heads of inserted edges
  int nHeads;
  int parentBits;                     // 0: distance-only words
  int optimistic;                     // maxDistance is below (n-1)*maxWeight: a dropped
                                      // candidate sets ObjCtl::dropped
  u64 noParent;                       // all-ones parent field
  u64 maxDistance;                    // larger candidates are dropped: (n-1)*maxWeight,
                                      // or the packed word's limit when optimistic
  u64 delta;                          // near-far bucket width
  const int64_t *distIn;              // old tree (update) -- may alias distOut
  const int *parentIn;
  int64_t *distOut;                   // new tree
  int *parentOut;
  ObjState s;
};

// Per-objective device counters of one launch (zeroed by the reset kernel).
struct ObjCtl {
  unsigned nearCount[3];   // near frontier sizes, slot = iteration % 3
  unsigned farCount[3];    // far pile sizes, slot = epoch % 3
  u64 farMin[3];           // min distance appended to the far pile of an epoch
  unsigned invalidated;
  unsigned overflow;       // bit 0: an input distance does not fit the packing;
                           // bit 1: an input parent is not in [-1, n)
  unsigned iterations;
  unsigned epochs;
  unsigned jumpCount[3];   // pointer jumping: still-jumping list, slot = round % 3
  unsigned pullHeavy;      // pull candidates left for whole workgroups (hub variant)
  unsigned pullCount;      // pull candidates listed for lane groups (hub variant)
  unsigned hubCount[3];    // grid-relaxed hubs listed for an iteration, slot = iteration % 3
  unsigned dropped;        // optimistic packing: a candidate or input distance was past
                           // maxDistance (the launch stops before writing results)
  unsigned lost[2];        // asynchronous near-far: a near entry of phase p found no room
                           // (slot p % 2); the next phase lists every vertex under the
                           // threshold again
  u64 pushes;
};

constexpr int BARRIER_GROUPS = 16;  // first-level arrival counters
constexpr int LINE_WORDS = 32;      // one 128-byte line per counter
// Words a barrier can publish with its epoch: five per objective (near count,
// far count, far minimum as two words, listed hubs) for MOSP_MAX_OBJ + 1
// searches.
constexpr int MAIL_WORDS = 5 * (MOSP_MAX_OBJ + 1);
// One mailbox per barrier group: the epoch line, then the published words.
constexpr int MAIL_STRIDE =
    LINE_WORDS * ((2 * LINE_WORDS + MAIL_WORDS - 1) / LINE_WORDS);

struct Ctl {
  // Two-level grid barrier: workgroup b arrives at sub[b % 16]; the last of
  // each group arrives at top; the last arrival publishes the epoch (and the
  // words the barrier carries) in every group's mailbox, which the group's
  // workgroups poll.  All only grow within a launch (zeroed by the reset
  // kernel).
  unsigned barrierSub[BARRIER_GROUPS * LINE_WORDS];
  unsigned barrierTop[LINE_WORDS];
  unsigned jumpActive[3];  // pointer jumping: some vertex still jumps
  unsigned rounds;
  unsigned generationOut;  // last stamp generation used
  u64 combEdges;           // combined graph: edge count
  u64 combWeightSum;       // combined graph: weight sum
  u64 combDelta;           // near-far width used for Step 3
  u64 clkStart, clkSospEnd, clkEnd;  // wall_clock64 stamps
  u64 clk[16];             // phase ends (PHASE_* below), wall_clock64
  unsigned barriers;       // grid barriers executed by the launch
  unsigned gaveUp;         // a workgroup's barrier wait ran out (the launch ended early)
  // Asynchronous near-far events: batches that dropped near entries (full
  // local queue), spills that found less room in their pool segment than
  // asked, phases that listed every vertex again.
  unsigned asyncLost, asyncShortSpills, asyncSweeps;
  u64 prof[16];            // sub-phase clock sums (MOSP_SUBPROFILE builds)
  ObjCtl obj[MOSP_MAX_OBJ + 1];
  // Barrier mailboxes (device-only: not part of the host read-back).
  alignas(256) unsigned mail[BARRIER_GROUPS * MAIL_STRIDE];
};

// Bytes of Ctl the host reads back after a launch (everything but mail).
constexpr size_t CTL_READBACK = offsetof(Ctl, mail);

// Asynchronous near-far (the single-update kernel without hub handling):
// between two grid barriers each workgroup relaxes the vertices of a local
// LDS queue, spills its surplus into shared pools and takes work from them
// when its queue runs dry; a phase ends when every appended entry has been
// processed (counted per workgroup, checked by workgroup 0).
constexpr int ASYNC_SEGS = 8;         // pool segments per objective (one per XCD)
constexpr u64 ASYNC_STATUS_INIT = 15ull << 60;  // a status word of no phase
// LaunchArgs::asyncFlags
constexpr unsigned ASYNC_CARRY = 1;   // local entries carry the word they were listed with
constexpr unsigned ASYNC_HUNGRY = 2;  // spill to XCDs with idle workgroups
constexpr unsigned ASYNC_FLAG_MASK = ASYNC_CARRY | ASYNC_HUNGRY;
constexpr int ASYNC_MAX_GRID = 2048;  // workgroups a launch with pools may have
// Entries of a workgroup's local queue (LDS ring) for `threads` threads.
constexpr unsigned async_ring_entries(int threads) { return threads >= 1024 ? 2048u : 1024u; }

struct AsyncCtl {
  // Per workgroup: phase (4 bits: the launch's phase sequence number mod 15)
  // | entries appended (30) | entries processed (30), counted from the start
  // of the phase; written only by its workgroup.
  u64 status[ASYNC_MAX_GRID];
  // Pool words {tail << 32 | head}, one 128-byte line each; two sets, used
  // by alternate phases (a set is cleared while the other is in use).
  u64 pool[2][MOSP_MAX_OBJ][ASYNC_SEGS][LINE_WORDS / 2];
  // Phase-done mailboxes: the launch's phase sequence number + 1 once that
  // phase ended; one line per barrier group.
  unsigned done[BARRIER_GROUPS][LINE_WORDS];
  // Idle workgroups per XCD (two sets, as the pool words): busy workgroups
  // spill into the segment of an XCD that has some.
  unsigned hungry[2][ASYNC_SEGS][LINE_WORDS];
};

struct LaunchArgs {
  const ObjDesc *objs;     // nobj descriptors (device memory)
  int nobj;
  int n;
  int source;
  int phases;
  int maxRounds;           // pointer-jumping rounds: ceil(log2 n) + 1
  int anyFallback;         // some objective of objs uses distance-only words
  int anyOptimistic;       // some objective of objs uses optimistic packing
  int laneBits;            // 2^laneBits lanes share a frontier vertex
  int hopsUpdate;          // local relaxation rounds per near-far iteration: K updates
  int hopsSearch;          // ... from-scratch and combined searches (kernels
                           // without hub handling only)
  int listCap;             // entries of each near / far list: 0 = n; a smaller
                           // value (tests) forces the sweeps that follow a full list
  int hubs;                // host only: launch the HUB_DEGREE-aware variant
  int queueCap;            // LDS queue entries the push loop uses: 256..MOSP_QUEUE_CAP_MAX
                           // (0: all of them)
  int initPush;            // Step 1 candidates of invalidated vertices: 0 pushed from the
                           // valid vertices when more than half are invalidated (graphs
                           // without hubs), else pulled; 1 always pulled; 2 always pushed
  int hubEdges;            // hub variant: push vertices of more out-edges go to
                           // the objective's hub list (HUB_GRID; tests lower it)
  int *genDev;             // device-resident stamp generation (persists)
  Ctl *ctl;
  // Asynchronous near-far (kernels without hub handling, single update):
  // its counters, or null for the iteration-synchronous loop; entries per
  // pool segment; spill thresholds of the local queue (entries).
  AsyncCtl *async;
  unsigned segCap;
  unsigned spillHi, spillLo;
  unsigned asyncFlags;     // ASYNC_CARRY | ASYNC_HUNGRY
  unsigned keepHungry;     // entries a spill to idle workgroups leaves in the local queue
  // Bounded barrier waits: a workgroup that waits longer than barrierTicks
  // (wall_clock64 ticks) at a grid barrier, or finds *abortWord set while it
  // waits, sets *abortWord and ends; so does every other workgroup at its
  // next wait.  A launch whose workgroups are not all resident then ends
  // with an error instead of spinning for ever.
  unsigned long long barrierTicks;
  unsigned *abortWord;
  // Combined graph (PH_COMBINED): parents of cK trees, objective-major.
  const int *cParents;
  int cK;
  int cBase;               // L * (K + 1)
  int cTerms[MOSP_MAX_OBJ];// L / Pref_k
  long long cDelta;        // > 0: use it; else the default rule on device ...
  int cDeltaScale;         // ... times this factor (>= 1)
  int *cDegree;            // n + 1 (counts, then fill cursors)
  int *cSlot;              // K * n, or nullptr: a combined edge's place in its parent's
                           // row, taken by the hub variant's count (slot[k * n + v] for
                           // the edge from v's parent in tree k), so its fill needs no
                           // atomics
  int *cRowPtr;            // n + 1 (written here, read through cDesc->outRow)
  int *cColInd;            // K * n
  int *cWeights;           // K * n
  int2 *cEdge;             // K * n {column, weight} records (the memory of cColInd + cWeights)
  int4 *cEdge4;            // K * n records with the column's row range as well, written when
                           // the launch runs the asynchronous loop (else null)
  int *cPartials;          // one per workgroup (scan)
  // The combined graph as a mask over the graph's out-edges (fused launches
  // of graphs without hubs that run the asynchronous loop; null otherwise):
  // every tree edge (p, v) is an out-edge of p, so the combined row of p is
  // part of p's row in the graph.  cMaskRec[e] for edge e = (p, v): {v,
  // combined weight, rowBegin(v), rowEnd(v)} when p is a parent of v (first
  // such edge of p's row), else {source, 0, 0, 0} (the source is never
  // relaxed); cDesc then names the graph's rows and these records.
  int4 *cMaskRec;
  const int4 *cMaskSrc;    // the graph's out-edge records with row ranges
  const int *cMaskRow;     // the graph's out-row pointers
  int4 *cPar4;             // with cMaskRec, 2 <= K <= 4 and no distance-only objective:
                           // each vertex's K new parents side by side (-1 beyond K),
                           // written by the unpack, so the mask pass reads one 16-byte
                           // record per edge instead of K words in K arrays (n records;
                           // may alias cColInd); nullptr: the pass reads cParents
  u64 *stage;              // nobj * n words (may alias cColInd..cWeights, which are
                           // written only after it is read): the fused parent
                           // recovery's per-vertex records; nullptr: none
  const ObjDesc *cDesc;    // search on the combined CSR: outRow = row pointers,
                           // outCol / outW = edges, packing, scratch, outputs
};

// A group of workgroups of a batched launch and the instances it runs, in
// order: LaunchArgs [first, last) of the launch's instance array.
struct BatchGroup {
  int size;   // workgroups in the group
  int first;
  int last;
};

}  // namespace mosp
