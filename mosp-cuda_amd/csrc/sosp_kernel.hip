// sosp_kernel.hip -- the persistent cooperative gfx950 kernel of the update.
//
// Semantics (what each phase computes) follow the reference's engine,
// src/sospUpdateGpu.cu:211-519 (Step 1 + near-far Step 2 + unpack / parent
// recovery) and src/combinedGraphGpu.cu:55-121 (combined graph).  The
// schedule is this kernel's own (see engine.hpp):
//   * up to 32 objectives are updated in ONE launch, in lock-step, so every
//     grid barrier serves all of them;
//   * one grid barrier per near-far iteration: counters rotate over three
//     slots (iteration % 3 for near frontiers, epoch % 3 for far piles) so a
//     slot is cleared two iterations before it is reused;
//   * the far-pile minimum is folded in when a vertex is appended (a wave
//     reduction + one atomicMin per wave), so raising the threshold needs no
//     separate reduction pass and barrier;
//   * list appends are wave64-aggregated (ballot + popcount + one atomicAdd
//     per wave, slot broadcast with readfirstlane);
//   * the grid barrier is a monotonic arrival counter with agent-scope
//     release/acquire (no reset between barriers).
// No MFMA: this is integer graph work bound by HBM latency and bandwidth.
#include <hip/hip_runtime.h>

#include <climits>

#include "engine.hpp"

namespace mosp {
namespace {

enum : int { MODE_DONE = 0, MODE_PUSH = 1, MODE_SPLIT = 2 };

// ---------------------------------------------------------------------------
// Device helpers
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned load_relaxed(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ u64 load_relaxed(const u64 *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(unsigned *p, unsigned v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(u64 *p, u64 v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for a cooperative launch (all workgroups co-resident).
// Two levels so that no counter sees more than ~gridDim/16 same-address
// atomics per barrier: workgroup b arrives at sub[b % 16]; the last arrival
// of a group arrives at top; the last group stores the epoch in flag, which
// every workgroup polls with relaxed loads (an acquire per poll would
// invalidate the XCD's L2 on every poll) before one acquire fence.  The
// counters only grow: barrier e completes when sub reaches e * groupSize and
// top reaches e * groups, so nothing is reset inside a launch.  The last
// arrivals re-fence (acquire + release) before passing the arrival on, so
// every workgroup's writes happen-before every workgroup's later reads.
struct GridBarrier {
  Ctl *ctl;
  unsigned epoch;

  __device__ void sync() {
    __syncthreads();
    ++epoch;
    if (threadIdx.x == 0) {
      const unsigned G = gridDim.x;
      const unsigned groups = G < BARRIER_GROUPS ? G : BARRIER_GROUPS;
      const unsigned g = blockIdx.x % BARRIER_GROUPS;
      const unsigned groupSize =
          G / BARRIER_GROUPS + (g < G % BARRIER_GROUPS ? 1u : 0u);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      const unsigned sub = __hip_atomic_fetch_add(
          &ctl->barrierSub[g * LINE_WORDS], 1u, __ATOMIC_RELAXED,
          __HIP_MEMORY_SCOPE_AGENT);
      if (sub + 1 == epoch * groupSize) {
        __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
        const unsigned top = __hip_atomic_fetch_add(
            &ctl->barrierTop[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (top + 1 == epoch * groups) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
          store_relaxed(&ctl->barrierFlag[0], epoch);
        }
      }
      while (load_relaxed(&ctl->barrierFlag[0]) < epoch) {
        __builtin_amdgcn_s_sleep(1);
      }
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
  }
};

// Reserve one slot of a shared list for every active lane: one atomicAdd
// per wave; the first active lane adds the lane count and its old value is
// broadcast with readfirstlane.  The counter must be wave-uniform.
__device__ __forceinline__ unsigned wave_append(unsigned *counter) {
  const u64 active = __ballot(1);
  const unsigned lane = __lane_id();
  const unsigned rank = __popcll(active & ((1ull << lane) - 1ull));
  unsigned base = 0;
  if (rank == 0) {
    base = atomicAdd(counter, static_cast<unsigned>(__popcll(active)));
  }
  base = static_cast<unsigned>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(base)));
  return base + rank;
}

// Minimum over a converged wave, folded into *target by lane 0.
__device__ __forceinline__ void wave_min_into(u64 value, u64 *target) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    const u64 other = __shfl_xor(value, offset);
    value = other < value ? other : value;
  }
  if (__lane_id() == 0 && value != PACKED_INF) {
    atomicMin(target, value);
  }
}

__device__ __forceinline__ bool claim(int *stamp, int v, int generation) {
  return atomicExch(&stamp[v], generation) != generation;
}

__device__ __forceinline__ u64 pack(u64 distance, int parent, int bits,
                                    u64 noParent) {
  if (bits == 0) {
    return distance;
  }
  return (distance << bits) | (parent < 0 ? noParent : static_cast<u64>(parent));
}

__device__ __forceinline__ u64 dist_of(u64 word, int bits) {
  return word >> bits;
}

// Stamp the end of a phase (called right after a grid barrier).
__device__ __forceinline__ void stamp(Ctl *ctl, int phase, unsigned tid) {
  if (tid == 0) {
    ctl->clk[phase] = wall_clock64();
  }
}

struct Threads {
  unsigned id;     // global thread index
  unsigned count;  // threads in the grid
};

// ---------------------------------------------------------------------------
// Near-far propagation (Step 2), all objectives of `objs` in lock-step.
// On entry each objective has its near frontier of iteration 0 in
// near[0] / nearCount[0] and/or its far pile of epoch 0 in far[0] /
// farCount[0] / farMin[0]; `initThr` is the first threshold.
// ---------------------------------------------------------------------------

struct NearFarShared {
  u64 thr[MOSP_MAX_OBJ];
  u64 newThr[MOSP_MAX_OBJ];
  u64 delta[MOSP_MAX_OBJ];
  u64 farMin[MOSP_MAX_OBJ];   // this workgroup's far appends, this iteration
  unsigned cnt[MOSP_MAX_OBJ];
  int epoch[MOSP_MAX_OBJ];
  int mode[MOSP_MAX_OBJ];
  int waveBegin[MOSP_MAX_OBJ];  // waves of the grid working on objective k
  int waveEnd[MOSP_MAX_OBJ];
  int any;
};

// Split the grid's waves over the active objectives: one wave each, the rest
// in proportion to their item counts, leftovers to the largest.  Integer
// arithmetic only, so every workgroup derives the same partition.
__device__ void partition_waves(NearFarShared &sh, int nobj, unsigned waves) {
  u64 total = 0;
  int active = 0, largest = -1;
  for (int k = 0; k < nobj; ++k) {
    sh.waveBegin[k] = sh.waveEnd[k] = 0;
    if (sh.mode[k] != MODE_DONE) {
      total += sh.cnt[k];
      ++active;
      if (largest < 0 || sh.cnt[k] > sh.cnt[largest]) largest = k;
    }
  }
  if (active == 0) return;
  const u64 spare = waves > static_cast<unsigned>(active) ? waves - active : 0;
  unsigned next = 0, given = 0;
  int extra[MOSP_MAX_OBJ];
  for (int k = 0; k < nobj; ++k) {
    extra[k] = 0;
    if (sh.mode[k] != MODE_DONE) {
      extra[k] = static_cast<int>(spare * sh.cnt[k] / (total > 0 ? total : 1));
      given += 1 + extra[k];
    }
  }
  if (given < waves) extra[largest] += static_cast<int>(waves - given);
  for (int k = 0; k < nobj; ++k) {
    if (sh.mode[k] == MODE_DONE) continue;
    const unsigned w = 1 + extra[k];
    sh.waveBegin[k] = static_cast<int>(next < waves ? next : waves - 1);
    next += w;
    sh.waveEnd[k] = static_cast<int>(next < waves ? next : waves);
    if (sh.waveEnd[k] <= sh.waveBegin[k]) sh.waveEnd[k] = sh.waveBegin[k] + 1;
  }
}

// Relax the out-edges of frontier vertices of one objective: groups of
// 2^laneBits lanes share a vertex and stride over its edges.
__device__ void push_items(const ObjDesc &D, ObjCtl &c, NearFarShared &sh, int k,
                           unsigned it, int source, int gen, unsigned local,
                           unsigned span, int laneBits) {
  const unsigned cnt = sh.cnt[k];
  const int e = sh.epoch[k];
  const int bits = D.parentBits;
  const u64 noParent = D.noParent;
  const u64 thr = sh.thr[k];
  const u64 maxDist = D.maxDistance;
  u64 *packed = D.s.packed;
  const int *in = D.s.near[it & 1];
  int *outNear = D.s.near[(it + 1) & 1];
  int *farBuf = D.s.far[e & 1];
  unsigned *nextCnt = &c.nearCount[(it + 1) % 3];
  unsigned *farCnt = &c.farCount[e % 3];
  const unsigned lanes = 1u << laneBits;
  const unsigned sub = local & (lanes - 1);
  const unsigned groups = span >> laneBits;
  for (unsigned i = local >> laneBits; i < cnt; i += groups) {
    const int u = in[i];
    const u64 word = packed[u];
    if (word == PACKED_INF) {
      continue;
    }
    const u64 du = dist_of(word, bits);
    const int end = D.outRow[u + 1];
    for (int x = D.outRow[u] + static_cast<int>(sub); x < end; x += lanes) {
      const int w = D.outCol[x];
      if (w == source) {
        continue;  // the source is never relaxed
      }
      const u64 nd = du + static_cast<u64>(D.outW[x]);
      if (nd > maxDist) {
        continue;  // never shortest; would not fit the packed word
      }
      const u64 cand = pack(nd, u, bits, noParent);
      if (cand >= packed[w]) {
        continue;  // stale pre-check; the atomic below decides
      }
      const u64 prev = atomicMin(&packed[w], cand);
      if (nd < dist_of(prev, bits)) {
        if (nd < thr) {
          if (D.s.stamp[w] != gen && claim(D.s.stamp, w, gen)) {
            outNear[wave_append(nextCnt)] = w;
          }
        } else if (D.s.inFar[w] == 0 && atomicExch(&D.s.inFar[w], 1) == 0) {
          farBuf[wave_append(farCnt)] = w;
          atomicMin(&sh.farMin[k], nd);
        }
      }
    }
  }
}

// Raise the threshold past the far pile: move the entries below it to the
// next near frontier, keep the rest in the other far buffer.
__device__ void split_items(const ObjDesc &D, ObjCtl &c, NearFarShared &sh, int k,
                            unsigned it, int gen, unsigned local, unsigned span) {
  const unsigned cnt = sh.cnt[k];
  const int e = sh.epoch[k];
  const int bits = D.parentBits;
  const u64 thr = sh.newThr[k];
  const int *in = D.s.far[e & 1];
  int *far2 = D.s.far[(e + 1) & 1];
  int *outNear = D.s.near[(it + 1) & 1];
  unsigned *nextCnt = &c.nearCount[(it + 1) % 3];
  unsigned *far2Cnt = &c.farCount[(e + 1) % 3];
  for (unsigned i = local; i < cnt; i += span) {
    const int v = in[i];
    const u64 word = D.s.packed[v];
    const u64 d = word == PACKED_INF ? PACKED_INF : dist_of(word, bits);
    if (d < thr) {
      D.s.inFar[v] = 0;
      if (claim(D.s.stamp, v, gen)) {
        outNear[wave_append(nextCnt)] = v;
      }
    } else {
      far2[wave_append(far2Cnt)] = v;
      atomicMin(&sh.farMin[k], d);
    }
  }
}

__device__ int near_far(const ObjDesc *objs, int nobj, ObjCtl *ctls,
                        int source, int generation, u64 initThr,
                        const u64 *deltaOverride, int laneBits,
                        GridBarrier &grid, Threads t, NearFarShared &sh) {
  if (threadIdx.x < static_cast<unsigned>(nobj)) {
    const int k = threadIdx.x;
    sh.thr[k] = initThr;
    sh.epoch[k] = 0;
    sh.delta[k] = deltaOverride != nullptr ? load_relaxed(deltaOverride)
                                           : objs[k].delta;
  }
  const unsigned waves = t.count / 64;
  const unsigned wave = t.id / 64;
  for (unsigned it = 0;; ++it) {
    const unsigned old = (it + 2) % 3;
    if (threadIdx.x == 0) {
      sh.any = 0;
    }
    __syncthreads();
    // Snapshot: every workgroup reads the same stable counters, so every
    // thread of the grid takes the same decisions.
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      const unsigned nc = load_relaxed(&c.nearCount[it % 3]);
      const int e = sh.epoch[k];
      sh.farMin[k] = PACKED_INF;
      if (nc > 0) {
        sh.mode[k] = MODE_PUSH;
        sh.cnt[k] = nc;
        sh.any = 1;
      } else {
        const unsigned fc = load_relaxed(&c.farCount[e % 3]);
        if (fc > 0) {
          const u64 fm = load_relaxed(&c.farMin[e % 3]);
          const u64 base = fm > sh.thr[k] && fm != PACKED_INF ? fm : sh.thr[k];
          sh.mode[k] = MODE_SPLIT;
          sh.cnt[k] = fc;
          sh.newThr[k] = base + sh.delta[k];
          sh.any = 1;
        } else {
          sh.mode[k] = MODE_DONE;
        }
      }
    }
    __syncthreads();
    if (!sh.any) {
      break;
    }
    if (threadIdx.x == 0) {
      partition_waves(sh, nobj, waves);
    }
    __syncthreads();
    const int gen = ++generation;
    if (t.id == 0) {
      // Slots that nobody reads or appends to in this iteration.
      for (int k = 0; k < nobj; ++k) {
        ObjCtl &c = ctls[k];
        store_relaxed(&c.nearCount[old], 0u);
        if (sh.mode[k] == MODE_SPLIT) {
          const int e2 = (sh.epoch[k] + 2) % 3;
          store_relaxed(&c.farCount[e2], 0u);
          store_relaxed(&c.farMin[e2], PACKED_INF);
          c.epochs += 1;
        } else if (sh.mode[k] == MODE_PUSH) {
          c.iterations += 1;
          c.pushes += sh.cnt[k];
        }
      }
    }
    // This wave's objective (wave-uniform, so descriptor loads are scalar).
    int mine = -1;
    for (int k = 0; k < nobj; ++k) {
      if (static_cast<int>(wave) >= sh.waveBegin[k] &&
          static_cast<int>(wave) < sh.waveEnd[k] && sh.mode[k] != MODE_DONE) {
        mine = k;
      }
    }
    mine = __builtin_amdgcn_readfirstlane(mine);
    if (mine >= 0) {
      const ObjDesc &D = objs[mine];
      ObjCtl &c = ctls[mine];
      const unsigned first = static_cast<unsigned>(sh.waveBegin[mine]) * 64;
      const unsigned span = static_cast<unsigned>(sh.waveEnd[mine] - sh.waveBegin[mine]) * 64;
      if (sh.mode[mine] == MODE_PUSH) {
        push_items(D, c, sh, mine, it, source, gen, t.id - first, span, laneBits);
      } else {
        split_items(D, c, sh, mine, it, gen, t.id - first, span);
      }
    }
    __syncthreads();
    // Fold this workgroup's far minimum into the objective's slot.
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      const u64 fm = sh.farMin[k];
      if (fm != PACKED_INF) {
        const int slot = sh.mode[k] == MODE_SPLIT ? (sh.epoch[k] + 1) % 3 : sh.epoch[k] % 3;
        atomicMin(&ctls[k].farMin[slot], fm);
      }
    }
    grid.sync();
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      if (sh.mode[k] == MODE_SPLIT) {
        sh.thr[k] = sh.newThr[k];
        sh.epoch[k] += 1;
      }
    }
  }
  return generation;
}

// Words -> (distance, parent); distance-only words get INT_MAX parents for
// the recovery pass (src/sospUpdateGpu.cu:478-497).
__device__ void unpack(const ObjDesc &D, int n, int source, Threads t) {
  const int bits = D.parentBits;
  const bool clearFlags = D.nChanged > 0 && D.s.flag != nullptr;
  for (int v = t.id; v < n; v += t.count) {
    if (clearFlags) {
      D.s.flag[v] = 0;  // invalidation marks are left clean for the next update
    }
    const u64 word = D.s.packed[v];
    if (bits > 0) {
      if (word == PACKED_INF) {
        D.distOut[v] = DISTANCE_INF;
        D.parentOut[v] = -1;
      } else {
        D.distOut[v] = static_cast<int64_t>(dist_of(word, bits));
        const u64 p = word & D.noParent;
        D.parentOut[v] = p == D.noParent ? -1 : static_cast<int>(p);
      }
    } else {
      D.distOut[v] = word == PACKED_INF ? DISTANCE_INF
                                        : static_cast<int64_t>(word);
      D.parentOut[v] = word == PACKED_INF || v == source ? -1 : INT_MAX;
    }
  }
}

// Lowest-id parent over tight out-edges (src/sospUpdateGpu.cu:490-511).
__device__ void recover_parents(const ObjDesc &D, int n, int source,
                                Threads t) {
  for (int u = t.id; u < n; u += t.count) {
    const u64 du = D.s.packed[u];
    if (du == PACKED_INF) {
      continue;
    }
    for (int x = D.outRow[u]; x < D.outRow[u + 1]; ++x) {
      const int w = D.outCol[x];
      if (w != source && du + static_cast<u64>(D.outW[x]) == D.s.packed[w]) {
        atomicMin(&D.parentOut[w], u);
      }
    }
  }
}

// Source-only start of a from-scratch search (src/sospUpdateGpu.cu:222-231).
__device__ void scratch_init(const ObjDesc &D, ObjCtl &c, int n, int source,
                             Threads t) {
  for (int v = t.id; v < n; v += t.count) {
    D.s.packed[v] = v == source ? pack(0, -1, D.parentBits, D.noParent)
                                : PACKED_INF;
  }
  if (t.id == 0) {
    D.s.near[0][0] = source;
    store_relaxed(&c.nearCount[0], 1u);
  }
}

// ---------------------------------------------------------------------------
// Step 1 of the update for every objective (src/sospUpdateGpu.cu:233-344)
// ---------------------------------------------------------------------------

__device__ int update_step1(const LaunchArgs &a, GridBarrier &grid, Threads t,
                            int generation) {
  const int n = a.n;
  Ctl *ctl = a.ctl;
  // Pack the old trees; ancestors; roots of changed tree edges.
  bool anyChanged = false;
  for (int k = 0; k < a.nobj; ++k) {
    const ObjDesc &D = a.objs[k];
    anyChanged = anyChanged || D.nChanged > 0;
    unsigned overflow = 0;
    for (int v = t.id; v < n; v += t.count) {
      const int64_t d = D.distIn[v];
      const int p = D.parentIn[v];
      u64 word;
      if (d >= DISTANCE_INF / 2) {
        word = PACKED_INF;
      } else if (d < 0 || static_cast<u64>(d) > D.maxDistance) {
        overflow = 1;
        word = PACKED_INF;
      } else {
        word = pack(static_cast<u64>(d), p, D.parentBits, D.noParent);
      }
      D.s.packed[v] = word;
      D.s.ancestor[v] = p;
    }
    if (overflow) {
      atomicOr(&ctl->obj[k].overflow, 1u);
    }
    for (int i = t.id; i < D.nChanged; i += t.count) {
      const int v = D.chTo[i];
      if (D.parentIn[v] == D.chFrom[i]) {
        D.s.flag[v] = 1;
      }
    }
  }
  grid.sync();
  stamp(ctl, PHASE_PACK, t.id);

  // Subtree invalidation by pointer jumping.  Invariant for every x:
  // flag[x] == 1 implies a root among x and its ancestors; flag[x] == 0
  // implies no root on the tree path from x up to (excluding) ancestor[x].
  // Only x's own thread writes (ancestor[x], flag[x]) and every value another
  // thread can observe keeps the invariant, so the update is in place.
  if (anyChanged) {
    for (int round = 0; round < a.maxRounds; ++round) {
      if (t.id == 0) {
        store_relaxed(&ctl->jumpActive[(round + 1) % 3], 0u);
        ctl->rounds = round + 1;
      }
      int jumping = 0;
      for (int k = 0; k < a.nobj; ++k) {
        const ObjDesc &D = a.objs[k];
        if (D.nChanged == 0) {
          continue;
        }
        int *anc = D.s.ancestor;
        int *flag = D.s.flag;
        for (int v = t.id; v < n; v += t.count) {
          const int x = anc[v];
          if (x < 0 || flag[v]) {
            continue;
          }
          if (flag[x]) {
            flag[v] = 1;
            continue;
          }
          const int next = anc[x];
          anc[v] = next;
          jumping |= next >= 0 ? 1 : 0;
        }
      }
      if (__any(jumping) && __lane_id() == 0) {
        store_relaxed(&ctl->jumpActive[round % 3], 1u);
      }
      grid.sync();
      if (load_relaxed(&ctl->jumpActive[round % 3]) == 0) {
        break;
      }
    }
  }
  stamp(ctl, PHASE_JUMP, t.id);

  // Invalidate flagged vertices; candidates = invalidated + insert heads.
  // Each flagged vertex is visited by exactly one thread, so it is listed
  // with a plain stamp store; insert heads that are not flagged are
  // de-duplicated among themselves by a stamp claim.  Flags stay set until
  // the unpack sweep clears them (they are read-only here).
  const int genInv = ++generation;
  for (int k = 0; k < a.nobj; ++k) {
    const ObjDesc &D = a.objs[k];
    ObjCtl &c = ctl->obj[k];
    unsigned count = 0;
    if (D.nChanged > 0) {
      for (int v = t.id; v < n; v += t.count) {
        if (D.s.flag[v]) {
          D.s.packed[v] = PACKED_INF;
          D.s.stamp[v] = genInv;
          D.s.cand[wave_append(&c.candCount)] = v;
          ++count;
        }
      }
    }
    for (int i = t.id; i < D.nHeads; i += t.count) {
      const int v = D.heads[i];
      if (v != a.source && !D.s.flag[v] && claim(D.s.stamp, v, genInv)) {
        D.s.cand[wave_append(&c.candCount)] = v;
      }
    }
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
      count += __shfl_xor(count, offset);
    }
    if (__lane_id() == 0 && count > 0) {
      atomicAdd(&c.invalidated, count);
    }
  }
  grid.sync();
  stamp(ctl, PHASE_INVALIDATE, t.id);

  // Pull: each candidate takes the best (distance, parent) over its
  // in-neighbours; improved ones start in the far pile of epoch 0 (the first
  // near-far iteration splits it at min + delta).
  const int genPull = ++generation;
  for (int k = 0; k < a.nobj; ++k) {
    const ObjDesc &D = a.objs[k];
    ObjCtl &c = ctl->obj[k];
    const unsigned cnt = load_relaxed(&c.candCount);
    const int bits = D.parentBits;
    u64 localMin = PACKED_INF;
    for (unsigned i = t.id; i < cnt; i += t.count) {
      const int v = D.s.cand[i];
      const u64 current = D.s.packed[v];
      u64 best = current;
      for (int x = D.inRow[v]; x < D.inRow[v + 1]; ++x) {
        const int u = D.inCol[x];
        const u64 word = D.s.packed[u];
        if (word == PACKED_INF) {
          continue;
        }
        const u64 nd = dist_of(word, bits) + static_cast<u64>(D.inW[x]);
        if (nd <= D.maxDistance) {  // larger: never shortest, may not fit
          const u64 cand = pack(nd, u, bits, D.noParent);
          best = cand < best ? cand : best;
        }
      }
      if (best < current) {
        const u64 prev = atomicMin(&D.s.packed[v], best);
        const u64 bd = dist_of(best, bits);
        if (bd < dist_of(prev, bits) && claim(D.s.stamp, v, genPull)) {
          D.s.inFar[v] = 1;
          D.s.far[0][wave_append(&c.farCount[0])] = v;
          localMin = bd < localMin ? bd : localMin;
        }
      }
    }
    wave_min_into(localMin, &c.farMin[0]);
  }
  grid.sync();
  stamp(ctl, PHASE_PULL, t.id);
  return generation;
}

// ---------------------------------------------------------------------------
// Combined graph of cK trees (src/combinedGraphGpu.cu:55-121) + its CSR.
// ---------------------------------------------------------------------------

// Parent k of v if it is the first occurrence of its value among the K
// parents; its weight is base - sum of terms over the trees containing it.
__device__ __forceinline__ bool combined_edge(const LaunchArgs &a, int v, int k,
                                              int &p, int &weight) {
  const int n = a.n;
  const int *par = a.cParents;
  p = par[static_cast<size_t>(k) * n + v];
  if (p < 0) {
    return false;
  }
  for (int j = 0; j < k; ++j) {
    if (par[static_cast<size_t>(j) * n + v] == p) {
      return false;
    }
  }
  weight = a.cBase - a.cTerms[k];
  for (int j = k + 1; j < a.cK; ++j) {
    if (par[static_cast<size_t>(j) * n + v] == p) {
      weight -= a.cTerms[j];
    }
  }
  return true;
}

__device__ void combined_build(const LaunchArgs &a, GridBarrier &grid,
                               Threads t, int *sScan) {
  const int n = a.n;
  Ctl *ctl = a.ctl;
  // Clear the degree counters (n + 1).
  for (int i = t.id; i <= n; i += t.count) {
    a.cDegree[i] = 0;
  }
  grid.sync();
  // Count: out-degree of every parent; edge count and weight sum.
  u64 edges = 0, wsum = 0;
  for (int v = t.id; v < n; v += t.count) {
    if (v == a.source) {
      continue;
    }
    for (int k = 0; k < a.cK; ++k) {
      int p, weight;
      if (combined_edge(a, v, k, p, weight)) {
        atomicAdd(&a.cDegree[p], 1);
        edges += 1;
        wsum += static_cast<u64>(weight);
      }
    }
  }
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    edges += __shfl_xor(edges, offset);
    wsum += __shfl_xor(wsum, offset);
  }
  if (__lane_id() == 0 && edges > 0) {
    atomicAdd(&ctl->combEdges, edges);
    atomicAdd(&ctl->combWeightSum, wsum);
  }
  grid.sync();
  stamp(ctl, PHASE_COMB_COUNT, t.id);

  // Exclusive scan of the n + 1 degrees into row pointers: every workgroup
  // owns one contiguous chunk; chunk totals, then offsets, then local scans.
  const int total = n + 1;
  const int chunk = (total + gridDim.x - 1) / gridDim.x;
  const int lo = min(total, static_cast<int>(blockIdx.x) * chunk);
  const int hi = min(total, lo + chunk);
  const int per = (chunk + BLOCK - 1) / BLOCK;
  const int mlo = min(hi, lo + static_cast<int>(threadIdx.x) * per);
  const int mhi = min(hi, mlo + per);
  int mine = 0;
  for (int i = mlo; i < mhi; ++i) {
    mine += a.cDegree[i];
  }
  // Workgroup inclusive scan of `mine` (Hillis-Steele in LDS).
  sScan[threadIdx.x] = mine;
  __syncthreads();
  for (int offset = 1; offset < BLOCK; offset <<= 1) {
    const int add = threadIdx.x >= static_cast<unsigned>(offset)
                        ? sScan[threadIdx.x - offset]
                        : 0;
    __syncthreads();
    sScan[threadIdx.x] += add;
    __syncthreads();
  }
  const int chunkTotal = sScan[BLOCK - 1];
  const int myExclusive = sScan[threadIdx.x] - mine;
  if (threadIdx.x == 0) {
    a.cPartials[blockIdx.x] = chunkTotal;
  }
  grid.sync();
  int offset = 0;
  for (unsigned b = threadIdx.x; b < blockIdx.x; b += BLOCK) {
    offset += a.cPartials[b];
  }
  __syncthreads();
  sScan[threadIdx.x] = offset;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if (threadIdx.x < static_cast<unsigned>(s)) {
      sScan[threadIdx.x] += sScan[threadIdx.x + s];
    }
    __syncthreads();
  }
  int running = sScan[0] + myExclusive;
  for (int i = mlo; i < mhi; ++i) {
    const int d = a.cDegree[i];
    a.cRowPtr[i] = running;
    a.cDegree[i] = running;  // becomes the fill cursor
    running += d;
  }
  grid.sync();
  stamp(ctl, PHASE_COMB_SCAN, t.id);

  // Fill (row order is irrelevant: Step 3 breaks ties by the lowest id).
  for (int v = t.id; v < n; v += t.count) {
    if (v == a.source) {
      continue;
    }
    for (int k = 0; k < a.cK; ++k) {
      int p, weight;
      if (combined_edge(a, v, k, p, weight)) {
        const int pos = atomicAdd(&a.cDegree[p], 1);
        a.cColInd[pos] = v;
        a.cWeights[pos] = weight;
      }
    }
  }
  // Near-far width for Step 3: the caller's, or 32 * average weight /
  // average out-degree (src/sospUpdateGpu.cu:607-616), in double as there.
  if (t.id == 0) {
    long long delta = a.cDelta;
    if (delta <= 0) {
      const long long m = static_cast<long long>(load_relaxed(&ctl->combEdges));
      const long long ws = static_cast<long long>(load_relaxed(&ctl->combWeightSum));
      delta = 1;
      if (m > 0 && n > 0) {
        const double avgW = static_cast<double>(ws) / static_cast<double>(m);
        const double avgDeg = static_cast<double>(m) / static_cast<double>(n);
        const long long d = static_cast<long long>(32.0 * avgW / avgDeg);
        delta = d > 1 ? d : 1;
      }
    }
    store_relaxed(&ctl->combDelta, static_cast<u64>(delta));
  }
}

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------

__global__ void __launch_bounds__(BLOCK) sosp_persistent(LaunchArgs a) {
  __shared__ NearFarShared sh;
  __shared__ int sScan[BLOCK];
  GridBarrier grid{a.ctl, 0};
  const Threads t{blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK};
  const int n = a.n;
  if (t.id == 0) {
    a.ctl->clkStart = wall_clock64();
  }
  // Stamp generations persist in device memory across launches; restart
  // them (and clear the stamps) well before the counter could wrap.
  int generation = *a.genDev;
  if (generation > (1 << 30)) {
    for (int k = 0; k < a.nobj; ++k) {
      for (int v = t.id; v < n; v += t.count) {
        a.objs[k].s.stamp[v] = 0;
      }
    }
    if (a.cDesc != nullptr) {
      for (int v = t.id; v < n; v += t.count) {
        a.cDesc->s.stamp[v] = 0;
      }
    }
    generation = 0;
    grid.sync();
  }

  if (a.phases & PH_UPDATE) {
    generation = update_step1(a, grid, t, generation);
    generation = near_far(a.objs, a.nobj, a.ctl->obj, a.source, generation, 0,
                          nullptr, a.laneBits, grid, t, sh);
    stamp(a.ctl, PHASE_NEARFAR, t.id);
    for (int k = 0; k < a.nobj; ++k) {
      unpack(a.objs[k], n, a.source, t);
    }
    if (a.anyFallback) {
      grid.sync();
      for (int k = 0; k < a.nobj; ++k) {
        if (a.objs[k].parentBits == 0) {
          recover_parents(a.objs[k], n, a.source, t);
        }
      }
    }
    if (t.id == 0) {
      a.ctl->clkSospEnd = wall_clock64();
      a.ctl->clk[PHASE_UNPACK] = a.ctl->clkSospEnd;
    }
  }

  if (a.phases & PH_SCRATCH) {
    const ObjDesc &D = a.objs[0];
    scratch_init(D, a.ctl->obj[0], n, a.source, t);
    grid.sync();
    generation = near_far(a.objs, 1, a.ctl->obj, a.source, generation,
                          D.delta, nullptr, a.laneBits, grid, t, sh);
    unpack(D, n, a.source, t);
    if (D.parentBits == 0) {
      grid.sync();
      recover_parents(D, n, a.source, t);
    }
  }

  if (a.phases & PH_COMBINED) {
    if (a.phases & PH_UPDATE) {
      grid.sync();  // the K new trees are complete
    }
    const ObjDesc &C = *a.cDesc;
    ObjCtl &cc = a.ctl->obj[MOSP_MAX_OBJ];
    combined_build(a, grid, t, sScan);
    scratch_init(C, cc, n, a.source, t);
    grid.sync();
    stamp(a.ctl, PHASE_COMB_FILL, t.id);
    const u64 delta = load_relaxed(&a.ctl->combDelta);
    generation = near_far(a.cDesc, 1, &cc, a.source, generation, delta,
                          &a.ctl->combDelta, a.laneBits, grid, t, sh);
    stamp(a.ctl, PHASE_COMB_SEARCH, t.id);
    unpack(C, n, a.source, t);
    if (C.parentBits == 0) {
      grid.sync();
      recover_parents(C, n, a.source, t);
    }
  }

  if (t.id == 0) {
    a.ctl->generationOut = static_cast<unsigned>(generation);
    *a.genDev = generation;
    a.ctl->clkEnd = wall_clock64();
    a.ctl->clk[PHASE_COMB_UNPACK] = a.ctl->clkEnd;
    a.ctl->barriers = grid.epoch;
  }
}

__global__ void ctl_reset(Ctl *ctl) {
  unsigned *words = reinterpret_cast<unsigned *>(ctl);
  const unsigned count = sizeof(Ctl) / sizeof(unsigned);
  for (unsigned i = threadIdx.x; i < count; i += blockDim.x) {
    words[i] = 0;
  }
  __syncthreads();
  for (unsigned k = threadIdx.x; k <= static_cast<unsigned>(MOSP_MAX_OBJ);
       k += blockDim.x) {
    for (int s = 0; s < 3; ++s) {
      ctl->obj[k].farMin[s] = PACKED_INF;
    }
  }
}

}  // namespace

// Host-visible handles of the kernels (used by engine.hip).
const void *sosp_kernel_symbol() {
  return reinterpret_cast<const void *>(&sosp_persistent);
}

hipError_t launch_ctl_reset(Ctl *ctl, hipStream_t stream) {
  hipLaunchKernelGGL(ctl_reset, dim3(1), dim3(256), 0, stream, ctl);
  return hipGetLastError();
}

}  // namespace mosp
