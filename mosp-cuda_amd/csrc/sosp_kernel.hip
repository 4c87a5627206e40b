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

__device__ __forceinline__ unsigned load_acquire(const unsigned *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
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
// The counter only grows: barrier number e completes when it reaches
// e * gridDim.x, so it never needs a reset inside the launch.
struct GridBarrier {
  unsigned *counter;
  unsigned epoch;

  __device__ void sync() {
    __syncthreads();
    ++epoch;
    if (threadIdx.x == 0) {
      const unsigned target = epoch * gridDim.x;
      __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELEASE,
                             __HIP_MEMORY_SCOPE_AGENT);
      while (load_acquire(counter) < target) {
        __builtin_amdgcn_s_sleep(1);
      }
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
  unsigned cnt[MOSP_MAX_OBJ];
  int epoch[MOSP_MAX_OBJ];
  int mode[MOSP_MAX_OBJ];
  int any;
};

__device__ int near_far(const ObjDesc *objs, int nobj, ObjCtl *ctls,
                        int source, int generation, u64 initThr,
                        const u64 *deltaOverride, GridBarrier &grid,
                        Threads t, NearFarShared &sh) {
  if (threadIdx.x < static_cast<unsigned>(nobj)) {
    const int k = threadIdx.x;
    sh.thr[k] = initThr;
    sh.epoch[k] = 0;
    sh.delta[k] = deltaOverride != nullptr ? load_relaxed(deltaOverride)
                                           : objs[k].delta;
  }
  for (unsigned it = 0;; ++it) {
    const unsigned cur = it % 3, nxt = (it + 1) % 3, old = (it + 2) % 3;
    if (threadIdx.x == 0) {
      sh.any = 0;
    }
    __syncthreads();
    // Snapshot: every workgroup reads the same stable counters, so every
    // thread of the grid takes the same decisions.
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      const unsigned nc = load_relaxed(&c.nearCount[cur]);
      const int e = sh.epoch[k];
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
    for (int k = 0; k < nobj; ++k) {
      const int mode = sh.mode[k];
      if (mode == MODE_DONE) {
        continue;
      }
      const ObjDesc &D = objs[k];
      ObjCtl &c = ctls[k];
      const unsigned cnt = sh.cnt[k];
      const int e = sh.epoch[k];
      const int bits = D.parentBits;
      const u64 noParent = D.noParent;
      u64 *packed = D.s.packed;
      int *outNear = D.s.near[(it + 1) & 1];
      u64 localMin = PACKED_INF;
      if (mode == MODE_PUSH) {
        const u64 thr = sh.thr[k];
        const u64 maxDist = D.maxDistance;
        const int *in = D.s.near[it & 1];
        int *farBuf = D.s.far[e & 1];
        unsigned *farCnt = &c.farCount[e % 3];
        for (unsigned i = t.id; i < cnt; i += t.count) {
          const int u = in[i];
          const u64 word = packed[u];
          if (word == PACKED_INF) {
            continue;
          }
          const u64 du = dist_of(word, bits);
          const int begin = D.outRow[u], end = D.outRow[u + 1];
          for (int x = begin; x < end; ++x) {
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
                if (claim(D.s.stamp, w, gen)) {
                  outNear[wave_append(&c.nearCount[nxt])] = w;
                }
              } else if (atomicExch(&D.s.inFar[w], 1) == 0) {
                farBuf[wave_append(farCnt)] = w;
                localMin = nd < localMin ? nd : localMin;
              }
            }
          }
        }
        wave_min_into(localMin, &c.farMin[e % 3]);
      } else {  // MODE_SPLIT: raise the threshold past the far pile
        const u64 thr = sh.newThr[k];
        const int *in = D.s.far[e & 1];
        int *far2 = D.s.far[(e + 1) & 1];
        unsigned *far2Cnt = &c.farCount[(e + 1) % 3];
        for (unsigned i = t.id; i < cnt; i += t.count) {
          const int v = in[i];
          const u64 word = packed[v];
          const u64 d = word == PACKED_INF ? PACKED_INF : dist_of(word, bits);
          if (d < thr) {
            D.s.inFar[v] = 0;
            if (claim(D.s.stamp, v, gen)) {
              outNear[wave_append(&c.nearCount[nxt])] = v;
            }
          } else {
            far2[wave_append(far2Cnt)] = v;
            localMin = d < localMin ? d : localMin;
          }
        }
        wave_min_into(localMin, &c.farMin[(e + 1) % 3]);
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
  for (int v = t.id; v < n; v += t.count) {
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

  // Invalidate flagged vertices; candidates = invalidated + insert heads,
  // de-duplicated by a stamp claim.  Flags are left clean.
  const int genInv = ++generation;
  for (int k = 0; k < a.nobj; ++k) {
    const ObjDesc &D = a.objs[k];
    ObjCtl &c = ctl->obj[k];
    for (int v = t.id; v < n; v += t.count) {
      if (D.s.flag[v]) {
        D.s.flag[v] = 0;
        D.s.packed[v] = PACKED_INF;
        wave_append(&c.invalidated);
        if (claim(D.s.stamp, v, genInv)) {
          D.s.cand[wave_append(&c.candCount)] = v;
        }
      }
    }
    for (int i = t.id; i < D.nHeads; i += t.count) {
      const int v = D.heads[i];
      if (v != a.source && claim(D.s.stamp, v, genInv)) {
        D.s.cand[wave_append(&c.candCount)] = v;
      }
    }
  }
  grid.sync();

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
  GridBarrier grid{&a.ctl->barrier, 0};
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
                          nullptr, grid, t, sh);
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
    }
  }

  if (a.phases & PH_SCRATCH) {
    const ObjDesc &D = a.objs[0];
    scratch_init(D, a.ctl->obj[0], n, a.source, t);
    grid.sync();
    generation = near_far(a.objs, 1, a.ctl->obj, a.source, generation,
                          D.delta, nullptr, grid, t, sh);
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
    const u64 delta = load_relaxed(&a.ctl->combDelta);
    generation = near_far(a.cDesc, 1, &cc, a.source, generation, delta,
                          &a.ctl->combDelta, grid, t, sh);
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
