// sosp_kernel.hip -- the persistent cooperative gfx950 kernel of the update.
//
// Semantics (what each phase computes) follow the reference's engine,
// src/sospUpdateGpu.cu:211-519 (Step 1 + near-far Step 2 + unpack / parent
// recovery) and src/combinedGraphGpu.cu:55-121 (combined graph).  The
// schedule is this kernel's own (see engine.hpp):
//   * up to 32 objectives are updated in ONE launch, in lock-step, so every
//     grid barrier serves all of them; in each near-far iteration the
//     workgroups are split over the active objectives in proportion to their
//     frontier sizes;
//   * one grid barrier per near-far iteration: counters rotate over three
//     slots (iteration % 3 for near frontiers, epoch % 3 for far piles) so a
//     slot is cleared two iterations before it is reused;
//   * list appends are staged in LDS per workgroup (wave-aggregated LDS
//     reservations) and published with ONE global atomicAdd per workgroup
//     and list: same-address global atomics serialize (~10 ns each), so a
//     per-wave or per-lane counter would dominate an iteration;
//   * the far-pile minimum is folded in at append time (LDS atomicMin, one
//     global atomicMin per workgroup), so raising the threshold needs no
//     separate reduction pass and barrier;
//   * 2^laneBits lanes share a frontier vertex and stride over its edges,
//     shortening each lane's chain of dependent atomics;
//   * pointer jumping compacts its still-jumping set every round; the pull
//     pass sweeps the vertices instead of building a candidate list;
//   * the grid barrier is two-level (16 arrival counters, then one) with
//     agent-scope release/acquire fences and relaxed polling.
// No MFMA: this is integer graph work bound by memory latency, atomics and
// HBM bandwidth.
#include <hip/hip_runtime.h>

#include <climits>

#include "engine.hpp"

namespace mosp {
namespace {

enum : int { MODE_DONE = 0, MODE_PUSH = 1, MODE_SPLIT = 2 };

// Launch arguments and objective descriptors are written by the host before
// a launch and only read by the kernels, so they are read through the
// constant address space: uniform reads become scalar loads (results in
// SGPRs), and stores and atomics elsewhere cannot force them to be read
// again.  Through generic pointers every field would be a vector load
// repeated after each atomic, on the dependent path of every edge.
using Args = const __attribute__((address_space(4))) LaunchArgs;
using Desc = const __attribute__((address_space(4))) ObjDesc;

__device__ __forceinline__ Desc *desc(const ObjDesc *p) {
  return (Desc *)p;
}

// The same pointer, known to be uniform (SGPRs): arguments of calls arrive in
// VGPRs, and loads through them would be per-lane vector loads.
template <typename T>
__device__ __forceinline__ T *uniform(T *p) {
  const unsigned long long bits = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(bits));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(bits >> 32));
  return reinterpret_cast<T *>((static_cast<unsigned long long>(hi) << 32) | lo);
}

// uniform() of a pointer into device (global) memory, rebuilt in the global
// address space: atomics through it stay global atomics (through a rebuilt
// generic pointer they would be flat atomics with aperture checks).
template <typename T>
__device__ __forceinline__ T *uniform_global(T *p) {
  using G = __attribute__((address_space(1))) T;
  const unsigned long long bits = reinterpret_cast<unsigned long long>(p);
  const unsigned lo = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(bits));
  const unsigned hi = __builtin_amdgcn_readfirstlane(static_cast<unsigned>(bits >> 32));
  return (T *)reinterpret_cast<G *>((static_cast<unsigned long long>(hi) << 32) | lo);
}

constexpr unsigned QCAP = MOSP_QUEUE_CAP_MAX;  // LDS queue entries per list per workgroup
// Asynchronous near-far: local queue entries (a power of 2; 20 bytes each).
constexpr unsigned LQ_CAP = async_ring_entries(BLOCK);

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
__device__ __forceinline__ int load_relaxed(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_relaxed(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sub-phase clock sums for MOSP_SUBPROFILE builds: thread 0 of every
// workgroup adds its own intervals to LDS slots, written out to ctl->prof at
// the end of the launch (no contended atomics inside the timed loops).
#ifdef MOSP_SUBPROFILE
#define SUBPROF_MARK(var) const u64 var = wall_clock64()
#define SUBPROF_ADD(slot, from, to)                          \
  do {                                                       \
    if (threadIdx.x == 0) subprof_lds()[slot] += (to) - (from); \
  } while (0)
__device__ __forceinline__ u64 *subprof_lds() {
  __shared__ u64 slots[16];
  return slots;
}
__device__ __forceinline__ int *subprof_last() {
  __shared__ int last;
  return &last;
}
#else
#define SUBPROF_MARK(var) (void)0
#define SUBPROF_ADD(slot, from, to) (void)0
#endif

// Grid-wide barrier for a cooperative launch (all workgroups co-resident).
// Two arrival levels so that no counter sees more than ~gridDim/16
// same-address atomics per barrier: workgroup b arrives at sub[b % 16]; the
// last arrival of a group arrives at top.  Coherent loads of one address
// from many workgroups serialize like atomics (~20 ns each), so the release
// is replicated: the last arrival's wave writes the epoch -- and up to
// MAIL_WORDS words the caller names, read once after every arrival -- into
// one mailbox per group, and each workgroup polls only its group's mailbox
// (relaxed loads: an acquire per poll would invalidate the XCD's L2 on every
// poll) before one acquire fence.  Counters and epochs only grow, so nothing
// is reset inside a launch.  The last arrivals re-fence (acquire + release)
// before passing the arrival on, so every workgroup's writes happen-before
// every workgroup's later reads.
// The wait for the release is bounded (limit wall_clock64 ticks, and the
// launch's abort word, read every 64 polls): a workgroup that gives up sets
// the abort word, and sync() returns false to all of its threads (it stays
// false for every later call); each phase then returns at once, so the
// workgroup only ever works on the state of the last barrier it passed, as
// in a complete run.  The others give up at their next wait.  A launch
// whose workgroups are not all resident ends with its abort word set
// instead of spinning for ever.
__device__ __forceinline__ int *gave_up_lds() {
  __shared__ int flag;
  return &flag;
}

struct GridBarrier {
  Ctl *ctl;
  unsigned epoch;
  unsigned rank;  // workgroup rank within the group that synchronizes
  unsigned size;  // workgroups in that group
  unsigned *abortWord;
  u64 limit;      // wall_clock64 ticks a wait may last
  bool aborted = false;
#ifdef MOSP_SUBPROFILE
  bool wasLast = false;  // this workgroup released the last barrier
#endif

  // src: nsrc word addresses in a workgroup-uniform LDS array (set before
  // the call); dst: LDS array that receives their values as of the barrier.
  __device__ bool sync(unsigned *const *src = nullptr, int nsrc = 0,
                       unsigned *dst = nullptr) {
    if (aborted) {
      return false;
    }
    __syncthreads();
    // Locals: a store through a global pointer could alias *this, so every
    // member access after one would be reloaded.
    const unsigned e = ++epoch;
    Ctl *const c = ctl;
    unsigned *const abortWord = this->abortWord;
    const u64 limit = this->limit;
    const unsigned G = size;
    const unsigned g = rank % BARRIER_GROUPS;
    unsigned *const mail = c->mail;
    unsigned *const box = mail + g * MAIL_STRIDE;
    if (threadIdx.x < 64) {
      const unsigned lane = threadIdx.x;
      const unsigned groups = G < BARRIER_GROUPS ? G : BARRIER_GROUPS;
      int last = 0;
      if (lane == 0) {
        const unsigned groupSize =
            G / BARRIER_GROUPS + (g < G % BARRIER_GROUPS ? 1u : 0u);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        const unsigned sub = __hip_atomic_fetch_add(
            &c->barrierSub[g * LINE_WORDS], 1u, __ATOMIC_RELAXED,
            __HIP_MEMORY_SCOPE_AGENT);
        if (sub + 1 == e * groupSize) {
          __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
          const unsigned top = __hip_atomic_fetch_add(
              &c->barrierTop[0], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          if (top + 1 == e * groups) {
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
            last = 1;
          }
        }
      }
#ifdef MOSP_SUBPROFILE
      if (lane == 0) *subprof_last() = last;
#endif
      if (__shfl(last, 0)) {
        // Lane i copies word i % nsrc into mailbox i / nsrc (and further
        // mailboxes in steps of 64 lanes).
        if (nsrc > 0) {
          for (int i = static_cast<int>(lane); i < nsrc * static_cast<int>(groups);
               i += 64) {
            const int word = i % nsrc, h = i / nsrc;
            mail[h * MAIL_STRIDE + LINE_WORDS + word] = load_relaxed(src[word]);
          }
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        }
        if (lane < groups) {
          store_relaxed(&mail[lane * MAIL_STRIDE], e);
        }
      }
      if (lane == 0) {
        int gaveUp = 0;
        const u64 start = wall_clock64();
        for (unsigned polls = 1; load_relaxed(box) < e; ++polls) {
          if ((polls & 63u) == 0 &&
              (load_relaxed(abortWord) != 0 || wall_clock64() - start > limit)) {
            store_relaxed(abortWord, 1u);
            gaveUp = 1;
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        *gave_up_lds() = gaveUp;
      }
    }
    __syncthreads();
    if (*gave_up_lds() != 0) {
      aborted = true;
      return false;
    }
#ifdef MOSP_SUBPROFILE
    wasLast = *subprof_last() != 0;
#endif
    if (nsrc > 0) {
      for (int i = static_cast<int>(threadIdx.x); i < nsrc; i += BLOCK) {
        dst[i] = load_relaxed(&box[LINE_WORDS + i]);
      }
      __syncthreads();
    }
    return true;
  }
};

// One slot per active lane from a wave-uniform counter: one atomicAdd per
// wave (LDS or global), slot broadcast from the first active lane.
__device__ __forceinline__ unsigned wave_reserve(unsigned *counter) {
  const u64 active = __ballot(1);
  const unsigned rank = __popcll(active & ((1ull << __lane_id()) - 1ull));
  unsigned base = 0;
  if (rank == 0) {
    base = atomicAdd(counter, static_cast<unsigned>(__popcll(active)));
  }
  base = static_cast<unsigned>(
      __builtin_amdgcn_readfirstlane(static_cast<int>(base)));
  return base + rank;
}

// Lanes that own a vertex with more edges than this hand it to their whole
// wave: on power-law graphs a hub's tens of thousands of edges would
// otherwise be one lane's (or one lane group's) serial chain of dependent
// loads and atomics, which every other workgroup waits for at the next grid
// barrier.
constexpr int HEAVY_GROUP = 64;   // lane groups (push): more than one wave pass
constexpr unsigned HEAVY_LIST = 128;  // per-wave LDS list of heavy push vertices
// Push vertices of more edges than this go to the workgroup's list (drained
// by all its threads) rather than the wave's, when the team is the whole
// workgroup; those of more than the launch's hubEdges (HUB_GRID unless a
// test lowers it) to the objective's hub list.
constexpr int WAVE_EDGES = 512;
constexpr unsigned WG_LIST = 256;     // workgroup LDS list of heavy push vertices
static_assert(HUB_LIST <= BLOCK, "one thread per listed hub");
constexpr int HEAVY_THREAD = HUB_DEGREE;  // one lane per vertex (parent recovery)
// Combined-graph count in the hub variant: increments of a parent are first
// summed in a per-workgroup LDS table (open addressing, AGG_PROBES probes;
// misses go straight to the global counter), then added once per parent:
// a hub is the parent of ~10^5 children, whose same-address atomics would
// serialize.
constexpr int AGG_BITS = 10;
constexpr int AGG_SLOTS = 1 << AGG_BITS;
constexpr int AGG_PROBES = 8;
constexpr int STAGE_MAX = 8;    // objectives of a staged parent recovery
constexpr int PULL_BLOCK = 1024;   // pull: more in-edges -> a whole workgroup, after the sweep

__device__ __forceinline__ u64 shfl_u64(u64 x, int src) {
  const unsigned lo = __shfl(static_cast<unsigned>(x), src);
  const unsigned hi = __shfl(static_cast<unsigned>(x >> 32), src);
  return (static_cast<u64>(hi) << 32) | lo;
}

__device__ __forceinline__ u64 wave_min_u64(u64 x) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    const unsigned lo = __shfl_xor(static_cast<unsigned>(x), offset);
    const unsigned hi = __shfl_xor(static_cast<unsigned>(x >> 32), offset);
    const u64 y = (static_cast<u64>(hi) << 32) | lo;
    x = y < x ? y : x;
  }
  return x;
}

// Sum of x over the lanes below this one.
__device__ __forceinline__ u64 wave_exclusive_sum_u64(u64 x) {
  const int lane = static_cast<int>(__lane_id());
  u64 run = x;
#pragma unroll
  for (int offset = 1; offset < 64; offset <<= 1) {
    const u64 y = shfl_u64(run, lane >= offset ? lane - offset : lane);
    run += lane >= offset ? y : 0;
  }
  return run - x;
}

// Calls f(src) for every lane src of the wave with `mine` set, in lane order;
// wave-uniform control flow.
template <typename F>
__device__ __forceinline__ void wave_each(bool mine, F &&f) {
  u64 bits = __ballot(mine);
  while (bits != 0) {
    const int src = __ffsll(static_cast<long long>(bits)) - 1;
    bits &= bits - 1;
    f(src);
  }
}

// Largest u in [lo, hi] with rowPtr[u] <= e: the row holding edge e when
// rowPtr[lo] <= e < rowPtr[hi + 1].
__device__ __forceinline__ int row_of(const int *rowPtr, int lo, int hi, long long e) {
  while (lo < hi) {
    const int mid = lo + (hi - lo + 1) / 2;
    if (rowPtr[mid] <= e) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// List appends staged in LDS.  The buffer is split into equal slices, one
// per team of waves working on one objective (slice 0 = the whole buffer
// when the whole workgroup is one team); lanes reserve slots in their slice
// (one LDS atomic per wave) and a slice's entries are published with ONE
// global atomicAdd.  A full slice sends entries straight to the global list.
constexpr unsigned WAVES = BLOCK / 64;
// Every per-wave LDS array below is dimensioned by WAVES and indexed by
// threadIdx.x / 64 (or a team rank below WAVES): a workgroup of BLOCK
// threads has exactly WAVES full waves.
static_assert(BLOCK % 64 == 0 && WAVES * 64 == BLOCK && WAVES <= 16, "whole wave64s, at most 16");
// HIP's second launch bound is the minimum number of waves per SIMD: 4 SIMDs
// per CU share the workgroups' waves.
constexpr unsigned WAVES_PER_SIMD = MOSP_MIN_BLOCKS * WAVES / 4;

struct BlockQueue {
  int buf[QCAP];
  unsigned count[WAVES];  // entries reserved per slice
  unsigned base[WAVES];   // list offset of a slice being flushed
  unsigned keep;          // entries of a whole-buffer flush that fit the list
};

// Global lists hold listCap entries.  Every append still counts, so a list
// that would run past its end has a count above listCap and loses the
// entries beyond it: readers take such a count to mean "every vertex".  The
// two stores name their address spaces (LDS queue, global list): otherwise
// the compiler merges them into one flat store through a pointer that may
// be either.
__device__ __forceinline__ void slice_push(BlockQueue &q, unsigned slice, unsigned cap, int v,
                                           int *list, unsigned *listCount, unsigned listCap) {
  const unsigned slot = wave_reserve(&q.count[slice]);
  if (slot < cap) {
    *(__attribute__((address_space(3))) int *)(q.buf + slice * cap + slot) = v;
  } else {
    const unsigned at = wave_reserve(listCount);
    if (at < listCap) {
      *(__attribute__((address_space(1))) int *)(list + at) = v;
    }
  }
}

__device__ __forceinline__ void queue_push(BlockQueue &q, int v, int *list,
                                           unsigned *listCount, unsigned listCap) {
  slice_push(q, 0, QCAP, v, list, listCount, listCap);
}

// Workgroup-uniform: publish the whole-buffer queue (its first `cap`
// entries hold appends; later ones went straight to the list) and empty it.
// Thread 0 alone reads the count and empties the queue before the barrier;
// the others read the published base and length after it.
__device__ void queue_flush(BlockQueue &q, int *list, unsigned *listCount, unsigned listCap,
                            unsigned cap = QCAP) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned count = q.count[0] < cap ? q.count[0] : cap;
    const unsigned base = count > 0 ? atomicAdd(listCount, count) : 0;
    q.base[0] = base;
    q.keep = base < listCap ? min(count, listCap - base) : 0;  // entries that fit
    q.count[0] = 0;
  }
  __syncthreads();
  const unsigned keep = q.keep;
  for (unsigned i = threadIdx.x; i < keep; i += BLOCK) {
    list[q.base[0] + i] = q.buf[i];
  }
  __syncthreads();
}

// Workgroup-uniform: publish slices [0, teams) of `cap` entries to their
// lists (one global atomicAdd per slice) and empty them.
__device__ void flush_slices(BlockQueue &q, unsigned teams, unsigned cap,
                             int *const *lists, unsigned *const *counts, unsigned listCap) {
  __syncthreads();
  if (threadIdx.x < teams) {
    const unsigned s = threadIdx.x;
    const unsigned n = q.count[s] < cap ? q.count[s] : cap;
    const unsigned base = n > 0 ? atomicAdd(counts[s], n) : 0;
    q.base[s] = base;
    q.count[s] = base < listCap ? min(n, listCap - base) : 0;  // entries that fit
  }
  __syncthreads();
  for (unsigned s = 0; s < teams; ++s) {
    const unsigned n = q.count[s];
    int *list = lists[s];
    for (unsigned i = threadIdx.x; i < n; i += BLOCK) {
      list[q.base[s] + i] = q.buf[s * cap + i];
    }
  }
  __syncthreads();
  if (threadIdx.x < teams) {
    q.count[threadIdx.x] = 0;
  }
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

// Edge x of an objective as {column, weight}: one 8-byte record when the
// graph holds records (device graphs, the combined graph), else the
// separate column and weight arrays.  The choice is wave-uniform.
__device__ __forceinline__ int2 out_edge(Desc &D, long long x) {
  const int2 *rec = D.outEdge;
  if (rec != nullptr) {
    return rec[x];
  }
  return make_int2(D.outCol[x], D.outW[x]);
}

__device__ __forceinline__ int2 in_edge(Desc &D, long long x) {
  const int2 *rec = D.inEdge;
  if (rec != nullptr) {
    return rec[x];
  }
  return make_int2(D.inCol[x], D.inW[x]);
}

// Stamp the end of a phase (called right after a grid barrier).
__device__ __forceinline__ void stamp(Ctl *ctl, int phase, unsigned tid) {
  if (tid == 0) {
    ctl->clk[phase] = wall_clock64();
  }
}

// Workgroup sum of one value per thread (uses `red`, BLOCK entries).
__device__ unsigned block_sum(unsigned value, unsigned *red) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    value += __shfl_xor(value, offset);
  }
  __syncthreads();
  if (__lane_id() == 0) {
    red[threadIdx.x / 64] = value;
  }
  __syncthreads();
  unsigned total = 0;
  for (unsigned w = 0; w < BLOCK / 64; ++w) {
    total += red[w];
  }
  __syncthreads();
  return total;
}

// Workgroup minimum of one value per thread (uses `red`, WAVES entries).
__device__ u64 block_min_u64(u64 value, u64 *red) {
  value = wave_min_u64(value);
  __syncthreads();
  if (__lane_id() == 0) {
    red[threadIdx.x / 64] = value;
  }
  __syncthreads();
  u64 best = PACKED_INF;
  for (unsigned w = 0; w < WAVES; ++w) {
    best = red[w] < best ? red[w] : best;
  }
  __syncthreads();
  return best;
}

struct Threads {
  unsigned id;      // thread index within the group of workgroups
  unsigned count;   // threads in the group
  unsigned block;   // workgroup rank within the group
  unsigned blocks;  // workgroups in the group
};

// Dense sweep over the vertices [0, n) of the launch's threads, four
// vertices per thread in flight: load(v) of the four is issued before any
// use(v, loaded), so a sweep waits for one round trip per four vertices
// instead of one per vertex (stores in use() would otherwise keep the next
// vertex's loads behind them).
template <typename L, typename U>
__device__ __forceinline__ void sweep4(int n, Threads t, L &&load, U &&use) {
  const long long stride = t.count;
  for (long long v0 = t.id; v0 < n; v0 += 4 * stride) {
    decltype(load(0)) x[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long v = v0 + j * stride;
      if (v < n) {
        x[j] = load(static_cast<int>(v));
      }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const long long v = v0 + j * stride;
      if (v < n) {
        use(static_cast<int>(v), x[j]);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Workgroup-shared state
// ---------------------------------------------------------------------------

struct Shared {
  // near-far per objective
  u64 thr[MOSP_MAX_OBJ];
  u64 newThr[MOSP_MAX_OBJ];
  u64 delta[MOSP_MAX_OBJ];
  unsigned cnt[MOSP_MAX_OBJ];
  int epoch[MOSP_MAX_OBJ];
  int mode[MOSP_MAX_OBJ];
  int blockBegin[MOSP_MAX_OBJ];  // workgroups working on objective k
  int blockEnd[MOSP_MAX_OBJ];
  int any;
  u64 farMin;                    // this workgroup's far appends (pull)
  int mine[MOSP_MAX_OBJ];        // objectives this workgroup works on now
  int nmine;
  u64 sliceFarMin[WAVES];        // per team: min distance appended to far
  int *sliceList[2][WAVES];      // per team: near / far target lists
  unsigned *sliceCount[2][WAVES];
  int sliceObj[WAVES];
  int sliceFarSlot[WAVES];
  unsigned *src[MAIL_WORDS];     // counters the next grid barrier publishes
  unsigned snap[MAIL_WORDS];     // their values after it
  unsigned statIter[MOSP_MAX_OBJ];
  unsigned statEpoch[MOSP_MAX_OBJ];
  u64 statPush[MOSP_MAX_OBJ];
  unsigned red[BLOCK / 64];
  int scan[BLOCK];
  union {
    int heavy[WAVES][HEAVY_LIST];  // push: per-wave list of heavy vertices
    struct {                       // push: this iteration's hubs (before the lists fill)
      u64 end[HUB_LIST];           // end of each hub's edges in their concatenation
      int u[HUB_LIST];
      int begin[HUB_LIST];         // its first edge
    } hub;
  };
  unsigned hubs[MOSP_MAX_OBJ];   // hubs listed for this iteration (hub variant)
  int hubEdges;                  // push: more out-edges -> the hub list (hub variant)
  unsigned queueCap;             // push: entries of the LDS queues used (<= QCAP)
  unsigned wgCount;              // push: entries of wgList (hub variant)
  int wgList[WG_LIST];
  int span[2];                   // edge-parallel passes: rows of the first/last edge
  int aggKey[AGG_SLOTS];         // combined count (hub variant): parent ids ...
  int aggCount[AGG_SLOTS];       // ... and their increments in this workgroup
  u64 redMin[WAVES];             // workgroup minimum
  union {
    BlockQueue near;
    // Asynchronous near-far: the local queue, a ring of vertices and the
    // words they were listed with.
    struct {
      int v[LQ_CAP];
      int b[LQ_CAP];   // first out-edge, or -1: read the row pointers
      int e[LQ_CAP];   // end of the out-edges
      u64 w[LQ_CAP];
    } lq;
  };
  BlockQueue far;
  // Asynchronous near-far: the local queue is the ring lq.
  struct {
    unsigned lqHead, lqTail;     // entries [lqHead, lqTail) are listed (ring indices mod LQ_CAP)
    int cur;                     // objective of the local queue and of the far queue
    int step;                    // what the workgroup does next (STEP_*)
    int claimed;                 // outcome of a claim: STEP_POOL or STEP_IDLE
    unsigned count;              // entries of the current batch
    int seg;                     // pool segment of a pool batch ...
    unsigned base;               // ... and its first slot
    unsigned sent, recv;         // entries this workgroup appended / processed (status word)
    int seedIdx;                 // index into mine[] of the objective being seeded
    unsigned seedPos, seedEnd;   // the rest of this workgroup's share of its seeds
    int exit;                    // the phase is over
    // Set by thread 0 before the barrier that ends a batch, read by every
    // thread after it: the next batch may already change exit and the far
    // queue while slower threads still decide on the last one.
    int leave;                   // exit was 2: leave the loop
    int flushFar;                // publish the far queue now
    int lostNear;                // a near entry found no room in this batch
    unsigned polls;              // consecutive idle polls
    unsigned batches;
    u64 farMin;                  // smallest distance in the far queue since its last flush
    unsigned spillBase, spillKeep;
    int spillSeg;                // pool segment of the spill
    int obj;                     // objective of the entries a pool batch took
    int hungry;                  // this workgroup counts in its XCD's hungry word
  } as;
};

// Split the workgroups over the active objectives: one each, the rest in
// proportion to their item counts, leftovers to the largest.  Integer
// arithmetic only, so every workgroup derives the same partition.  With more
// active objectives than workgroups, objective k goes to workgroup k % G.
// A listed hub (hub variant) weighs as much as HUB_WEIGHT items.
constexpr u64 HUB_WEIGHT = HUB_GRID / 16;

__device__ void partition_blocks(Shared &sh, int nobj, unsigned G) {
  u64 total = 0;
  unsigned active = 0;
  int largest = -1;
  auto weight = [&](int k) { return sh.cnt[k] + HUB_WEIGHT * sh.hubs[k]; };
  for (int k = 0; k < nobj; ++k) {
    sh.blockBegin[k] = sh.blockEnd[k] = 0;
    if (sh.mode[k] != MODE_DONE) {
      total += weight(k);
      ++active;
      if (largest < 0 || weight(k) > weight(largest)) largest = k;
    }
  }
  if (active == 0) return;
  if (active > G) {
    for (int k = 0; k < nobj; ++k) {
      if (sh.mode[k] != MODE_DONE) {
        sh.blockBegin[k] = k % G;
        sh.blockEnd[k] = k % G + 1;
      }
    }
    return;
  }
  const u64 spare = G - active;
  unsigned given = 0;
  for (int k = 0; k < nobj; ++k) {
    if (sh.mode[k] != MODE_DONE) {
      sh.blockEnd[k] = 1 + static_cast<int>(spare * weight(k) / (total > 0 ? total : 1));
      given += sh.blockEnd[k];
    }
  }
  sh.blockEnd[largest] += static_cast<int>(G - given);
  int next = 0;
  for (int k = 0; k < nobj; ++k) {
    if (sh.mode[k] == MODE_DONE) continue;
    const int size = sh.blockEnd[k];
    sh.blockBegin[k] = next;
    next += size;
    sh.blockEnd[k] = next;
  }
}

// ---------------------------------------------------------------------------
// Near-far propagation (Step 2), all objectives of `objs` in lock-step.
// On entry each objective has its near frontier of iteration 0 in
// near[0] / nearCount[0] and/or its far pile of epoch 0 in far[0] /
// farCount[0] / farMin[0]; `initThr` is the first threshold.
// ---------------------------------------------------------------------------

// A team: a contiguous run of waves of one workgroup working on one
// objective in a near-far iteration, with its own LDS queue slices.  When a
// workgroup serves several objectives (small groups, many objectives) its
// waves split into teams that run them concurrently.
struct Team {
  unsigned slice;  // LDS queue slice of the team
  unsigned tid;    // thread index within the team (teams start on waves)
  unsigned size;   // threads in the team
  unsigned cap;    // entries per slice
  bool whole;      // the team is the whole workgroup: it may flush mid-pass
};

// Relax the out-edges of the frontier vertices of objective k: groups of
// 2^laneBits lanes share a vertex and stride over its edges.
// SPLIT: the iteration raises the threshold past the far pile instead: its
// entries below the new threshold move to the next near frontier, the rest
// to the next epoch's pile.  In the kernels with local hops, the hop rounds
// then relax the moved vertices at once, under the new threshold, instead of
// leaving them all to the next iteration.
template <bool HUBS, int LOCAL, bool SPLIT = false>
__device__ void push_team(Desc &D, ObjCtl &c, Shared &sh, int k,
                          unsigned it, int source, int gen, unsigned blockRank,
                          unsigned blocks, int laneBits, int localHops, unsigned n,
                          unsigned listCap, const Team &tm) {
  // A count above the lists' capacity means entries were lost: the frontier
  // is then every vertex (relaxing a vertex again is always valid); a split's
  // pile is then read from the far marks of every vertex, which name the
  // same set.
  const bool all = sh.cnt[k] > listCap;
  const unsigned cnt = all ? n : sh.cnt[k];
  const int e = sh.epoch[k] + (SPLIT ? 1 : 0);  // a split works for the next epoch
  const int bits = D.parentBits;
  const u64 noParent = D.noParent;
  const u64 thr = SPLIT ? sh.newThr[k] : sh.thr[k];
  const u64 maxDist = D.maxDistance;
  const bool optimistic = D.optimistic != 0;
  // Locals, read once; a team is whole waves, so they are wave-uniform and
  // live in SGPRs (k, hence D, comes from LDS: the compiler cannot tell).
  // All are device memory: global, not flat, accesses.
  u64 *const packed = uniform_global(D.s.packed);
  int *const stampArr = uniform_global(D.s.stamp);
  int *const inFar = uniform_global(D.s.inFar);
  const int *const outRow = uniform_global(D.outRow);
  const int *const outCol = uniform_global(D.outCol);
  const int *const outW = uniform_global(D.outW);
  const int2 *const outRec = uniform_global(D.outEdge);
  const int *in = uniform_global(SPLIT ? D.s.far[(e + 1) & 1] : D.s.near[it & 1]);
  int *outNear = uniform_global(D.s.near[(it + 1) & 1]);
  int *farBuf = uniform_global(D.s.far[e & 1]);
  unsigned *nextCnt = &c.nearCount[(it + 1) % 3];
  unsigned *farCnt = &c.farCount[e % 3];
  const unsigned lanes = 1u << laneBits;
  const unsigned sub = tm.tid & (lanes - 1);
  const unsigned perPass = tm.size >> laneBits;
  const int lane = static_cast<int>(__lane_id());
  // Relax edge x of u, whose word is `word`.  The edge back to u's parent
  // (the vertex whose relaxation gave u this word) is skipped: it can never
  // improve the parent (weights are at least 1).
  const u64 parentMask = bits != 0 ? (1ull << bits) - 1ull : 0ull;
  auto relax = [&](int u, u64 word, int x) {
    const u64 du = dist_of(word, bits);
    const int pu = bits != 0 ? static_cast<int>(word & parentMask) : -1;
    // One record load when the graph holds {column, weight} records.
    int w, weight;
    if (outRec != nullptr) {
      const int2 r = outRec[x];
      w = r.x;
      weight = r.y;
    } else {
      w = outCol[x];
      weight = outW[x];
    }
    if (w == source || w == pu) {
      return;  // the source is never relaxed
    }
    const u64 nd = du + static_cast<u64>(weight);
    if (nd > maxDist) {
      // Never shortest, unless the packing is optimistic (then the launch
      // is redone with a packing that holds it); would not fit the word.
      if (optimistic) {
        store_relaxed(&c.dropped, 1u);
      }
      return;
    }
    const u64 cand = pack(nd, u, bits, noParent);
    // Every improvement of w below the threshold lists w for the next
    // iteration.  In the light variant (at most HUB_DEGREE edges a vertex)
    // that is all: a vertex improved twice in an iteration is listed twice,
    // and its second entry relaxes a few edges again (about 3 % more entries
    // on a road grid), which costs less than reading a claim word -- one
    // more random line -- on every relaxation (4.17 -> 3.75 ms per update).
    // The hub variant claims near appends with w's stamp, and every variant
    // claims far appends with w's far mark (the pile lives for an epoch).
    // The claim word is read before the atomic so both round trips overlap.
    // Claims need no atomics: two lanes may both append w (a duplicate entry
    // only repeats work; racing lanes, and other XCDs, which see this XCD's
    // stores only after the next barrier, can make many), but an improvement
    // is never missed: the stamp only ever holds this iteration's generation
    // once w is appended in it (the entry reads w's word when processed,
    // after this iteration's atomics), and inFar is only set during push
    // iterations of this epoch.  Reading the claim word early can only see
    // an older value, i.e. add a duplicate.  Duplicates can run a list past
    // its n entries (a small graph that local hops cover in one iteration
    // does): the lost entries turn the next iteration into a sweep of every
    // vertex.
    const bool near = nd < thr;
    const int claim = near ? (HUBS ? stampArr[w] : 0) : inFar[w];
    const u64 prev = atomicMin(&packed[w], cand);
    if (nd < dist_of(prev, bits)) {
      if (near) {
        if (!HUBS || claim != gen) {
          if (HUBS) {
            stampArr[w] = gen;
          }
          slice_push(sh.near, tm.slice, tm.cap, w, outNear, nextCnt, listCap);
        }
      } else if (claim == 0) {
        inFar[w] = 1;
        slice_push(sh.far, tm.slice, tm.cap, w, farBuf, farCnt, listCap);
        atomicMin(&sh.sliceFarMin[tm.slice], nd);
      }
    }
  };
  // Contiguous, even shares: a short frontier still spreads over every
  // workgroup of the objective (random loads cost per-CU address cycles).
  const unsigned lo = static_cast<unsigned>(static_cast<u64>(cnt) * blockRank / blocks);
  const unsigned hi = static_cast<unsigned>(static_cast<u64>(cnt) * (blockRank + 1) / blocks);
  if constexpr (SPLIT) {
    // One thread per pile entry; `in` is the old epoch's pile, farBuf the new one's.
    for (unsigned base = lo; base < hi; base += tm.size) {
      const unsigned i = base + tm.tid;
      const int v = i < hi ? (all ? static_cast<int>(i) : in[i]) : -1;
      if (v >= 0 && (!all || inFar[v] != 0)) {
        const u64 word = packed[v];
        const u64 d = word == PACKED_INF ? PACKED_INF : dist_of(word, bits);
        if (d < thr) {
          inFar[v] = 0;
          if (stampArr[v] != gen) {  // duplicates of v in the pile: append once
            stampArr[v] = gen;
            slice_push(sh.near, tm.slice, tm.cap, v, outNear, nextCnt, listCap);
          }
        } else {
          slice_push(sh.far, tm.slice, tm.cap, v, farBuf, farCnt, listCap);
          atomicMin(&sh.sliceFarMin[tm.slice], d);
        }
      }
      if (tm.whole) {
        __syncthreads();
        if (sh.near.count[0] + BLOCK > tm.cap) queue_flush(sh.near, outNear, nextCnt, listCap, tm.cap);
        if (sh.far.count[0] + BLOCK > tm.cap) queue_flush(sh.far, farBuf, farCnt, listCap, tm.cap);
      }
    }
  } else if constexpr (!HUBS) {
    for (unsigned base = lo; base < hi; base += perPass) {
      const unsigned i = base + (tm.tid >> laneBits);
      if (i < hi) {
        const int u = all ? static_cast<int>(i) : in[i];
        // Independent loads, issued together.
        const u64 word = packed[u];
        const int begin = outRow[u];
        const int end = outRow[u + 1];
        if (word != PACKED_INF) {
          for (int x = begin + static_cast<int>(sub); x < end; x += static_cast<int>(lanes)) {
            relax(u, word, x);
          }
        }
      }
      if (tm.whole) {
        __syncthreads();
        if (sh.near.count[0] > tm.cap / 2) queue_flush(sh.near, outNear, nextCnt, listCap, tm.cap);
        if (sh.far.count[0] > tm.cap / 2) queue_flush(sh.far, farBuf, farCnt, listCap, tm.cap);
      }
    }
  } else {
    // Out-degree tiers.  Up to HEAVY_GROUP edges: the vertex's lane group, as
    // above.  Up to WAVE_EDGES (up to hubEdges when the team is not the whole
    // workgroup): this wave's LDS list, whose vertices the whole wave relaxes
    // one by one when the list could overflow and after the last pass.  Up
    // to hubEdges: the workgroup's LDS list, relaxed by all its threads, one
    // vertex after the other, when it is half full and after the last pass.
    // More: the objective's hub list, relaxed in the next iteration by every
    // workgroup of the objective, each over an even share of the listed
    // hubs' edges (the next iteration is then a push iteration, under the
    // same threshold).  A full list hands a vertex down a tier.  Listed
    // vertices' distances are re-read when they are relaxed: a newer one is
    // as valid a source.
    // First the hubs listed for this iteration.
    const unsigned hc = sh.hubs[k] < static_cast<unsigned>(HUB_LIST) ? sh.hubs[k] : HUB_LIST;
    if (hc > 0) {
      const int *const hubIn = uniform_global(D.s.hubs + (it & 1) * HUB_LIST);
      if (tm.whole) {
        // The listed hubs' edges as one range: each hub's end in it, in LDS.
        if (threadIdx.x < hc) {
          const int u = hubIn[threadIdx.x];
          const u64 word = packed[u];
          const int b = outRow[u];
          const int e2 = outRow[u + 1];
          sh.hub.u[threadIdx.x] = u;
          sh.hub.begin[threadIdx.x] = b;
          sh.hub.end[threadIdx.x] = word != PACKED_INF ? static_cast<u64>(e2 - b) : 0;
        }
        __syncthreads();
        if (threadIdx.x < 64) {
          constexpr unsigned per = HUB_LIST / 64;  // entries per lane of wave 0
          u64 own = 0;
          for (unsigned j = 0; j < per; ++j) {
            const unsigned at = static_cast<unsigned>(lane) * per + j;
            own += at < hc ? sh.hub.end[at] : 0;
          }
          u64 run = wave_exclusive_sum_u64(own);
          for (unsigned j = 0; j < per; ++j) {
            const unsigned at = static_cast<unsigned>(lane) * per + j;
            if (at < hc) {
              run += sh.hub.end[at];
              sh.hub.end[at] = run;
            }
          }
        }
        __syncthreads();
        const u64 total = sh.hub.end[hc - 1];
        const u64 from = total * blockRank / blocks;
        const u64 to = total * (blockRank + 1) / blocks;
        for (u64 base = from; base < to; base += BLOCK) {
          const u64 g = base + threadIdx.x;
          if (g < to) {
            // The first hub whose end is past g.
            unsigned j = 0;
#pragma unroll
            for (unsigned step = HUB_LIST / 2; step > 0; step >>= 1) {
              if (j + step < hc && sh.hub.end[j + step - 1] <= g) {
                j += step;
              }
            }
            const u64 start = j > 0 ? sh.hub.end[j - 1] : 0;
            const int u = sh.hub.u[j];
            relax(u, packed[u], sh.hub.begin[j] + static_cast<int>(g - start));
          }
          __syncthreads();
          if (sh.near.count[0] > tm.cap / 2) queue_flush(sh.near, outNear, nextCnt, listCap, tm.cap);
          if (sh.far.count[0] > tm.cap / 2) queue_flush(sh.far, farBuf, farCnt, listCap, tm.cap);
        }
        __syncthreads();  // the wave lists overlay the hub table
      } else {
        for (unsigned j = 0; j < hc; ++j) {
          const int u = hubIn[j];
          const u64 word = packed[u];
          const int b = outRow[u];
          const int len = word != PACKED_INF ? outRow[u + 1] - b : 0;
          const int x0 = b + static_cast<int>(static_cast<u64>(len) * blockRank / blocks);
          const int x1 = b + static_cast<int>(static_cast<u64>(len) * (blockRank + 1) / blocks);
          for (int x = x0 + static_cast<int>(tm.tid); x < x1; x += static_cast<int>(tm.size)) {
            relax(u, word, x);
          }
        }
      }
    }
    // Workgroup-uniform, after a __syncthreads: relax the workgroup list.
    auto drainGroup = [&]() {
      const unsigned m = sh.wgCount < WG_LIST ? sh.wgCount : WG_LIST;
      for (unsigned j = 0; j < m; ++j) {
        const int u = sh.wgList[j];
        const u64 word = packed[u];
        const int b = outRow[u];
        const int e2 = word != PACKED_INF ? outRow[u + 1] : b;
        for (int x = b + static_cast<int>(threadIdx.x); x < e2; x += BLOCK) {
          relax(u, word, x);
        }
        __syncthreads();
        if (sh.near.count[0] > tm.cap / 2) queue_flush(sh.near, outNear, nextCnt, listCap, tm.cap);
        if (sh.far.count[0] > tm.cap / 2) queue_flush(sh.far, farBuf, farCnt, listCap, tm.cap);
      }
      if (threadIdx.x == 0) {
        sh.wgCount = 0;
      }
      __syncthreads();
    };
    const int hubEdges = __builtin_amdgcn_readfirstlane(sh.hubEdges);
    const int groupEdges = hubEdges < WAVE_EDGES ? hubEdges : WAVE_EDGES;
    int *const hubOut = uniform_global(D.s.hubs + ((it + 1) & 1) * HUB_LIST);
    unsigned *const hubCnt = &c.hubCount[(it + 1) % 3];
    // One relax site serves the lane groups and the wave list: each lane
    // walks (vertex, first edge, stride, end) -- its group's light vertex,
    // then the list's vertices in turn.
    int *const heavyList = sh.heavy[threadIdx.x / 64];
    unsigned heavyCount = 0;  // wave-uniform
    const unsigned long long below = (1ull << lane) - 1ull;
    for (unsigned base = lo;; base += perPass) {
      const bool more = base < hi;  // team-uniform
      int cu = -1, cx = 0, cend = 0, cstep = static_cast<int>(lanes);
      u64 cword = 0;
      if (more) {
        const unsigned i = base + (tm.tid >> laneBits);
        bool heavy = false;
        if (i < hi) {
          const int u = all ? static_cast<int>(i) : in[i];
          // Independent loads, issued together.
          const u64 word = packed[u];
          const int begin = outRow[u];
          const int end = outRow[u + 1];
          if (word != PACKED_INF) {
            const int deg = end - begin;
            heavy = deg > HEAVY_GROUP;
            cu = u;
            if (!heavy) {
              cword = word;
              cx = begin + static_cast<int>(sub);
              cend = end;
            } else if (sub == 0) {
              if (deg > hubEdges) {
                const unsigned slot = atomicAdd(hubCnt, 1u);
                if (slot < static_cast<unsigned>(HUB_LIST)) {
                  hubOut[slot] = u;
                  heavy = false;
                }
              }
              if (heavy && tm.whole && deg > groupEdges) {
                const unsigned slot = atomicAdd(&sh.wgCount, 1u);
                if (slot < WG_LIST) {
                  sh.wgList[slot] = u;
                  heavy = false;
                }
              }
            }
          }
        }
        const bool lead = heavy && sub == 0;
        const unsigned long long hv = __ballot(lead);
        if (lead) {
          heavyList[heavyCount + __popcll(hv & below)] = cu;
        }
        heavyCount += __popcll(hv);
      }
      const bool drainNow = !more || heavyCount > HEAVY_LIST - 64;
      for (unsigned j = 0;;) {
        for (; cx < cend; cx += cstep) {
          relax(cu, cword, cx);
        }
        if (!drainNow || j >= heavyCount) {
          break;
        }
        __builtin_amdgcn_wave_barrier();
        cu = heavyList[j++];
        const u64 word = packed[cu];
        cx = outRow[cu] + lane;
        cend = word != PACKED_INF ? outRow[cu + 1] : 0;
        cword = word;
        cstep = 64;
      }
      if (drainNow) {
        heavyCount = 0;
        __builtin_amdgcn_wave_barrier();
      }
      if (!more) {
        break;
      }
      if (tm.whole) {
        __syncthreads();
        if (sh.wgCount > WG_LIST / 2) drainGroup();
        if (sh.near.count[0] > tm.cap / 2) queue_flush(sh.near, outNear, nextCnt, listCap, tm.cap);
        if (sh.far.count[0] > tm.cap / 2) queue_flush(sh.far, farBuf, farCnt, listCap, tm.cap);
      }
    }
    if (tm.whole) {
      __syncthreads();
      if (sh.wgCount > 0) drainGroup();
    }
  }
  // Local hops: relax, within this iteration, the near vertices this
  // workgroup has just listed (still in its LDS queue), then those these
  // list, localHops rounds.  They stay listed for the next iteration as well
  // (relaxing a vertex again is always valid), so this only brings
  // improvements forward: up to localHops + 1 hops advance per grid barrier,
  // which halves the iterations.  A round takes at most one pass (perPass
  // vertices): a workgroup that listed many must not hold the grid at the
  // next barrier, and the capped rounds make more of them affordable.  No
  // flush inside the rounds (the entries being read stay in place; a full
  // queue spills to the global list).  Heavy vertices are left to the next
  // iteration's balanced handling.  Compiled into the kernels without hub
  // handling (single update and batched): in the hub variants the extra loop
  // pushes register use past the budget and the spills cost more than the
  // hops save.
  if (LOCAL != 0 && tm.whole) {
    unsigned from = 0;
    for (int hop = 0; hop < localHops; ++hop) {
      __syncthreads();
      unsigned to = sh.near.count[0] < tm.cap ? sh.near.count[0] : tm.cap;
      to = to < from + perPass ? to : from + perPass;
      if (from >= to) {
        break;
      }
      for (unsigned base = from; base < to; base += perPass) {
        const unsigned i = base + (tm.tid >> laneBits);
        if (i < to) {
          const int u = sh.near.buf[i];
          const u64 word = packed[u];
          const int begin = outRow[u];
          const int end = outRow[u + 1];
          if (word != PACKED_INF && (!HUBS || end - begin <= HEAVY_GROUP)) {
            for (int x = begin + static_cast<int>(sub); x < end; x += static_cast<int>(lanes)) {
              relax(u, word, x);
            }
          }
        }
      }
      from = to;
    }
  }
  if (tm.tid == 0) {
    sh.sliceList[0][tm.slice] = outNear;
    sh.sliceCount[0][tm.slice] = nextCnt;
    sh.sliceList[1][tm.slice] = farBuf;
    sh.sliceCount[1][tm.slice] = farCnt;
    sh.sliceFarSlot[tm.slice] = e % 3;
  }
}

template <bool HUBS, int LOCAL>
__device__ int near_far(Desc *objsArg, int nobj, ObjCtl *ctlsArg,
                        int source, int generation, u64 initThr,
                        const u64 *deltaOverride, int laneBits, int localHops,
                        unsigned n, unsigned listCap, GridBarrier &grid, Threads t,
                        Shared &sh, int profBase) {
  Desc *const objs = uniform(objsArg);
  ObjCtl *const ctls = uniform_global(ctlsArg);
  (void)profBase;
  // Words the barrier publishes per objective: near count, far count, far
  // minimum (two words) and, in the hub variant, the listed hubs.
  constexpr int MW = HUBS ? 5 : 4;
  if (threadIdx.x < static_cast<unsigned>(nobj)) {
    const int k = threadIdx.x;
    sh.thr[k] = initThr;
    sh.epoch[k] = 0;
    sh.delta[k] = deltaOverride != nullptr ? *deltaOverride : objs[k].delta;
    sh.statIter[k] = 0;
    sh.statEpoch[k] = 0;
    sh.statPush[k] = 0;
  }
  for (unsigned it = 0;; ++it) {
    SUBPROF_MARK(tTop);
    const unsigned old = (it + 2) % 3;
    if (threadIdx.x == 0) {
      sh.any = 0;
    }
    __syncthreads();
    // Snapshot: every workgroup reads the same stable counters, so every
    // thread of the grid takes the same decisions.
    // Iteration 0 reads the counters; later ones get them from the barrier.
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      const int e = sh.epoch[k];
      unsigned nc, fc, hc = 0;
      u64 fm;
      if (it == 0) {
        nc = load_relaxed(&c.nearCount[0]);
        fc = load_relaxed(&c.farCount[e % 3]);
        fm = load_relaxed(&c.farMin[e % 3]);
        if (HUBS) hc = load_relaxed(&c.hubCount[0]);
      } else {
        nc = sh.snap[MW * k];
        fc = sh.snap[MW * k + 1];
        fm = static_cast<u64>(sh.snap[MW * k + 2]) |
             (static_cast<u64>(sh.snap[MW * k + 3]) << 32);
        if (HUBS) hc = sh.snap[MW * k + 4];
      }
      sh.hubs[k] = hc;
      if (nc > 0 || hc > 0) {
        sh.mode[k] = MODE_PUSH;
        sh.cnt[k] = nc;
        sh.any = 1;
      } else {
        if (fc > 0) {
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
      partition_blocks(sh, nobj, t.blocks);
    }
    __syncthreads();
    SUBPROF_MARK(tSnap);
    const int gen = ++generation;
    if (t.block == 0 && threadIdx.x < static_cast<unsigned>(nobj)) {
      // Slots that nobody reads or appends to in this iteration.
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      store_relaxed(&c.nearCount[old], 0u);
      if (HUBS) store_relaxed(&c.hubCount[old], 0u);
      if (sh.mode[k] == MODE_SPLIT) {
        const int e2 = (sh.epoch[k] + 2) % 3;
        store_relaxed(&c.farCount[e2], 0u);
        store_relaxed(&c.farMin[e2], PACKED_INF);
        sh.statEpoch[k] += 1;
      } else if (sh.mode[k] == MODE_PUSH) {
        sh.statIter[k] += 1;
        sh.statPush[k] += sh.cnt[k];
      }
    }
    // The objectives of this workgroup; with several, its waves split into
    // teams (up to one per wave) that run them concurrently.
    if (threadIdx.x == 0) {
      int mine = 0;
      for (int k = 0; k < nobj; ++k) {
        if (sh.mode[k] != MODE_DONE && static_cast<int>(t.block) >= sh.blockBegin[k] &&
            static_cast<int>(t.block) < sh.blockEnd[k]) {
          sh.mine[mine++] = k;
        }
      }
      sh.nmine = mine;
    }
    __syncthreads();
    const int nmine = sh.nmine;
    for (int first = 0; first < nmine; first += static_cast<int>(WAVES)) {
      const unsigned teams = min(WAVES, static_cast<unsigned>(nmine - first));
      const unsigned team = (threadIdx.x / 64) * teams / WAVES;
      const unsigned w0 = (team * WAVES + teams - 1) / teams;
      const unsigned w1 = ((team + 1) * WAVES + teams - 1) / teams;
      const Team tm{team, threadIdx.x - 64 * w0, 64 * (w1 - w0), sh.queueCap / teams, teams == 1};
      const int k = sh.mine[first + static_cast<int>(team)];
      if (threadIdx.x < teams) {
        sh.sliceFarMin[threadIdx.x] = PACKED_INF;
      }
      __syncthreads();
      const unsigned rank = t.block - static_cast<unsigned>(sh.blockBegin[k]);
      const unsigned blocks = static_cast<unsigned>(sh.blockEnd[k] - sh.blockBegin[k]);
      if (sh.mode[k] == MODE_PUSH) {
        push_team<HUBS, LOCAL>(objs[k], ctls[k], sh, k, it, source, gen, rank, blocks,
                               laneBits, localHops, n, listCap, tm);
      } else {
        push_team<HUBS, LOCAL, true>(objs[k], ctls[k], sh, k, it, source, gen, rank, blocks,
                                     laneBits, localHops, n, listCap, tm);
      }
      if (tm.tid == 0) {
        sh.sliceObj[team] = k;
      }
      flush_slices(sh.near, teams, tm.cap, sh.sliceList[0], sh.sliceCount[0], listCap);
      flush_slices(sh.far, teams, tm.cap, sh.sliceList[1], sh.sliceCount[1], listCap);
      if (threadIdx.x < teams && sh.sliceFarMin[threadIdx.x] != PACKED_INF) {
        atomicMin(&ctls[sh.sliceObj[threadIdx.x]].farMin[sh.sliceFarSlot[threadIdx.x]],
                  sh.sliceFarMin[threadIdx.x]);
      }
    }
    // The barrier publishes the counters the next iteration starts from.
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      const int eNext = sh.epoch[k] + (sh.mode[k] == MODE_SPLIT ? 1 : 0);
      unsigned *fm = reinterpret_cast<unsigned *>(&c.farMin[eNext % 3]);
      sh.src[MW * k] = &c.nearCount[(it + 1) % 3];
      sh.src[MW * k + 1] = &c.farCount[eNext % 3];
      sh.src[MW * k + 2] = fm;
      sh.src[MW * k + 3] = fm + 1;
      if (HUBS) sh.src[MW * k + 4] = &c.hubCount[(it + 1) % 3];
    }
    SUBPROF_MARK(tWork);
    if (!grid.sync(sh.src, MW * nobj, sh.snap)) {
      return generation;
    }
    SUBPROF_MARK(tSync);
    SUBPROF_ADD(profBase + 0, tTop, tSnap);
    SUBPROF_ADD(profBase + 1, tSnap, tWork);
    SUBPROF_ADD(profBase + 2, tWork, tSync);
#ifdef MOSP_SUBPROFILE
    if (grid.wasLast) SUBPROF_ADD(profBase == 8 ? 6 : 7, tWork, tSync);
#endif
    SUBPROF_ADD(profBase + 3, 0, 1);
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      if (sh.mode[k] == MODE_SPLIT) {
        sh.thr[k] = sh.newThr[k];
        sh.epoch[k] += 1;
      }
    }
  }
  if (t.block == 0 && threadIdx.x < static_cast<unsigned>(nobj)) {
    const int k = threadIdx.x;
    atomicAdd(&ctls[k].iterations, sh.statIter[k]);
    atomicAdd(&ctls[k].epochs, sh.statEpoch[k]);
    atomicAdd(&ctls[k].pushes, sh.statPush[k]);
  }
  return generation;
}

// ---------------------------------------------------------------------------
// Asynchronous near-far (Step 2 and the searches of the single-update kernel
// without hub handling).  In near_far() every iteration ends at a grid
// barrier, and on a road grid ~40 % of an iteration goes to waiting for the
// last workgroup and the release.  Here only the threshold raises (epochs)
// are grid-synchronous.  Within the phase between two barriers:
//   * each workgroup relaxes batches of up to perPass vertices taken from
//     its own LDS queue (a FIFO ring); the vertices a batch improves below
//     the threshold are appended to that queue, those at or above it to the
//     far pile, as in near_far();
//   * a workgroup whose queue holds more than spillHi entries moves its
//     newest entries down to spillLo into its XCD's segment of the
//     objective's shared pool (one atomic reservation); a workgroup whose
//     queue is empty takes up to perPass entries from a pool (load, then a
//     compare-and-swap of the pool word), its XCD's segment first;
//   * pool slots hold -1 when free: a taker waits (bounded) until its slot
//     is written, and writes -1 back, so the pools are free again when the
//     phase ends;
//   * the word of a listed vertex is read with a device-scope load: the
//     vertex was listed after an atomicMin improved it, and the relaxation
//     must start from at least that word (a stale copy of its line in this
//     XCD's L2 would lose the improvement);
//   * termination: every workgroup publishes (phase, entries appended,
//     entries processed) after each step; an appended entry is counted
//     before another workgroup can take it (the pool reservation is issued
//     after the exchange that published the count has returned).  Idle,
//     workgroup 0 reads every status word twice; if all carry this phase,
//     the sums did not change between the reads and appended == processed,
//     then at an instant between the reads no entry was left -- and none can
//     be appended any more -- so the phase is over, which it writes to 16
//     mailboxes that idle workgroups poll.  Each workgroup's seeding token
//     (one entry appended at the start of the phase, processed once its
//     share of the seeds is listed) keeps the phase open until all seeds are
//     listed;
//   * a near entry that finds no room (full LDS queue) is dropped and the
//     objective's lost flag set: the next phase lists every vertex under the
//     threshold again (relaxing a vertex again is always valid).
// Results are those of near_far(): every vertex improved below the
// threshold is relaxed again after its last improvement, with a word at
// least as good, in whatever order.
// ---------------------------------------------------------------------------

enum : int { MODE_SWEEP = 3 };  // list every vertex under the threshold again
enum : int { STEP_SEED = 0, STEP_LOCAL = 1, STEP_CLAIM = 2, STEP_POOL = 3, STEP_IDLE = 4 };

// Status word: phase (4 bits: the phase sequence number mod 15) | appended
// (30) | processed (30).  Phase 15 never occurs: launches start every
// workgroup's word with it (ASYNC_STATUS_INIT).
__device__ __forceinline__ unsigned status_phase(unsigned seq) { return seq % 15u; }
__device__ __forceinline__ u64 status_word(unsigned seq, unsigned sent, unsigned recv) {
  return (static_cast<u64>(status_phase(seq)) << 60) |
         (static_cast<u64>(sent & 0x3fffffffu) << 30) | static_cast<u64>(recv & 0x3fffffffu);
}

__device__ __forceinline__ u64 *pool_word(AsyncCtl *A, unsigned seq, int k, int seg) {
  return &A->pool[seq & 1][k][seg][0];
}

// Workgroup-uniform: publish the far queue of objective k (its pile of the
// epoch it appends to) and its distance minimum.
__device__ void async_flush_far(Desc *objs, ObjCtl *ctls, Shared &sh, int k, unsigned listCap,
                                unsigned qcap) {
  const int e = sh.epoch[k] + (sh.mode[k] == MODE_SPLIT ? 1 : 0);
  queue_flush(sh.far, objs[k].s.far[e & 1], &ctls[k].farCount[e % 3], listCap, qcap);
  if (threadIdx.x == 0 && sh.as.farMin != PACKED_INF) {
    atomicMin(&ctls[k].farMin[e % 3], sh.as.farMin);
    sh.as.farMin = PACKED_INF;
  }
  __syncthreads();
}

// Thread 0: this workgroup's share of objective k's seeds for the phase.
__device__ void async_seed_share(Shared &sh, int k, unsigned blockRank, unsigned n,
                                 unsigned listCap) {
  const unsigned rank = blockRank - static_cast<unsigned>(sh.blockBegin[k]);
  const unsigned blocks = static_cast<unsigned>(sh.blockEnd[k] - sh.blockBegin[k]);
  unsigned count = sh.cnt[k];
  if (sh.mode[k] == MODE_SWEEP || (sh.mode[k] == MODE_SPLIT && count > listCap)) {
    count = n;
  }
  sh.as.seedPos = static_cast<unsigned>(static_cast<u64>(count) * rank / blocks);
  sh.as.seedEnd = static_cast<unsigned>(static_cast<u64>(count) * (rank + 1) / blocks);
}

// near_far() without per-iteration barriers, for objectives objs[0, nobj):
// same entry state (near list 0 or far pile 0, first threshold initThr),
// same exit state (final words in the objectives' packed arrays).
// phaseSeq counts the launch's phases (every workgroup the same).  LOCAL:
// the caller's kernel (1 single update, 2 batched), so that each kernel has
// its own copy and register allocation.
template <int LOCAL>
__device__ int near_far_async(Desc *objsArg, int nobj, ObjCtl *ctlsArg, AsyncCtl *Aarg,
                              int source, int generation, u64 initThr,
                              const u64 *deltaOverride, int laneBits, unsigned n,
                              unsigned listCap, unsigned segCap, unsigned spillHi,
                              unsigned spillLo, unsigned flags, unsigned keepHungry,
                              unsigned *abortWord, unsigned &phaseSeq,
                              GridBarrier &grid, Threads t, Shared &sh) {
  Desc *const objs = uniform(objsArg);
  Ctl *const ctlAll = uniform_global(grid.ctl);
  ObjCtl *const ctls = uniform_global(ctlsArg);
  AsyncCtl *const A = uniform_global(Aarg);
  constexpr int MW = 4;  // per objective: far count, far minimum (two words), lost flag
  const unsigned qcap = sh.queueCap;                      // far queue
  const unsigned lcap = qcap < LQ_CAP ? qcap : LQ_CAP;    // local queue
  const unsigned lanes = 1u << laneBits;
  const unsigned perPass = BLOCK >> laneBits;
  const unsigned sub = threadIdx.x & (lanes - 1);
  const unsigned seedPass = lcap / 2 < static_cast<unsigned>(BLOCK) ? lcap / 2 : BLOCK;
  spillHi = spillHi < lcap ? spillHi : lcap - 1;
  spillLo = spillLo < spillHi ? spillLo : spillHi / 2;
  const bool carry = (flags & ASYNC_CARRY) != 0;
  const bool useHungry = (flags & ASYNC_HUNGRY) != 0;
  const int mySeg = static_cast<int>(t.block % ASYNC_SEGS);
  const u64 limit = grid.limit;
  if (threadIdx.x < static_cast<unsigned>(nobj)) {
    const int k = threadIdx.x;
    sh.thr[k] = initThr;
    sh.epoch[k] = 0;
    sh.delta[k] = deltaOverride != nullptr ? *deltaOverride : objs[k].delta;
    sh.statIter[k] = 0;
    sh.statEpoch[k] = 0;
    sh.statPush[k] = 0;
  }
  if (threadIdx.x == 0) {
    sh.as.lqHead = sh.as.lqTail = 0;
    sh.as.cur = 0;
    sh.as.farMin = PACKED_INF;
    sh.as.lostNear = 0;
    sh.as.batches = 0;
  }
  __syncthreads();
  // Thread 0: what the workgroup does next, given `have` listed entries;
  // published by the barrier that ends the current step.
  auto decide = [&](unsigned have) {
    const bool seeding = sh.as.seedIdx < sh.nmine;
    int step;
    if (seeding && have < perPass && have + seedPass <= lcap &&
        (have == 0 || sh.as.cur == sh.mine[sh.as.seedIdx])) {
      step = STEP_SEED;
    } else if (have > 0) {
      step = STEP_LOCAL;
      sh.as.count = have < perPass ? have : perPass;
    } else if (seeding) {
      step = STEP_SEED;
    } else {
      step = STEP_CLAIM;
    }
    sh.as.step = step;
  };
  for (bool first = true;; first = false) {
    const unsigned seq = phaseSeq;
    if (threadIdx.x == 0) {
      sh.any = 0;
    }
    __syncthreads();
    // Snapshot: every workgroup takes the same decisions.
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      const int e = sh.epoch[k];
      unsigned nc = 0, fc, lost = 0;
      u64 fm;
      if (first) {
        nc = load_relaxed(&c.nearCount[0]);
        fc = load_relaxed(&c.farCount[e % 3]);
        fm = load_relaxed(&c.farMin[e % 3]);
      } else {
        fc = sh.snap[MW * k];
        fm = static_cast<u64>(sh.snap[MW * k + 1]) | (static_cast<u64>(sh.snap[MW * k + 2]) << 32);
        lost = sh.snap[MW * k + 3];
      }
      sh.hubs[k] = 0;
      sh.newThr[k] = sh.thr[k];  // the threshold in force during the phase
      if (nc > 0) {
        sh.mode[k] = MODE_PUSH;  // the near list is the seed list
        sh.cnt[k] = nc;
        sh.any = 1;
      } else if (lost != 0) {
        sh.mode[k] = MODE_SWEEP;
        sh.cnt[k] = n;
        sh.any = 1;
        if (t.block == 0) {
          atomicAdd(&ctlAll->asyncSweeps, 1u);
        }
      } else if (fc > 0) {
        const u64 base = fm > sh.thr[k] && fm != PACKED_INF ? fm : sh.thr[k];
        sh.mode[k] = MODE_SPLIT;
        sh.cnt[k] = fc;
        sh.newThr[k] = base + sh.delta[k];
        sh.any = 1;
      } else {
        sh.mode[k] = MODE_DONE;
      }
    }
    __syncthreads();
    if (!sh.any) {
      break;
    }
    ++phaseSeq;
    const int gen = ++generation;
    if (threadIdx.x == 0) {
      partition_blocks(sh, nobj, t.blocks);
      int mine = 0;
      for (int k = 0; k < nobj; ++k) {
        if (sh.mode[k] != MODE_DONE && static_cast<int>(t.block) >= sh.blockBegin[k] &&
            static_cast<int>(t.block) < sh.blockEnd[k]) {
          sh.mine[mine++] = k;
        }
      }
      sh.nmine = mine;
    }
    if (t.block == 0) {
      // The other pool set, hungry words and lost slots were last used by the
      // previous phase, which every workgroup has left: clear them for the
      // next one.
      for (unsigned i = threadIdx.x; i < static_cast<unsigned>(MOSP_MAX_OBJ) * ASYNC_SEGS; i += BLOCK) {
        store_relaxed(pool_word(A, seq + 1, static_cast<int>(i / ASYNC_SEGS),
                                static_cast<int>(i % ASYNC_SEGS)),
                      0ull);
      }
      if (threadIdx.x < static_cast<unsigned>(ASYNC_SEGS)) {
        store_relaxed(&A->hungry[(seq + 1) & 1][threadIdx.x][0], 0u);
      }
      if (threadIdx.x < static_cast<unsigned>(nobj)) {
        const int k = threadIdx.x;
        ObjCtl &c = ctls[k];
        store_relaxed(&c.lost[(seq + 1) & 1], 0u);
        if (sh.mode[k] == MODE_SPLIT) {
          const int e2 = (sh.epoch[k] + 2) % 3;
          store_relaxed(&c.farCount[e2], 0u);
          store_relaxed(&c.farMin[e2], PACKED_INF);
          sh.statEpoch[k] += 1;
        }
        if (sh.mode[k] != MODE_DONE) {
          sh.statIter[k] += 1;
        }
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      sh.as.seedIdx = 0;
      if (sh.nmine > 0) {
        sh.as.cur = sh.mine[0];
        async_seed_share(sh, sh.mine[0], t.block, n, listCap);
      }
      // Counts of this phase; the seeding token counts as appended until
      // this workgroup's seeds are listed.
      sh.as.sent = 1;
      sh.as.recv = sh.nmine == 0 ? 1 : 0;
      sh.as.exit = 0;
      sh.as.polls = 0;
      sh.as.hungry = 0;
      store_relaxed(&A->status[t.block], status_word(seq, sh.as.sent, sh.as.recv));
      decide(sh.as.lqTail - sh.as.lqHead);
    }
    __syncthreads();
    const u64 phaseStart = wall_clock64();
    unsigned tailBefore = sh.as.lqTail;  // thread 0: the ring's tail when a step begins
    for (;;) {
      // The step was decided by thread 0 before the barrier that ended the
      // previous one (or the phase's start).  sh.as.step is written only by
      // thread 0 right before such a barrier, so every thread reads the same
      // value.
      SUBPROF_MARK(tStep);
      const int decided = sh.as.step;
      if (decided == STEP_SEED && sh.mine[sh.as.seedIdx] != sh.as.cur) {
        async_flush_far(objs, ctls, sh, sh.as.cur, listCap, qcap);
        if (threadIdx.x == 0) {
          sh.as.cur = sh.mine[sh.as.seedIdx];
        }
        __syncthreads();
      }
      if (decided == STEP_CLAIM) {
        // Wave 0 looks at two pool segments of every active objective --
        // its XCD's (lanes 0-31), where spills for its idle workgroups land,
        // and one other in turn (lanes 32-63) -- and takes up to perPass
        // entries of one: its own XCD's first, the current objective's first.
        if (threadIdx.x < 64) {
          const int lane = static_cast<int>(threadIdx.x);
          const int kk = lane & 31;
          const int seg = lane < 32 ? mySeg
                                    : static_cast<int>((mySeg + 1 + sh.as.polls % (ASYNC_SEGS - 1)) %
                                                       ASYNC_SEGS);
          unsigned avail = 0, head = 0;
          u64 word = 0;
          u64 *pw = nullptr;
          if (kk < nobj && sh.mode[kk] != MODE_DONE) {
            pw = pool_word(A, seq, kk, seg);
            word = load_relaxed(pw);
            head = static_cast<unsigned>(word);
            unsigned tail = static_cast<unsigned>(word >> 32);
            tail = tail < segCap ? tail : segCap;
            avail = tail > head ? tail - head : 0;
          }
          const u64 some = __ballot(avail > 0);
          const int cur = sh.as.cur;
          int pick = -1;
          if ((some >> cur) & 1ull) {
            pick = cur;
          } else if ((some & 0xffffffffull) != 0) {
            pick = __ffsll(static_cast<long long>(some & 0xffffffffull)) - 1;
          } else if ((some >> (32 + cur)) & 1ull) {
            pick = 32 + cur;
          } else if (some != 0) {
            pick = __ffsll(static_cast<long long>(some)) - 1;
          }
          // The outcome goes to its own word: slower waves may still be
          // reading sh.as.step.
          if (lane == 0) {
            sh.as.claimed = STEP_IDLE;
          }
          if (lane == pick) {
            const unsigned c = avail < perPass ? avail : perPass;
            const u64 old = atomicCAS(pw, word, word + c);
            if (old == word) {
              sh.as.claimed = STEP_POOL;
              sh.as.count = c;
              sh.as.seg = seg;
              sh.as.base = head;
              sh.as.obj = kk;
            }
          }
        }
        __syncthreads();
        if (threadIdx.x == 0 && sh.as.hungry != 0 && sh.as.claimed == STEP_POOL) {
          atomicSub(&A->hungry[seq & 1][mySeg][0], 1u);
          sh.as.hungry = 0;
        }
        if (sh.as.claimed == STEP_POOL) {
          if (sh.as.obj != sh.as.cur) {
            async_flush_far(objs, ctls, sh, sh.as.cur, listCap, qcap);
            if (threadIdx.x == 0) {
              sh.as.cur = sh.as.obj;
            }
          }
          if (threadIdx.x == 0) {
            sh.as.polls = 0;
          }
          __syncthreads();
        }
      }
      const int step = decided == STEP_CLAIM ? sh.as.claimed : decided;
      const unsigned batchCount = sh.as.count;
      (void)batchCount;
      SUBPROF_MARK(tDecided);
      if (step == STEP_IDLE) {
        // Nothing to take: workgroup 0 checks for the end of the phase, the
        // others poll its mailbox; every wait is bounded.
        if (t.block == 0) {
          if (threadIdx.x < 64) {
            const unsigned lane = threadIdx.x;
            bool sameOk = true;
            u64 s1 = 0, r1 = 0, s2 = 0, r2 = 0;
            for (int pass = 0; pass < 2; ++pass) {
              u64 s = 0, r = 0;
              bool ok = true;
              for (unsigned j = lane; j < t.blocks; j += 64) {
                const u64 w = load_relaxed(&A->status[j]);
                ok = ok && static_cast<unsigned>(w >> 60) == status_phase(seq);
                s += (w >> 30) & 0x3fffffffu;
                r += w & 0x3fffffffu;
              }
              sameOk = sameOk && __all(ok);
#pragma unroll
              for (int offset = 32; offset > 0; offset >>= 1) {
                s += shfl_u64(s, static_cast<int>(lane ^ offset));
                r += shfl_u64(r, static_cast<int>(lane ^ offset));
              }
              if (pass == 0) {
                s1 = s;
                r1 = r;
              } else {
                s2 = s;
                r2 = r;
              }
            }
            if (sameOk && s1 == s2 && r1 == r2 && s2 == r2) {
              if (lane < static_cast<unsigned>(BARRIER_GROUPS)) {
                store_relaxed(&A->done[lane][0], seq + 1);
              }
              if (lane == 0) {
                sh.as.exit = 1;
              }
            }
          }
        } else if (threadIdx.x == 0) {
          if (load_relaxed(&A->done[t.block % BARRIER_GROUPS][0]) >= seq + 1) {
            sh.as.exit = 1;
          }
        }
        if (threadIdx.x == 0 && sh.as.exit == 0) {
          sh.as.step = STEP_CLAIM;  // the ring is empty and the seeds are listed
          if (useHungry && sh.as.hungry == 0) {
            atomicAdd(&A->hungry[seq & 1][mySeg][0], 1u);
            sh.as.hungry = 1;
          }
          const unsigned polls = ++sh.as.polls;
          if ((polls & 15u) == 0 &&
              (load_relaxed(abortWord) != 0 || wall_clock64() - phaseStart > limit)) {
            store_relaxed(abortWord, 1u);
            sh.as.exit = 2;
          } else if (polls > ASYNC_SEGS) {
            // Every segment looked at once: back off (~1 us).
            __builtin_amdgcn_s_sleep(32);
            __builtin_amdgcn_s_sleep(32);
          }
        }
        __syncthreads();
        SUBPROF_MARK(tIdle);
        SUBPROF_ADD(1, tStep, tIdle);
        SUBPROF_ADD(7, 0, 1);
        if (sh.as.exit != 0) {
          break;
        }
        continue;
      }
      // A batch of the local queue, a batch from a pool, or a seeding pass.
      const int k = sh.as.cur;
      Desc &D = objs[k];
      ObjCtl &c = ctls[k];
      const int bits = D.parentBits;
      const u64 noParent = D.noParent;
      const u64 thr = sh.newThr[k];
      const u64 maxDist = D.maxDistance;
      const bool optimistic = D.optimistic != 0;
      const int ea = sh.epoch[k] + (sh.mode[k] == MODE_SPLIT ? 1 : 0);  // far pile appended to
      u64 *const packed = uniform_global(D.s.packed);
      int *const inFar = uniform_global(D.s.inFar);
      const int *const outRow = uniform_global(D.outRow);
      const int2 *const outRec = uniform_global(D.outEdge);
      const int *const outCol = uniform_global(D.outCol);
      const int *const outW = uniform_global(D.outW);
      int *const farBuf = uniform_global(D.s.far[ea & 1]);
      unsigned *const farCnt = &c.farCount[ea % 3];
      const unsigned headBatch = sh.as.lqHead;
      // An XCD with idle workgroups, read while the batch runs (thread 0).
      const int hseg = static_cast<int>((mySeg + sh.as.batches) % ASYNC_SEGS);
      unsigned hungryThere = 0;
      if (useHungry && threadIdx.x == 0) {
        hungryThere = load_relaxed(&A->hungry[seq & 1][hseg][0]);
      }
      const int4 *const rec4 = uniform_global(D.outEdge4);
      auto listNear = [&](int v, u64 word, int begin, int end) {
        const unsigned slot = wave_reserve(&sh.as.lqTail);
        if (slot - headBatch < lcap) {
          const unsigned at = slot & (LQ_CAP - 1);
          *(__attribute__((address_space(3))) int *)(sh.lq.v + at) = v;
          *(__attribute__((address_space(3))) int *)(sh.lq.b + at) = begin;
          *(__attribute__((address_space(3))) int *)(sh.lq.e + at) = end;
          *(__attribute__((address_space(3))) u64 *)(sh.lq.w + at) = word;
        } else {
          sh.as.lostNear = 1;
        }
      };
      auto listFar = [&](int v, u64 d) {
        slice_push(sh.far, 0, qcap, v, farBuf, farCnt, listCap);
        atomicMin(&sh.as.farMin, d);
      };
      if (step == STEP_SEED) {
        const unsigned i = sh.as.seedPos + threadIdx.x;
        if (threadIdx.x < seedPass && i < sh.as.seedEnd) {
          const int mode = sh.mode[k];
          if (mode == MODE_PUSH) {
            const int v = D.s.near[0][i];
            listNear(v, load_relaxed(&packed[v]), -1, 0);
          } else if (mode == MODE_SWEEP) {
            const u64 word = load_relaxed(&packed[i]);
            if (word != PACKED_INF && dist_of(word, bits) < thr) {
              listNear(static_cast<int>(i), word, -1, 0);
            }
          } else {
            const bool all = sh.cnt[k] > listCap;
            const int v = all ? static_cast<int>(i) : D.s.far[sh.epoch[k] & 1][i];
            if (!all || inFar[v] != 0) {
              const u64 word = load_relaxed(&packed[v]);
              const u64 d = word == PACKED_INF ? PACKED_INF : dist_of(word, bits);
              if (d < thr) {
                inFar[v] = 0;
                if (D.s.stamp[v] != gen) {  // duplicates of v in the pile: list once
                  D.s.stamp[v] = gen;
                  listNear(v, word, -1, 0);
                }
              } else {
                listFar(v, d);
              }
            }
          }
        }
      } else {
        const unsigned count = sh.as.count;
        const unsigned i = threadIdx.x >> laneBits;
        int u = -1, begin = -1, end = 0;
        u64 word = PACKED_INF;
        if (i < count) {
          if (step == STEP_LOCAL) {
            // The word the vertex was listed with (a later improvement lists
            // it again with its own word) and its edge range.
            const unsigned at = (headBatch + i) & (LQ_CAP - 1);
            u = sh.lq.v[at];
            word = sh.lq.w[at];
            begin = sh.lq.b[at];
            end = sh.lq.e[at];
          } else {
            int *const slot = D.s.pool + static_cast<size_t>(sh.as.seg) * segCap + sh.as.base + i;
            u = load_relaxed(slot);
            if (u == -1) {
              const u64 start = wall_clock64();
              while ((u = load_relaxed(slot)) == -1) {
                if (wall_clock64() - start > limit) {
                  sh.as.exit = 2;
                  break;
                }
                __builtin_amdgcn_s_sleep(1);
              }
            }
            if (sub == 0 && u != -1) {
              store_relaxed(slot, -1);
            }
          }
          if (u < -1 || u >= static_cast<int>(n)) {
            sh.as.exit = 2;  // not a vertex: never used as an index
            u = -1;
          }
        }
        if (u >= 0) {
          if (begin < 0) {
            begin = outRow[u];
            end = outRow[u + 1];
          }
          // Edge x as {column, weight, rowBegin, rowEnd} (row range -1: not
          // known), from whichever records the graph holds.
          auto edgeAt = [&](int x) -> int4 {
            if (rec4 != nullptr) {
              return rec4[x];
            }
            if (outRec != nullptr) {
              const int2 r = outRec[x];
              return make_int4(r.x, r.y, -1, 0);
            }
            return make_int4(outCol[x], outW[x], -1, 0);
          };
          // The first record is requested before the word is awaited.
          const int L = static_cast<int>(lanes);
          int x = begin + static_cast<int>(sub);
          int4 next = make_int4(-1, 0, -1, 0);
          if (x < end) {
            next = edgeAt(x);
          }
          if (step == STEP_POOL || !carry) {
            word = load_relaxed(&packed[u]);
          }
          if (word != PACKED_INF) {
            const u64 du = dist_of(word, bits);
            const int pu = bits != 0 ? static_cast<int>(word & ((1ull << bits) - 1ull)) : -1;
            // Two edges per lane and round, both atomics in flight together
            // (a batch otherwise waits for a second round trip whenever one
            // of its vertices has more edges than lanes).
            for (; x < end; x += 2 * L) {
              const int4 ea = next;
              const int4 eb = x + L < end ? edgeAt(x + L) : make_int4(-1, 0, -1, 0);
              if (x + 2 * L < end) {
                next = edgeAt(x + 2 * L);
              }
              const u64 ndA = du + static_cast<u64>(ea.y);
              const u64 ndB = du + static_cast<u64>(eb.y);
              bool goA = ea.x != source && ea.x != pu;  // the source is never relaxed; nor is u's parent
              bool goB = eb.x >= 0 && eb.x != source && eb.x != pu;
              if ((goA && ndA > maxDist) || (goB && ndB > maxDist)) {
                if (optimistic) {
                  store_relaxed(&c.dropped, 1u);
                }
                goA = goA && ndA <= maxDist;
                goB = goB && ndB <= maxDist;
              }
              const bool nearA = ndA < thr, nearB = ndB < thr;
              const int claimA = goA && !nearA ? inFar[ea.x] : 0;
              const int claimB = goB && !nearB ? inFar[eb.x] : 0;
              const u64 candA = pack(ndA, u, bits, noParent);
              const u64 candB = pack(ndB, u, bits, noParent);
              u64 prevA = 0, prevB = 0;
              if (goA) {
                prevA = atomicMin(&packed[ea.x], candA);
              }
              if (goB) {
                prevB = atomicMin(&packed[eb.x], candB);
              }
              if (goA && ndA < dist_of(prevA, bits)) {
                if (nearA) {
                  listNear(ea.x, candA, ea.z, ea.w);
                } else if (claimA == 0) {
                  inFar[ea.x] = 1;
                  listFar(ea.x, ndA);
                }
              }
              if (goB && ndB < dist_of(prevB, bits)) {
                if (nearB) {
                  listNear(eb.x, candB, eb.z, eb.w);
                } else if (claimB == 0) {
                  inFar[eb.x] = 1;
                  listFar(eb.x, ndB);
                }
              }
            }
          }
        }
      }
      __syncthreads();
      SUBPROF_MARK(tRelaxed);
      // Bookkeeping (thread 0): consume the batch, count, publish, spill.
      if (threadIdx.x == 0) {
        unsigned tail = sh.as.lqTail;
        if (tail - headBatch > lcap) {
          tail = headBatch + lcap;  // appends past the room were dropped
        }
        if (sh.as.lostNear != 0) {
          store_relaxed(&c.lost[seq & 1], 1u);
          sh.as.lostNear = 0;
          atomicAdd(&ctlAll->asyncLost, 1u);
          SUBPROF_ADD(13, 0, 1);
        }
        sh.as.lqTail = tail;
        sh.as.sent += tail - tailBefore;
        if (step == STEP_SEED) {
          sh.as.seedPos += seedPass;
          if (sh.as.seedPos >= sh.as.seedEnd) {
            if (++sh.as.seedIdx < sh.nmine) {
              async_seed_share(sh, sh.mine[sh.as.seedIdx], t.block, n, listCap);
            } else {
              sh.as.recv += 1;  // the seeding token
            }
          }
        } else {
          if (step == STEP_LOCAL) {
            sh.as.lqHead += sh.as.count;
          }
          sh.as.recv += sh.as.count;
          sh.statPush[k] += sh.as.count;
        }
        const unsigned have = sh.as.lqTail - sh.as.lqHead;
        const u64 status = status_word(seq, sh.as.sent, sh.as.recv);
        sh.as.spillKeep = 0;
        // Spill: the queue holds more than spillHi entries (down to spillLo,
        // into this XCD's segment), or an XCD has idle workgroups (all but
        // keepHungry entries, up to perPass per idle workgroup, into its
        // segment).
        unsigned s = 0;
        int spillSeg = mySeg;
        if (hungryThere > 0 && have > keepHungry) {
          s = min(have - keepHungry, hungryThere * perPass);
          spillSeg = hseg;
        }
        if (have > spillHi && have - spillLo > s) {
          s = have - spillLo;
          spillSeg = mySeg;
        }
        if (s > 0) {
          // The published count must be seen before the entries can be
          // taken: wait for the exchange, then reserve.
          const u64 prev = __hip_atomic_exchange(&A->status[t.block], status, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
          asm volatile("" ::"v"(prev) : "memory");
          const u64 old = atomicAdd(pool_word(A, seq, k, spillSeg), static_cast<u64>(s) << 32);
          const unsigned t0 = static_cast<unsigned>(old >> 32);
          sh.as.spillBase = t0;
          sh.as.spillSeg = spillSeg;
          sh.as.spillKeep = t0 < segCap ? min(s, segCap - t0) : 0;
          if (sh.as.spillKeep < s) {
            atomicAdd(&ctlAll->asyncShortSpills, 1u);
          }
        } else {
          store_relaxed(&A->status[t.block], status);
        }
        if ((++sh.as.batches & 63u) == 0 &&
            (load_relaxed(abortWord) != 0 || wall_clock64() - phaseStart > limit)) {
          store_relaxed(abortWord, 1u);
          sh.as.exit = 2;
        }
        // The spilled entries leave the ring (copied out below, before any
        // later append), then the next step.
        sh.as.lqTail -= sh.as.spillKeep;
        decide(sh.as.lqTail - sh.as.lqHead);
        tailBefore = sh.as.lqTail;
        sh.as.leave = sh.as.exit == 2 ? 1 : 0;
        sh.as.flushFar = sh.far.count[0] > qcap / 2 ? 1 : 0;
      }
      __syncthreads();
      if (sh.as.leave) {
        break;
      }
      const unsigned keep = sh.as.spillKeep;
      if (keep > 0) {
        // The newest `keep` entries, now past the tail, go to slots
        // [spillBase, spillBase + keep) before the next step appends.
        const unsigned from = sh.as.lqTail;
        int *const seg = D.s.pool + static_cast<size_t>(sh.as.spillSeg) * segCap + sh.as.spillBase;
        for (unsigned j = threadIdx.x; j < keep; j += BLOCK) {
          store_relaxed(seg + j, sh.lq.v[(from + j) & (LQ_CAP - 1)]);
        }
        __syncthreads();
      }
      if (sh.as.flushFar) {
        async_flush_far(objs, ctls, sh, k, listCap, qcap);
      }
      SUBPROF_MARK(tBusy);
      SUBPROF_ADD(0, tStep, tBusy);
      if (step != STEP_SEED) {
        SUBPROF_ADD(step == STEP_LOCAL ? 2 : 3, 0, batchCount);
      }
      SUBPROF_ADD(6, tStep, tDecided);
      SUBPROF_ADD(14, tDecided, tRelaxed);
      SUBPROF_ADD(step == STEP_LOCAL ? 4 : step == STEP_POOL ? 5 : 15, 0, 1);
      SUBPROF_ADD(8, 0, keep > 0 ? 1 : 0);
      SUBPROF_ADD(9, 0, keep);

    }
    if (sh.as.exit == 2) {
      if (threadIdx.x == 0) {
        store_relaxed(abortWord, 1u);
      }
      grid.aborted = true;
      return generation;
    }
    // The phase is over everywhere: publish the far queue, then the barrier
    // carries every objective's far count, minimum and lost flag.
    async_flush_far(objs, ctls, sh, sh.as.cur, listCap, qcap);
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      ObjCtl &c = ctls[k];
      const int eNext = sh.epoch[k] + (sh.mode[k] == MODE_SPLIT ? 1 : 0);
      unsigned *fm = reinterpret_cast<unsigned *>(&c.farMin[eNext % 3]);
      sh.src[MW * k] = &c.farCount[eNext % 3];
      sh.src[MW * k + 1] = fm;
      sh.src[MW * k + 2] = fm + 1;
      sh.src[MW * k + 3] = &c.lost[seq & 1];
    }
    SUBPROF_MARK(tEnd);
    if (!grid.sync(sh.src, MW * nobj, sh.snap)) {
      return generation;
    }
    SUBPROF_MARK(tSynced);
    SUBPROF_ADD(10, tEnd, tSynced);
    SUBPROF_ADD(11, 0, 1);
    SUBPROF_ADD(12, phaseStart, tEnd);
    if (threadIdx.x < static_cast<unsigned>(nobj)) {
      const int k = threadIdx.x;
      if (sh.mode[k] == MODE_SPLIT) {
        sh.thr[k] = sh.newThr[k];
        sh.epoch[k] += 1;
      }
    }
  }
  if (threadIdx.x < static_cast<unsigned>(nobj)) {
    const int k = threadIdx.x;
    if (t.block == 0) {
      atomicAdd(&ctls[k].iterations, sh.statIter[k]);
      atomicAdd(&ctls[k].epochs, sh.statEpoch[k]);
    }
    if (sh.statPush[k] != 0) {
      atomicAdd(&ctls[k].pushes, sh.statPush[k]);
    }
  }
  // The local queue overlays the near queue: leave that one empty.
  if (threadIdx.x < WAVES) {
    sh.near.count[threadIdx.x] = 0;
  }
  __syncthreads();
  return generation;
}

// Words -> (distance, parent); distance-only words get INT_MAX parents for
// the recovery pass (src/sospUpdateGpu.cu:478-497).
__device__ void unpack(Desc &D, int n, int source, Threads t) {
  const int bits = D.parentBits;
  const u64 *const packed = D.s.packed;
  int64_t *const distOut = D.distOut;
  int *const parentOut = D.parentOut;
  const u64 noParent = D.noParent;
  sweep4(n, t, [&](int v) { return packed[v]; }, [&](int v, u64 word) {
    if (bits > 0) {
      if (word == PACKED_INF) {
        distOut[v] = DISTANCE_INF;
        parentOut[v] = -1;
      } else {
        distOut[v] = static_cast<int64_t>(dist_of(word, bits));
        const u64 p = word & noParent;
        parentOut[v] = p == noParent ? -1 : static_cast<int>(p);
      }
    } else {
      distOut[v] = word == PACKED_INF ? DISTANCE_INF : static_cast<int64_t>(word);
      parentOut[v] = word == PACKED_INF || v == source ? -1 : INT_MAX;
    }
  });
}

// unpack for every objective at once (NOBJ = 2-4 objectives, all with
// packed parents), also writing each vertex's parents side by side for the
// masked combined graph: par4[v] = {parent in tree 0, 1, 2, 3} (-1 beyond
// NOBJ).
template <int NOBJ>
__device__ __forceinline__ void unpack_interleaved(Desc *objs, int n, Threads t, int4 *par4) {
  for (int v = t.id; v < n; v += t.count) {
    u64 w[NOBJ];
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
      w[k] = objs[k].s.packed[v];
    }
    int p[4] = {-1, -1, -1, -1};
#pragma unroll
    for (int k = 0; k < NOBJ; ++k) {
      Desc &D = objs[k];
      int64_t d = DISTANCE_INF;
      if (w[k] != PACKED_INF) {
        d = static_cast<int64_t>(dist_of(w[k], D.parentBits));
        const u64 q = w[k] & D.noParent;
        p[k] = q == D.noParent ? -1 : static_cast<int>(q);
      }
      D.distOut[v] = d;
      D.parentOut[v] = p[k];
    }
    par4[v] = make_int4(p[0], p[1], p[2], p[3]);
  }
}

// Edge-parallel sweep of this workgroup's even, contiguous share of the m
// = rowPtr[n] edges, BLOCK edges a pass: f(u, e) for edge e of row u.
// Balanced whatever the degrees (power-law generators put the hubs side by
// side).  Rows come from a window of BLOCK row pointers staged in LDS: a lane
// binary-searches its edge there, and the window is reloaded (one coalesced
// load, two barriers) only when the passes have used up its edges, so a pass
// costs one round of edge loads.  Only the share's first row is searched in
// global memory.
template <typename F>
__device__ void sweep_edges(const int *rowPtr, int n, Threads t, Shared &sh, F &&f) {
  const long long m = rowPtr[n];
  const long long lo = m * t.block / t.blocks;
  const long long hi = m * (t.block + 1) / t.blocks;
  if (lo >= hi) {
    return;
  }
  int next = row_of(rowPtr, 0, n - 1, lo);  // rowPtr[next] <= base: the next window
  int first = -1, rows = 0;                 // current window: rows [first, first + rows)
  long long end = 0;                        // ... and their edges [sh.scan[0], end)
  for (long long base = lo; base < hi;) {
    if (next != first) {
      __syncthreads();  // the last window's searches are done
      first = next;
      rows = n - first < static_cast<int>(BLOCK) ? n - first : static_cast<int>(BLOCK);
      for (int i = static_cast<int>(threadIdx.x); i < rows; i += BLOCK) {
        sh.scan[i] = rowPtr[first + i];
      }
      end = rowPtr[first + rows];
      __syncthreads();
    }
    long long stop = base + BLOCK < hi ? base + BLOCK : hi;
    stop = stop < end ? stop : end;
    const long long e = base + threadIdx.x;
    if (e < stop) {
      int a = 0, b = rows - 1;  // the last window row starting at or before e
      while (a < b) {
        const int mid = a + (b - a + 1) / 2;
        if (sh.scan[mid] <= e) {
          a = mid;
        } else {
          b = mid - 1;
        }
      }
      f(first + a, e);
    }
    if (stop == end) {
      next = first + rows;  // rowPtr[first + rows] == end == the next base
    }
    base = stop;
  }
}

// Lowest-id parent over tight out-edges for `count` distance-only
// objectives sharing one out-CSR (src/sospUpdateGpu.cu:490-511), edge-
// parallel (sweep_edges); one sweep serves every objective (row search,
// column read).
__device__ void recover_edges(Desc *objs, int count, int n, int source, Threads t,
                              Shared &sh) {
  sweep_edges(objs[0].outRow, n, t, sh, [&](int u, long long e) {
    const int w = out_edge(objs[0], e).x;
    if (w == source) {
      return;
    }
    for (int k = 0; k < count; ++k) {
      Desc &D = objs[k];
      const u64 du = D.s.packed[u];
      if (du != PACKED_INF && du + static_cast<u64>(out_edge(D, e).y) == D.s.packed[w]) {
        atomicMin(&D.parentOut[w], u);
      }
    }
  });
}

// unpack() of `count` distance-only objectives in one pass that also writes
// each vertex's K words side by side (stage[v * count + k]) for
// recover_staged: one record per vertex instead of K arrays.
__device__ void unpack_staged(Desc *objs, int count, int n, int source, Threads t,
                              u64 *stage) {
  for (int v = t.id; v < n; v += t.count) {
    for (int k = 0; k < count; ++k) {
      Desc &D = objs[k];
      const u64 word = D.s.packed[v];
      stage[static_cast<size_t>(v) * count + k] = word;
      D.distOut[v] = word == PACKED_INF ? DISTANCE_INF : static_cast<int64_t>(word);
      D.parentOut[v] = word == PACKED_INF || v == source ? -1 : INT_MAX;
    }
  }
}

// recover_edges over the staged records: a tight-edge test reads the head's
// K distances from one record (one or two cache lines) instead of K arrays
// (K lines); the sweep and the lowest-id rule are recover_edges'
// (src/sospUpdateGpu.cu:490-511).  N objectives, known at compile time, so
// every load of an edge is issued before the first test waits.
template <int N>
__device__ void recover_staged_n(Desc *objs, int n, int source, Threads t, Shared &sh,
                                 const u64 *stage) {
  const int *const outCol = objs[0].outCol;
  sweep_edges(objs[0].outRow, n, t, sh, [&](int u, long long e) {
    const int w = outCol[e];
    if (w == source) {
      return;
    }
    const u64 *const tail = stage + static_cast<size_t>(u) * N;
    const u64 *const head = stage + static_cast<size_t>(w) * N;
    u64 du[N], dw[N];
    int wt[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
      du[k] = tail[k];
      dw[k] = head[k];
      wt[k] = objs[k].outW[e];
    }
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (du[k] != PACKED_INF && du[k] + static_cast<u64>(wt[k]) == dw[k]) {
        atomicMin(&objs[k].parentOut[w], u);
      }
    }
  });
}

// count <= STAGE_MAX.
__device__ void recover_staged(Desc *objs, int count, int n, int source, Threads t,
                               Shared &sh, const u64 *stage) {
  switch (count) {
    case 1: recover_staged_n<1>(objs, n, source, t, sh, stage); break;
    case 2: recover_staged_n<2>(objs, n, source, t, sh, stage); break;
    case 3: recover_staged_n<3>(objs, n, source, t, sh, stage); break;
    case 4: recover_staged_n<4>(objs, n, source, t, sh, stage); break;
    case 5: recover_staged_n<5>(objs, n, source, t, sh, stage); break;
    case 6: recover_staged_n<6>(objs, n, source, t, sh, stage); break;
    case 7: recover_staged_n<7>(objs, n, source, t, sh, stage); break;
    default: recover_staged_n<STAGE_MAX>(objs, n, source, t, sh, stage); break;
  }
}

// Lowest-id parent over tight out-edges (src/sospUpdateGpu.cu:490-511).
template <bool HUBS>
__device__ void recover_parents(Desc &D, int n, int source,
                                Threads t, Shared &sh) {
  auto scan = [&](int u, u64 du, int x) {
    const int2 ew = out_edge(D, x);
    if (ew.x != source && du + static_cast<u64>(ew.y) == D.s.packed[ew.x]) {
      atomicMin(&D.parentOut[ew.x], u);
    }
  };
  if constexpr (!HUBS) {
    for (int u = t.id; u < n; u += t.count) {
      const u64 du = D.s.packed[u];
      if (du == PACKED_INF) {
        continue;
      }
      for (int x = D.outRow[u]; x < D.outRow[u + 1]; ++x) {
        scan(u, du, x);
      }
    }
    return;
  }
  recover_edges(&D, 1, n, source, t, sh);
}

// Source-only start of a from-scratch search (src/sospUpdateGpu.cu:222-231).
__device__ void scratch_init(Desc &D, ObjCtl &c, int n, int source,
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

// Pointer jumping over one word per vertex: bit 31 = "a root is among this
// vertex and its ancestors" (JUMP_FLAG), bits 0-30 = ancestor + 1 (0: none).
// Invariant for every x: a flagged word implies a root among x and its
// ancestors; an unflagged one, no root on the tree path from x up to
// (excluding) its ancestor.  Only v's thread writes word[v] and every word
// another thread can observe keeps the invariant, so the update is in place;
// one 32-bit word carries both fields, so a read sees a consistent pair.
constexpr int JUMP_FLAG = INT_MIN;

// Jumps per round and vertex: a later word of an ancestor is as valid to
// follow as an earlier one (every vertex between v and the ancestor its word
// names is known not to be a root), so each round follows up to JUMPS words
// instead of one and fewer rounds (grid barriers) are needed.  On a
// 1.96 M-vertex road grid with 3 objectives and 50,000 changes: 0.468 /
// 0.359 / 0.328 ms for the whole invalidation at 1 / 2 / 3 jumps.
constexpr int JUMPS = 3;

// One round for vertex v: every root carries its flag before round 0, so
// each ancestor's word alone decides.  Round 0 takes every vertex (a root
// is done at once); later rounds the vertices still jumping (unflagged,
// with an ancestor).  Returns true if v still jumps.
__device__ __forceinline__ bool jump_next(int *word, int v) {
  const int w0 = word[v];
  if (w0 & JUMP_FLAG) {
    return false;
  }
  int w = w0;
  for (int j = 0; j < JUMPS && w != 0; ++j) {
    const int wx = word[w - 1];
    if (wx & JUMP_FLAG) {
      word[v] = w0 | JUMP_FLAG;
      return false;
    }
    w = wx;
  }
  word[v] = w;
  return w != 0;
}

// Inlined into its instance: as a call (its own register allocation and
// frame) the update was ~10 % slower.
template <bool HUBS>
__device__ __forceinline__ int update_step1(Args &a, GridBarrier &grid, Threads t,
                            int generation, Shared &sh) {
  const int n = a.n;
  Ctl *ctl = a.ctl;
  // Pack the old trees; ancestors; roots of changed tree edges.
  bool anyChanged = false;
  for (int k = 0; k < a.nobj; ++k) {
    Desc &D = desc(a.objs)[k];
    anyChanged = anyChanged || D.nChanged > 0;
    unsigned overflow = 0;
    const int64_t *const distIn = D.distIn;
    const int *const parentIn = D.parentIn;
    struct Old {
      int64_t d;
      int p;
    };
    sweep4(n, t, [&](int v) { return Old{distIn[v], parentIn[v]}; }, [&](int v, Old old) {
      const int64_t d = old.d;
      int p = old.p;
      if (p < -1 || p >= n) {
        overflow |= 2;  // not a vertex: reported, and never used as an index
        p = -1;
      }
      u64 word;
      if (d >= DISTANCE_INF / 2) {
        word = PACKED_INF;
      } else if (d >= 0 && static_cast<u64>(d) > D.maxDistance && D.optimistic) {
        store_relaxed(&ctl->obj[k].dropped, 1u);  // redone with a wider packing
        word = PACKED_INF;
      } else if (d < 0 || static_cast<u64>(d) > D.maxDistance) {
        overflow |= 1;
        word = PACKED_INF;
      } else {
        word = pack(static_cast<u64>(d), p, D.parentBits, D.noParent);
      }
      D.s.packed[v] = word;
      if (D.nChanged > 0) {
        D.s.ancestor[v] = p + 1;  // jump word: ancestor + 1, no flag
      }
    });
    if (overflow) {
      atomicOr(&ctl->obj[k].overflow, overflow);
    }
  }
  if (!grid.sync()) {
    return generation;
  }
  // The roots (heads of changed tree edges) carry their flag in their jump
  // word before any round reads it: the rounds then read words only.
  if (anyChanged) {
    for (int k = 0; k < a.nobj; ++k) {
      Desc &D = desc(a.objs)[k];
      for (int i = t.id; i < D.nChanged; i += t.count) {
        const int v = D.chTo[i];
        const int p = D.parentIn[v];
        if (p == D.chFrom[i]) {
          D.s.ancestor[v] = (p + 1) | JUMP_FLAG;
        }
      }
    }
    if (!grid.sync()) {
      return generation;
    }
  }
  stamp(ctl, PHASE_PACK, t.id);

  // Subtree invalidation by pointer jumping.  Round 0 sweeps every vertex;
  // the vertices still jumping are compacted into a list (near buffers are
  // free before Step 2) that the next round walks, until no list is left.
  if (anyChanged) {
    for (int round = 0; round < a.maxRounds; ++round) {
      if (t.id == 0) {
        for (int k = 0; k < a.nobj; ++k) {
          store_relaxed(&ctl->obj[k].jumpCount[(round + 2) % 3], 0u);
        }
        ctl->rounds = round + 1;
      }
      for (int k = 0; k < a.nobj; ++k) {
        Desc &D = desc(a.objs)[k];
        if (D.nChanged == 0) {
          continue;
        }
        int *anc = D.s.ancestor;
        int *outList = D.s.near[(round + 1) & 1];
        unsigned *outCount = &ctl->obj[k].jumpCount[(round + 1) % 3];
        const unsigned cnt = round == 0 ? static_cast<unsigned>(n) : sh.snap[k];
        const int *inList = D.s.near[round & 1];
        for (unsigned base = t.block * BLOCK; base < cnt; base += t.count) {
          const unsigned i = base + threadIdx.x;
          if (i < cnt) {
            const int v = round == 0 ? static_cast<int>(i) : inList[i];
            if (jump_next(anc, v)) {
              queue_push(sh.near, v, outList, outCount, static_cast<unsigned>(n));
            }
          }
          __syncthreads();
          if (sh.near.count[0] > QCAP - BLOCK) queue_flush(sh.near, outList, outCount, static_cast<unsigned>(n));
        }
        queue_flush(sh.near, outList, outCount, static_cast<unsigned>(n));
      }
      if (threadIdx.x < static_cast<unsigned>(a.nobj)) {
        sh.src[threadIdx.x] = &ctl->obj[threadIdx.x].jumpCount[(round + 1) % 3];
      }
      if (!grid.sync(sh.src, a.nobj, sh.snap)) {
        return generation;
      }
      unsigned left = 0;
      for (int k = 0; k < a.nobj; ++k) {
        left += sh.snap[k];
      }
      if (left == 0) {
        break;
      }
    }
  }
  stamp(ctl, PHASE_JUMP, t.id);

  // Invalidate flagged vertices (INF); mark insert heads (stamp).  The jump
  // words keep the flags until the next update's pack pass rewrites them.
  const int genInv = ++generation;
  for (int k = 0; k < a.nobj; ++k) {
    Desc &D = desc(a.objs)[k];
    unsigned count = 0;
    if (D.nChanged > 0) {
      const int *const anc = D.s.ancestor;
      u64 *const packed = D.s.packed;
      sweep4(n, t, [&](int v) { return anc[v]; }, [&](int v, int w) {
        if (w & JUMP_FLAG) {
          packed[v] = PACKED_INF;
          ++count;
        }
      });
    }
    for (int i = t.id; i < D.nHeads; i += t.count) {
      const int v = D.heads[i];
      if (v != a.source) {
        D.s.stamp[v] = genInv;
      }
    }
    count = block_sum(count, sh.red);
    if (threadIdx.x == 0 && count > 0) {
      atomicAdd(&ctl->obj[k].invalidated, count);
    }
  }
  if (threadIdx.x < static_cast<unsigned>(a.nobj)) {
    sh.src[threadIdx.x] = &ctl->obj[threadIdx.x].invalidated;
  }
  if (!grid.sync(sh.src, a.nobj, sh.snap)) {
    return generation;
  }
  stamp(ctl, PHASE_INVALIDATE, t.id);

  // Pull: every invalidated vertex and insert head takes the best (distance,
  // parent) over its in-neighbours; improved ones start in the far pile of
  // epoch 0 (the first near-far iteration splits it at min + delta).  In the
  // hub variant, candidates with more than PULL_BLOCK in-edges are listed
  // (near[0]) and pulled afterwards by one whole workgroup each, round robin
  // over the grid: the heaviest ones would otherwise hold their wave -- and
  // the whole grid at the next barrier -- for milliseconds.
  auto pulled = [&](Desc &D, ObjCtl &c, int v, u64 best, u64 current) {
    if (best < current) {
      const u64 prev = atomicMin(&D.s.packed[v], best);
      const u64 bd = dist_of(best, D.parentBits);
      if (bd < dist_of(prev, D.parentBits)) {
        D.s.inFar[v] = 1;
        queue_push(sh.far, v, D.s.far[0], &c.farCount[0], static_cast<unsigned>(n));
        atomicMin(&sh.farMin, bd);
      }
    }
  };
  const unsigned listCap = static_cast<unsigned>(a.listCap > 0 && a.listCap < n ? a.listCap : n);
  for (int k = 0; k < a.nobj; ++k) {
    Desc &D = desc(a.objs)[k];
    ObjCtl &c = ctl->obj[k];
    const int bits = D.parentBits;
    if (threadIdx.x == 0) {
      sh.farMin = PACKED_INF;
    }
    __syncthreads();
    // Most vertices invalidated (graphs without hubs): the valid vertices
    // offer their candidates to their invalidated out-neighbours instead,
    // the same (valid in-neighbour, invalidated vertex) pairs as the pull,
    // so the same minima; insert heads still pull over all their in-edges.
    // Improved vertices get their far mark, and the far count is raised past
    // the list capacity so that epoch 0's split reads the pile from the far
    // marks (no list appends here).
    const bool push = !HUBS && D.nChanged > 0 && a.initPush != 1 &&
                      (a.initPush == 2 || 2ull * sh.snap[k] > static_cast<u64>(n));
    if (push) {
      int *const anc = D.s.ancestor;
      u64 *const packed = D.s.packed;
      auto offer = [&](int v, u64 cand) {
        const u64 prev = atomicMin(&packed[v], cand);
        const u64 cd = dist_of(cand, bits);
        if (cd < dist_of(prev, bits)) {
          D.s.inFar[v] = 1;
          atomicMin(&sh.farMin, cd);
        }
      };
      sweep4(n, t, [&](int u) { return anc[u]; }, [&](int u, int w) {
        if (w & JUMP_FLAG) {
          return;
        }
        const u64 word = packed[u];
        if (word == PACKED_INF) {
          return;
        }
        const u64 du = dist_of(word, bits);
        const int end = D.outRow[u + 1];
        for (int x = D.outRow[u]; x < end; ++x) {
          const int2 ew = out_edge(D, x);
          const int v = ew.x;
          if (v == a.source || !(anc[v] & JUMP_FLAG)) {
            continue;
          }
          const u64 nd = du + static_cast<u64>(ew.y);
          if (nd > D.maxDistance) {
            if (D.optimistic) {
              store_relaxed(&c.dropped, 1u);
            }
            continue;
          }
          offer(v, pack(nd, u, bits, D.noParent));
        }
      });
      for (int i = t.id; i < D.nHeads; i += t.count) {
        const int v = D.heads[i];
        if (v == a.source) {
          continue;
        }
        const u64 current = packed[v];
        u64 best = current;
        const int end = D.inRow[v + 1];
        for (int x = D.inRow[v]; x < end; ++x) {
          const int2 ew = in_edge(D, x);
          const u64 word = packed[ew.x];
          if (word == PACKED_INF) {
            continue;
          }
          const u64 nd = dist_of(word, bits) + static_cast<u64>(ew.y);
          if (nd > D.maxDistance) {
            if (D.optimistic) {
              store_relaxed(&c.dropped, 1u);
            }
            continue;
          }
          const u64 cand = pack(nd, ew.x, bits, D.noParent);
          best = cand < best ? cand : best;
        }
        if (best < current) {
          offer(v, best);
        }
      }
      __syncthreads();
      if (threadIdx.x == 0) {
        if (sh.farMin != PACKED_INF) {
          atomicMin(&c.farMin[0], sh.farMin);
        }
        if (t.block == 0) {
          atomicAdd(&c.farCount[0], listCap + 1);
        }
      }
      continue;
    }
    auto candidate = [&](int x) -> u64 {
      const int2 ew = in_edge(D, x);
      const int u = ew.x;
      const u64 word = D.s.packed[u];
      if (word == PACKED_INF) {
        return PACKED_INF;
      }
      const u64 nd = dist_of(word, bits) + static_cast<u64>(ew.y);
      // Larger than maxDistance: never shortest (unless the packing is
      // optimistic: the launch is redone), may not fit the word.
      if (nd > D.maxDistance) {
        if (D.optimistic) {
          store_relaxed(&c.dropped, 1u);
        }
        return PACKED_INF;
      }
      return pack(nd, u, bits, D.noParent);
    };
    for (unsigned base = t.block * BLOCK; base < static_cast<unsigned>(n);
         base += t.count) {
      const int v = static_cast<int>(base + threadIdx.x);
      bool want = v < n && v != a.source &&
                  ((D.nChanged > 0 && (D.s.ancestor[v] & JUMP_FLAG)) || D.s.stamp[v] == genInv);
      u64 current = PACKED_INF;
      int begin = 0, end = 0;
      if (want) {
        current = D.s.packed[v];
        begin = D.inRow[v];
        end = D.inRow[v + 1];
      }
      u64 best = current;
      if constexpr (!HUBS) {
        for (int x = begin; x < end; ++x) {
          const u64 cand = candidate(x);
          best = cand < best ? cand : best;
        }
      } else {
        // Hub graphs: the sweep only lists the wanted vertices -- more than
        // PULL_BLOCK in-edges: near[0], for whole workgroups; the others:
        // near[1] (through the far queue, unused here), for lane groups
        // over the whole grid (a vertex's pull is a chain of dependent
        // reads: one lane per vertex, or a wave per vertex in turn, kept
        // the grid waiting for its slowest wave).
        (void)best;
        if (want) {
          if (end - begin > PULL_BLOCK) {
            queue_push(sh.near, v, D.s.near[0], &c.pullHeavy, static_cast<unsigned>(n));
          } else {
            queue_push(sh.far, v, D.s.near[1], &c.pullCount, static_cast<unsigned>(n));
          }
          want = false;
        }
      }
      if (want) {
        pulled(D, c, v, best, current);
      }
      __syncthreads();
      if constexpr (HUBS) {
        if (sh.far.count[0] > QCAP - BLOCK) queue_flush(sh.far, D.s.near[1], &c.pullCount, static_cast<unsigned>(n));
        if (sh.near.count[0] > QCAP - BLOCK) queue_flush(sh.near, D.s.near[0], &c.pullHeavy, static_cast<unsigned>(n));
      } else {
        if (sh.far.count[0] > QCAP - BLOCK) queue_flush(sh.far, D.s.far[0], &c.farCount[0], static_cast<unsigned>(n));
      }
    }
    if constexpr (HUBS) {
      queue_flush(sh.far, D.s.near[1], &c.pullCount, static_cast<unsigned>(n));
      queue_flush(sh.near, D.s.near[0], &c.pullHeavy, static_cast<unsigned>(n));
    } else {
      queue_flush(sh.far, D.s.far[0], &c.farCount[0], static_cast<unsigned>(n));
    }
    if (threadIdx.x == 0 && sh.farMin != PACKED_INF) {
      atomicMin(&c.farMin[0], sh.farMin);
    }
  }
  if constexpr (HUBS) {
    if (threadIdx.x < static_cast<unsigned>(a.nobj)) {
      sh.src[threadIdx.x] = &ctl->obj[threadIdx.x].pullHeavy;
      sh.src[a.nobj + threadIdx.x] = &ctl->obj[threadIdx.x].pullCount;
    }
    if (!grid.sync(sh.src, 2 * a.nobj, sh.snap)) {
      return generation;
    }
    // The listed vertices: lane groups of PULL_GROUP over the whole grid,
    // each group one vertex at a time, striding over its in-edges.
    constexpr unsigned PULL_GROUP = 16;
    const unsigned gsub = threadIdx.x & (PULL_GROUP - 1);
    const unsigned gid = t.id / PULL_GROUP;
    const unsigned groups = t.count / PULL_GROUP;
    for (int k = 0; k < a.nobj; ++k) {
      Desc &D = desc(a.objs)[k];
      ObjCtl &c = ctl->obj[k];
      const int bits = D.parentBits;
      if (threadIdx.x == 0) {
        sh.farMin = PACKED_INF;
      }
      __syncthreads();
      const unsigned cnt = min(sh.snap[a.nobj + k], static_cast<unsigned>(n));
      for (unsigned base = 0; base < cnt; base += groups) {
        const unsigned i = base + gid;
        int v = -1;
        u64 current = PACKED_INF, part = PACKED_INF;
        if (i < cnt) {
          v = D.s.near[1][i];
          if (v < 0 || v >= n) {
            v = -1;  // not a vertex: never used as an index
          } else {
            current = D.s.packed[v];
            const int end = D.inRow[v + 1];
            for (int x = D.inRow[v] + static_cast<int>(gsub); x < end;
                 x += static_cast<int>(PULL_GROUP)) {
              const int2 ew = in_edge(D, x);
              const u64 word = D.s.packed[ew.x];
              if (word != PACKED_INF) {
                const u64 nd = dist_of(word, bits) + static_cast<u64>(ew.y);
                if (nd <= D.maxDistance) {
                  const u64 cand = pack(nd, ew.x, bits, D.noParent);
                  part = cand < part ? cand : part;
                } else if (D.optimistic) {
                  store_relaxed(&c.dropped, 1u);
                }
              }
            }
          }
        }
        // Group minimum (every lane takes part).
#pragma unroll
        for (unsigned off = PULL_GROUP / 2; off > 0; off >>= 1) {
          const u64 o = shfl_u64(part, static_cast<int>(__lane_id() ^ off));
          part = o < part ? o : part;
        }
        if (gsub == 0 && v >= 0) {
          pulled(D, c, v, part < current ? part : current, current);
        }
        __syncthreads();
        if (sh.far.count[0] > QCAP - BLOCK) queue_flush(sh.far, D.s.far[0], &c.farCount[0], static_cast<unsigned>(n));
      }
      queue_flush(sh.far, D.s.far[0], &c.farCount[0], static_cast<unsigned>(n));
      if (threadIdx.x == 0 && sh.farMin != PACKED_INF) {
        atomicMin(&c.farMin[0], sh.farMin);
      }
    }
#ifdef MOSP_SUBPROFILE
    stamp(ctl, PHASE_COUNT, t.id);  // the pull's sweep is over; the heavy candidates follow
    if (t.id == 0) {
      unsigned heavy = 0;
      for (int k = 0; k < a.nobj; ++k) {
        heavy += sh.snap[k];
      }
      ctl->clk[PHASE_COUNT + 1] = heavy;
    }
#endif
    for (int k = 0; k < a.nobj; ++k) {
      Desc &D = desc(a.objs)[k];
      ObjCtl &c = ctl->obj[k];
      const int bits = D.parentBits;
      if (threadIdx.x == 0) {
        sh.farMin = PACKED_INF;
      }
      __syncthreads();
      const unsigned heavyCount = sh.snap[k];
      for (unsigned i = t.block; i < heavyCount; i += t.blocks) {
        const int v = D.s.near[0][i];
        const u64 current = D.s.packed[v];
        u64 part = PACKED_INF;
        for (int x = D.inRow[v] + static_cast<int>(threadIdx.x); x < D.inRow[v + 1]; x += BLOCK) {
          const int2 ew = in_edge(D, x);
          const int u = ew.x;
          const u64 word = D.s.packed[u];
          if (word != PACKED_INF) {
            const u64 nd = dist_of(word, bits) + static_cast<u64>(ew.y);
            if (nd <= D.maxDistance) {
              const u64 cand = pack(nd, u, bits, D.noParent);
              part = cand < part ? cand : part;
            } else if (D.optimistic) {
              store_relaxed(&c.dropped, 1u);
            }
          }
        }
        const u64 best = block_min_u64(part, sh.redMin);
        if (threadIdx.x == 0) {
          pulled(D, c, v, best < current ? best : current, current);
        }
      }
      queue_flush(sh.far, D.s.far[0], &c.farCount[0], static_cast<unsigned>(n));
      if (threadIdx.x == 0 && sh.farMin != PACKED_INF) {
        atomicMin(&c.farMin[0], sh.farMin);
      }
    }
  }
  grid.sync();  // the caller checks grid.aborted
  stamp(ctl, PHASE_PULL, t.id);
  return generation;
}

// ---------------------------------------------------------------------------
// Combined graph of cK trees (src/combinedGraphGpu.cu:55-121) + its CSR.
// ---------------------------------------------------------------------------

// slot != nullptr: a parent that takes the global path gets its place in
// the parent's row written there.
__device__ __forceinline__ void agg_add(int *keys, int *counts, int *global, int p,
                                        int *slot = nullptr) {
  unsigned h = (static_cast<unsigned>(p) * 2654435761u) >> (32 - AGG_BITS);
  for (int probe = 0; probe < AGG_PROBES; ++probe) {
    const int key = keys[h];  // a slot's key only ever goes from -1 to a parent
    if (key == p) {
      atomicAdd(&counts[h], 1);
      return;
    }
    if (key == -1) {
      const int prev = atomicCAS(&keys[h], -1, p);
      if (prev == -1 || prev == p) {
        atomicAdd(&counts[h], 1);
        return;
      }
    }
    h = (h + 1) & (AGG_SLOTS - 1);
  }
  const int at = atomicAdd(&global[p], 1);
  if (slot != nullptr) {
    *slot = at;
  }
}

// Slot of key p in the table agg_add filled, or -1 (p took the global
// path).  Insertion takes the first free slot of p's probe sequence, so an
// empty slot before p means p is not in the table.
__device__ __forceinline__ int agg_find(const int *keys, int p) {
  unsigned h = (static_cast<unsigned>(p) * 2654435761u) >> (32 - AGG_BITS);
  for (int probe = 0; probe < AGG_PROBES; ++probe) {
    const int key = keys[h];
    if (key == p) {
      return static_cast<int>(h);
    }
    if (key == -1) {
      return -1;
    }
    h = (h + 1) & (AGG_SLOTS - 1);
  }
  return -1;
}

// Parent k of v if it is the first occurrence of its value among the K
// parents; its weight is base - sum of terms over the trees containing it.
__device__ __forceinline__ bool combined_edge(Args &a, int v, int k,
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

constexpr int COMB_REG = 8;  // trees whose parents a lane holds in registers

// f(p, weight, k) for every combined edge (p, v), k the first tree where p is
// v's parent: v's parents are read once
// (K <= COMB_REG: into registers, every test from them), in tree order.
template <typename F>
__device__ __forceinline__ void combined_edges(Args &a, int v, F &&f) {
  const int K = a.cK;
  if (K > COMB_REG) {
    for (int k = 0; k < K; ++k) {
      int p, weight;
      if (combined_edge(a, v, k, p, weight)) {
        f(p, weight, k);
      }
    }
    return;
  }
  const int *par = a.cParents;
  const size_t n = static_cast<size_t>(a.n);
  int pv[COMB_REG];
#pragma unroll
  for (int k = 0; k < COMB_REG; ++k) {
    pv[k] = k < K ? par[k * n + v] : -1;
  }
#pragma unroll
  for (int k = 0; k < COMB_REG; ++k) {
    if (k >= K || pv[k] < 0) {
      continue;
    }
    bool first = true;
#pragma unroll
    for (int j = 0; j < k; ++j) {
      first = first && pv[j] != pv[k];
    }
    if (!first) {
      continue;
    }
    int weight = a.cBase - a.cTerms[k];
#pragma unroll
    for (int j = k + 1; j < COMB_REG; ++j) {
      if (j < K && pv[j] == pv[k]) {
        weight -= a.cTerms[j];
      }
    }
    f(pv[k], weight, k);
  }
}

__device__ __forceinline__ void combined_delta(Args &a, Ctl *ctl, Threads t);

// The unpack writes the side-by-side parents (LaunchArgs::cPar4) and the
// mask pass reads them.
__device__ __forceinline__ bool par4_written(Args &a) {
  return a.cPar4 != nullptr && a.cMaskRec != nullptr && !a.anyFallback &&
         (a.phases & PH_COMBINED) && a.nobj == a.cK && a.cK >= 2 && a.cK <= 4;
}

// Sums of the combined edges and weights a thread counted, into the launch's
// totals (one atomic per wave).
__device__ __forceinline__ void combined_totals(Ctl *ctl, u64 edges, u64 wsum) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) {
    edges += __shfl_xor(edges, offset);
    wsum += __shfl_xor(wsum, offset);
  }
  if (__lane_id() == 0 && edges > 0) {
    atomicAdd(&ctl->combEdges, edges);
    atomicAdd(&ctl->combWeightSum, wsum);
  }
}

// The combined graph as a mask over the graph's out-edges (cMaskRec): one
// pass over the rows, no counters, scan or per-edge atomics.  Same edges,
// weights, edge count and weight sum as combined_build.
__device__ __attribute__((noinline)) void combined_masked(Args &args, GridBarrier &grid,
                                                          Threads t) {
  Args &a = *uniform(&args);
  const int n = a.n;
  const int K = a.cK;
  const int source = a.source;
  Ctl *ctl = a.ctl;
  const int *const row = uniform_global(a.cMaskRow);
  const int4 *const src = uniform_global(a.cMaskSrc);
  int4 *const dst = uniform_global(a.cMaskRec);
  const int *const par = uniform_global(a.cParents);
  const int4 *const par4 = par4_written(a) ? uniform_global(a.cPar4) : nullptr;
  const size_t nn = static_cast<size_t>(n);
  // A row is taken MU edges at a time: their records, then every parent
  // word they need, are requested together (two round trips per chunk).
  constexpr int MU = 4;
  u64 edges = 0, wsum = 0;
  for (int p = t.id; p < n; p += t.count) {
    const int b = row[p], e = row[p + 1];
    for (int x0 = b; x0 < e; x0 += MU) {
      int4 r[MU];
#pragma unroll
      for (int j = 0; j < MU; ++j) {
        r[j] = x0 + j < e ? src[x0 + j] : make_int4(source, 0, 0, 0);
      }
      int w[MU];
#pragma unroll
      for (int j = 0; j < MU; ++j) {
        w[j] = a.cBase;
      }
      if (par4 != nullptr) {
        // One 16-byte record per edge holds the K parents of its head.
        int4 q[MU];
#pragma unroll
        for (int j = 0; j < MU; ++j) {
          q[j] = r[j].x != source ? par4[r[j].x] : make_int4(-1, -1, -1, -1);
        }
#pragma unroll
        for (int j = 0; j < MU; ++j) {
          w[j] -= (q[j].x == p ? a.cTerms[0] : 0) + (q[j].y == p ? a.cTerms[1] : 0) +
                  (K > 2 && q[j].z == p ? a.cTerms[2] : 0) +
                  (K > 3 && q[j].w == p ? a.cTerms[3] : 0);
        }
      } else {
        for (int k = 0; k < K; ++k) {
          const int *const pk = par + k * nn;
          const int term = a.cTerms[k];
#pragma unroll
          for (int j = 0; j < MU; ++j) {
            if (r[j].x != source && pk[r[j].x] == p) {
              w[j] -= term;
            }
          }
        }
      }
#pragma unroll
      for (int j = 0; j < MU; ++j) {
        if (x0 + j >= e) {
          break;
        }
        // A combined edge if some tree's parent of v is p; a repeated (p, v)
        // edge of the row after the first is masked.
        bool keep = w[j] != a.cBase;
#pragma unroll
        for (int i = 0; i < j; ++i) {
          keep = keep && r[i].x != r[j].x;
        }
        for (int y = b; y < x0 && keep; ++y) {
          keep = src[y].x != r[j].x;
        }
        if (keep) {
          dst[x0 + j] = make_int4(r[j].x, w[j], r[j].z, r[j].w);
          edges += 1;
          wsum += static_cast<u64>(w[j]);
        } else {
          dst[x0 + j] = make_int4(source, 0, 0, 0);
        }
      }
    }
  }
  combined_totals(ctl, edges, wsum);
  if (!grid.sync()) {
    return;
  }
  stamp(ctl, PHASE_COMB_COUNT, t.id);
  stamp(ctl, PHASE_COMB_SCAN, t.id);
  combined_delta(a, ctl, t);
}

// A call, not inlined: it then does not raise the register use of the
// near-far loops around it.
template <bool HUBS>
__device__ __attribute__((noinline)) void combined_build(Args &args, GridBarrier &grid,
                               Threads t, Shared &sh) {
  Args &a = *uniform(&args);
  int *sScan = sh.scan;
  const int n = a.n;
  Ctl *ctl = a.ctl;
  // Clear the degree counters (n + 1).
  for (int i = t.id; i <= n; i += t.count) {
    a.cDegree[i] = 0;
  }
  if (!grid.sync()) {
    return;
  }
  // Count: out-degree of every parent; edge count and weight sum.
  if constexpr (HUBS) {
    for (int i = threadIdx.x; i < AGG_SLOTS; i += BLOCK) {
      sh.aggKey[i] = -1;
      sh.aggCount[i] = 0;
    }
    __syncthreads();
  }
  // With cSlot (hub variant), every increment's return value is kept: a
  // direct one's is the edge's place in its parent's row (cSlot), a
  // workgroup's aggregated one the base of its run there (aggCount), so the
  // fill needs no atomics on global counters.  Graphs without hubs build
  // this CSR only in batched launches or without the mask, where it did not
  // pay, and their variants keep their smaller frames.
  int *const slotOf = HUBS ? a.cSlot : nullptr;
  const size_t nn = static_cast<size_t>(n);
  u64 edges = 0, wsum = 0;
  for (int v = t.id; v < n; v += t.count) {
    if (v == a.source) {
      continue;
    }
    combined_edges(a, v, [&](int p, int weight, int k) {
      int *global = slotOf != nullptr ? &slotOf[k * nn + v] : nullptr;
      if constexpr (HUBS) {
        agg_add(sh.aggKey, sh.aggCount, a.cDegree, p, global);
      } else {
        const int at = atomicAdd(&a.cDegree[p], 1);
        if (global != nullptr) {
          *global = at;
        }
      }
      edges += 1;
      wsum += static_cast<u64>(weight);
    });
  }
  if constexpr (HUBS) {
    __syncthreads();
    for (int i = threadIdx.x; i < AGG_SLOTS; i += BLOCK) {
      if (sh.aggCount[i] > 0) {
        const int at = atomicAdd(&a.cDegree[sh.aggKey[i]], sh.aggCount[i]);
        if (slotOf != nullptr) {
          sh.aggCount[i] = at;  // the run's base; the fill counts from it
        }
      }
    }
  }
  combined_totals(ctl, edges, wsum);
  if (!grid.sync()) {
    return;
  }
  stamp(ctl, PHASE_COMB_COUNT, t.id);

  // Exclusive scan of the n + 1 degrees into row pointers: every workgroup
  // owns one contiguous chunk; chunk totals, then offsets, then local scans.
  const int total = n + 1;
  const int chunk = (total + static_cast<int>(t.blocks) - 1) / static_cast<int>(t.blocks);
  const int lo = min(total, static_cast<int>(t.block) * chunk);
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
    a.cPartials[t.block] = chunkTotal;
  }
  if (!grid.sync()) {
    return;
  }
  int offset = 0;
  for (unsigned b = threadIdx.x; b < t.block; b += BLOCK) {
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
  if (!grid.sync()) {
    return;
  }
  stamp(ctl, PHASE_COMB_SCAN, t.id);

  // Fill (row order is irrelevant: Step 3 breaks ties by the lowest id).
  // Hub variant: the workgroup visits the same vertices as in the count, so
  // its LDS table still holds its edge count per parent; one atomic per
  // parent reserves the workgroup's run of the parent's row (hub parents
  // have hundreds of thousands of children), and LDS atomics hand out the
  // run's places.  Parents that took the global path count there.
  if constexpr (HUBS) {
    if (slotOf == nullptr) {
      for (int i = threadIdx.x; i < AGG_SLOTS; i += BLOCK) {
        if (sh.aggCount[i] > 0) {
          sh.aggCount[i] = atomicAdd(&a.cDegree[sh.aggKey[i]], sh.aggCount[i]);
        }
      }
    }
    __syncthreads();
  }
  // Records with the child's row range as well, for the asynchronous loop.
  int4 *const rec4 = a.async != nullptr ? a.cEdge4 : nullptr;
  for (int v = t.id; v < n; v += t.count) {
    if (v == a.source) {
      continue;
    }
    int vBegin = 0, vEnd = 0;
    if (rec4 != nullptr) {
      vBegin = a.cRowPtr[v];
      vEnd = a.cRowPtr[v + 1];
    }
    combined_edges(a, v, [&](int p, int weight, int k) {
      int pos;
      if (slotOf != nullptr) {
        // The place the count took: the workgroup's run (hub parents in its
        // table) or the edge's own.
        int at;
        if constexpr (HUBS) {
          const int slot = agg_find(sh.aggKey, p);
          at = slot >= 0 ? atomicAdd(&sh.aggCount[slot], 1) : slotOf[k * nn + v];
        } else {
          at = slotOf[k * nn + v];
        }
        pos = a.cRowPtr[p] + at;
      } else if constexpr (HUBS) {
        const int slot = agg_find(sh.aggKey, p);
        pos = slot >= 0 ? atomicAdd(&sh.aggCount[slot], 1) : atomicAdd(&a.cDegree[p], 1);
      } else {
        pos = atomicAdd(&a.cDegree[p], 1);
      }
      a.cEdge[pos] = make_int2(v, weight);
      if (rec4 != nullptr) {
        rec4[pos] = make_int4(v, weight, vBegin, vEnd);
      }
    });
  }
  combined_delta(a, ctl, t);
}

// Near-far width for Step 3: the caller's, or 32 * average weight / average
// out-degree (src/sospUpdateGpu.cu:607-616), in double as there, times the
// launch's scale (a tuning factor; outputs do not depend on it).  Needs the
// edge count and weight sum complete.
__device__ __forceinline__ void combined_delta(Args &a, Ctl *ctl, Threads t) {
  const int n = a.n;
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
        delta = (d > 1 ? d : 1) * (a.cDeltaScale > 1 ? a.cDeltaScale : 1);
      }
    }
    store_relaxed(&ctl->combDelta, static_cast<u64>(delta));
  }
}

// ---------------------------------------------------------------------------
// The kernel
// ---------------------------------------------------------------------------

// One update instance (graph, K objectives, combined step) on the group of
// workgroups `t` describes; the group synchronizes through a.ctl.
// LOCAL: 0 no local hops (hub variants); 1 local hops, single-update
// kernel; 2 local hops, batched kernel.  The two values with hops compile
// the same code as separate instantiations down the call tree, so neither
// kernel's register allocation depends on the other's.
// Optimistic packing: whether some objective of the first `count` met a
// distance past its packed words' limit.  Read after the last barrier of the
// search (every drop happened before it), so every workgroup agrees; then
// the launch ends without writing results, clearing the root marks that
// Step 1 set (the host redoes it with a wider packing).
__device__ bool dropped_any(Args &a, int count, Threads t) {
  if (!a.anyOptimistic) {
    return false;
  }
  bool dropped = false;
  for (int k = 0; k < count; ++k) {
    dropped = dropped || load_relaxed(&a.ctl->obj[k].dropped) != 0;
  }
  return dropped;
}

// Every phase of one launch (instance); `generation` is the stamp
// generation, advanced as the phases use it.  Returns early when a barrier
// gives up or an optimistic packing meets a larger distance.
template <bool HUBS, int LOCAL>
__device__ __forceinline__ int run_phases(Args &a, GridBarrier &grid, const Threads t, Shared &sh,
                          int generation) {
  const int n = a.n;
  if (t.id == 0) {
    a.ctl->clkStart = wall_clock64();
  }
  if (threadIdx.x < WAVES) {
    sh.near.count[threadIdx.x] = 0;
    sh.far.count[threadIdx.x] = 0;
  }
  if (threadIdx.x == 0) {
    sh.wgCount = 0;
    sh.hubEdges = a.hubEdges;
    sh.queueCap = a.queueCap >= 256 && a.queueCap <= QCAP ? a.queueCap : QCAP;
  }
#ifdef MOSP_SUBPROFILE
  if (threadIdx.x < 16) subprof_lds()[threadIdx.x] = 0;
#endif
  // Asynchronous near-far: every launch starts its counters afresh.  Each
  // phase is preceded by a grid barrier, which orders these stores before
  // any use.
  constexpr bool kAsync = !HUBS && LOCAL != 0;
  // The batched kernel's light instances run the asynchronous loop only (the
  // iteration-synchronous one is not compiled into them: both together
  // spilled 640 bytes per lane).
  constexpr bool kAsyncOnly = !HUBS && LOCAL == 2;
  unsigned phaseSeq = 0;
  if constexpr (kAsync) {
    if (a.async != nullptr) {
      AsyncCtl *const A = a.async;
      if (threadIdx.x == 0) {
        store_relaxed(&A->status[t.block], ASYNC_STATUS_INIT);
        sh.as.sent = 0;
        sh.as.recv = 0;
      }
      if (t.block == 0) {
        for (unsigned i = threadIdx.x; i < static_cast<unsigned>(MOSP_MAX_OBJ) * ASYNC_SEGS;
             i += BLOCK) {
          store_relaxed(&A->pool[0][i / ASYNC_SEGS][i % ASYNC_SEGS][0], 0ull);
        }
        if (threadIdx.x < static_cast<unsigned>(BARRIER_GROUPS)) {
          store_relaxed(&A->done[threadIdx.x][0], 0u);
        }
        if (threadIdx.x < static_cast<unsigned>(ASYNC_SEGS)) {
          store_relaxed(&A->hungry[0][threadIdx.x][0], 0u);
        }
      }
    }
  }
  // Stamp generations persist in device memory across launches; restart
  // them (and clear the stamps) well before the counter could wrap.
  if (generation > (1 << 30)) {
    for (int k = 0; k < a.nobj; ++k) {
      for (int v = t.id; v < n; v += t.count) {
        desc(a.objs)[k].s.stamp[v] = 0;
      }
    }
    if (a.cDesc != nullptr) {
      for (int v = t.id; v < n; v += t.count) {
        desc(a.cDesc)->s.stamp[v] = 0;
      }
    }
    generation = 0;
    if (!grid.sync()) {
      return generation;
    }
  }
  __syncthreads();

  // Entries of each near / far list: n, or fewer when a caller asks for it
  // (a smaller capacity exercises the sweep that follows a full list).
  const unsigned listCap = static_cast<unsigned>(a.listCap > 0 && a.listCap < n ? a.listCap : n);
  if (a.phases & PH_UPDATE) {
    generation = update_step1<HUBS>(a, grid, t, generation, sh);
    if (grid.aborted) {
      return generation;
    }
    if (kAsyncOnly || (kAsync && a.async != nullptr)) {
      generation = near_far_async<LOCAL>(desc(a.objs), a.nobj, a.ctl->obj, a.async, a.source, generation,
                                  0, nullptr, a.laneBits, static_cast<unsigned>(n), listCap,
                                  a.segCap, a.spillHi, a.spillLo, a.asyncFlags, a.keepHungry, a.abortWord, phaseSeq, grid, t,
                                  sh);
    } else {
      if constexpr (!kAsyncOnly) {
        generation = near_far<HUBS, LOCAL>(desc(a.objs), a.nobj, a.ctl->obj, a.source, generation,
                                           0, nullptr, a.laneBits, a.hopsUpdate,
                                           static_cast<unsigned>(n), listCap, grid, t, sh, 8);
      }
    }
    if (grid.aborted || dropped_any(a, a.nobj, t)) {
      return generation;
    }
    stamp(a.ctl, PHASE_NEARFAR, t.id);
    // One edge sweep for every distance-only objective when they are all
    // views of one out-CSR (the fused update's case), else one each.
    bool fused = HUBS && a.anyFallback;
    for (int k = 0; k < a.nobj; ++k) {
      fused = fused && desc(a.objs)[k].parentBits == 0 && desc(a.objs)[k].outRow == desc(a.objs)[0].outRow &&
              desc(a.objs)[k].outCol == desc(a.objs)[0].outCol;
    }
    u64 *const stage = fused && a.nobj <= STAGE_MAX ? a.stage : nullptr;
    if (stage != nullptr) {
      unpack_staged(desc(a.objs), a.nobj, n, a.source, t, stage);
    } else if (!HUBS && par4_written(a)) {
      if (a.nobj == 2) {
        unpack_interleaved<2>(desc(a.objs), n, t, a.cPar4);
      } else if (a.nobj == 3) {
        unpack_interleaved<3>(desc(a.objs), n, t, a.cPar4);
      } else {
        unpack_interleaved<4>(desc(a.objs), n, t, a.cPar4);
      }
    } else {
      for (int k = 0; k < a.nobj; ++k) {
        unpack(desc(a.objs)[k], n, a.source, t);
      }
    }
    if (a.anyFallback) {
      if (!grid.sync()) {
        return generation;
      }
      if (stage != nullptr) {
        recover_staged(desc(a.objs), a.nobj, n, a.source, t, sh, stage);
      } else if (fused) {
        recover_edges(desc(a.objs), a.nobj, n, a.source, t, sh);
      } else {
        for (int k = 0; k < a.nobj; ++k) {
          if (desc(a.objs)[k].parentBits == 0) {
            recover_parents<HUBS>(desc(a.objs)[k], n, a.source, t, sh);
          }
        }
      }
    }
    if ((a.phases & PH_COMBINED) && !grid.sync()) {  // the K new trees are complete
      return generation;
    }
    if (t.id == 0) {
      a.ctl->clkSospEnd = wall_clock64();
      a.ctl->clk[PHASE_UNPACK] = a.ctl->clkSospEnd;
    }
  }

  if (a.phases & PH_SCRATCH) {
    Desc &D = desc(a.objs)[0];
    scratch_init(D, a.ctl->obj[0], n, a.source, t);
    if (!grid.sync()) {
      return generation;
    }
    if (kAsyncOnly || (kAsync && a.async != nullptr)) {
      generation = near_far_async<LOCAL>(desc(a.objs), 1, a.ctl->obj, a.async, a.source, generation,
                                  D.delta, nullptr, a.laneBits, static_cast<unsigned>(n), listCap,
                                  a.segCap, a.spillHi, a.spillLo, a.asyncFlags, a.keepHungry, a.abortWord, phaseSeq, grid, t,
                                  sh);
    } else {
      if constexpr (!kAsyncOnly) {
        generation = near_far<HUBS, LOCAL>(desc(a.objs), 1, a.ctl->obj, a.source, generation,
                                           D.delta, nullptr, a.laneBits, a.hopsSearch,
                                           static_cast<unsigned>(n), listCap, grid, t, sh, 8);
      }
    }
    if (grid.aborted || dropped_any(a, 1, t)) {
      return generation;
    }
    unpack(D, n, a.source, t);
    if (D.parentBits == 0) {
      if (!grid.sync()) {
        return generation;
      }
      recover_parents<HUBS>(D, n, a.source, t, sh);
    }
  }

  if (a.phases & PH_COMBINED) {
    Desc &C = *desc(a.cDesc);
    ObjCtl &cc = a.ctl->obj[MOSP_MAX_OBJ];
    if (!HUBS && a.cMaskRec != nullptr) {
      combined_masked(a, grid, t);
    } else {
      combined_build<HUBS>(a, grid, t, sh);
    }
    if (grid.aborted) {
      return generation;
    }
    scratch_init(C, cc, n, a.source, t);
    if (threadIdx.x < 2) {
      sh.src[threadIdx.x] = reinterpret_cast<unsigned *>(&a.ctl->combDelta) + threadIdx.x;
    }
    if (!grid.sync(sh.src, 2, sh.snap)) {
      return generation;
    }
    stamp(a.ctl, PHASE_COMB_FILL, t.id);
    const u64 delta = static_cast<u64>(sh.snap[0]) | (static_cast<u64>(sh.snap[1]) << 32);
    if (kAsyncOnly || (kAsync && a.async != nullptr)) {
      generation = near_far_async<LOCAL>(desc(a.cDesc), 1, &cc, a.async, a.source, generation, delta,
                                  &delta, a.laneBits, static_cast<unsigned>(n), listCap, a.segCap,
                                  a.spillHi, a.spillLo, a.asyncFlags, a.keepHungry, a.abortWord,
                                  phaseSeq, grid, t, sh);
    } else {
      if constexpr (!kAsyncOnly) {
        generation = near_far<HUBS, LOCAL>(desc(a.cDesc), 1, &cc, a.source, generation, delta,
                                           &delta, a.laneBits, a.hopsSearch,
                                           static_cast<unsigned>(n), listCap, grid, t, sh, 12);
      }
    }
    if (grid.aborted) {
      return generation;
    }
    stamp(a.ctl, PHASE_COMB_SEARCH, t.id);
    unpack(C, n, a.source, t);
    if (C.parentBits == 0) {
      if (!grid.sync()) {
        return generation;
      }
      recover_parents<HUBS>(C, n, a.source, t, sh);
    }
  }

  return generation;
}

template <bool HUBS, int LOCAL>
__device__ void run_instance(Args &args, const Threads t, Shared &sh) {
  Args &a = *uniform(&args);
  GridBarrier grid{a.ctl, 0, t.block, t.blocks, a.abortWord, a.barrierTicks};
  const int generation = run_phases<HUBS, LOCAL>(a, grid, t, sh, *a.genDev);
#ifdef MOSP_SUBPROFILE
  __syncthreads();
  if (threadIdx.x < 16) atomicAdd(&a.ctl->prof[threadIdx.x], subprof_lds()[threadIdx.x]);
#endif
  if (t.id == 0) {
    // Every exit records the generations it used, so that the next launch
    // never meets one of its stamps: after a barrier gave up, other
    // workgroups may have passed one more barrier than this one and used a
    // few more generations.
    const int used = grid.aborted ? generation + 16 : generation;
    a.ctl->generationOut = static_cast<unsigned>(used);
    *a.genDev = used;
    a.ctl->clkEnd = wall_clock64();
    a.ctl->clk[PHASE_COMB_UNPACK] = a.ctl->clkEnd;
    a.ctl->barriers = grid.epoch;
  }
}

// HUBS: the variant with wave-cooperative handling of high-degree vertices
// (chosen by the host when some vertex has more than HUB_DEGREE edges); the
// other variant keeps the lighter code, and fewer registers, of graphs
// without them.  Both compute the same results.
template <bool HUBS>
__global__ void __launch_bounds__(BLOCK, WAVES_PER_SIMD) sosp_persistent(LaunchArgs a) {
  __shared__ Shared sh;
  (void)a;  // read in place: the first kernel argument starts the kernarg segment
  Args &args = *(Args *)__builtin_amdgcn_kernarg_segment_ptr();
  run_instance<HUBS, HUBS ? 0 : 1>(args, Threads{blockIdx.x * BLOCK + threadIdx.x, gridDim.x * BLOCK,
                                          blockIdx.x, gridDim.x}, sh);
}

// One instance of a batched launch, as a real call: the single-update
// kernel inlines its instance (120 VGPRs, no spills), but inlined into the
// batched kernel's loop over instances the same code spills 78 VGPRs
// (batched 64 road instances: 32.5 ms inlined, 31.4 ms called).
template <bool HUBS>
__device__ __noinline__ void run_batched_instance(Args &args, const Threads t, Shared &sh) {
  run_instance<HUBS, HUBS ? 0 : 2>(args, t, sh);
}

// Many independent instances in one launch: workgroup b belongs to group
// groupOf[b] with rank rankOf[b]; a group runs its instances one after the
// other (each instance synchronizes its group through its own Ctl).
template <bool HUBS>
__global__ void __launch_bounds__(BLOCK, WAVES_PER_SIMD)
    sosp_batched(const LaunchArgs *__restrict__ insts, const BatchGroup *groups,
                 const int *groupOf, const int *rankOf) {
  __shared__ Shared sh;
  const int g = groupOf[blockIdx.x];
  if (g < 0) {
    return;
  }
  const BatchGroup G = groups[g];
  const unsigned rank = static_cast<unsigned>(rankOf[blockIdx.x]);
  const Threads t{rank * BLOCK + threadIdx.x, static_cast<unsigned>(G.size) * BLOCK, rank,
                  static_cast<unsigned>(G.size)};
  for (int i = G.first; i < G.last; ++i) {
    __syncthreads();
    // Read in place (uniform, read-only: scalar loads); an LDS copy would
    // put every field in vector registers and the kernel would spill.
    run_batched_instance<HUBS>(*(Args *)(insts + i), t, sh);
  }
}

// Zero control block blockIdx.x of ctls (one workgroup each) and open the
// far-pile minima; workgroup 0 also clears the launch's shared abort word
// (when there is one besides the blocks' own).
__global__ void ctl_reset(Ctl *ctls, unsigned *abortWord) {
  Ctl *ctl = ctls + blockIdx.x;
  if (blockIdx.x == 0 && threadIdx.x == 0 && abortWord != nullptr) {
    *abortWord = 0;
  }
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

// Host-visible handles of the kernels (used by engine.hip and batch.hip).
// The library builds this file once per workgroup shape (MOSP_BLOCK threads,
// MOSP_MIN_BLOCKS per CU); each build's handles live in the namespace
// MOSP_SHAPE_NS names.
#ifndef MOSP_SHAPE_NS
#define MOSP_SHAPE_NS shape1024
#endif
namespace MOSP_SHAPE_NS {

const void *sosp_kernel_symbol(bool hubs) {
  return hubs ? reinterpret_cast<const void *>(&sosp_persistent<true>)
              : reinterpret_cast<const void *>(&sosp_persistent<false>);
}

const void *sosp_batched_symbol(bool hubs) {
  return hubs ? reinterpret_cast<const void *>(&sosp_batched<true>)
              : reinterpret_cast<const void *>(&sosp_batched<false>);
}

}  // namespace MOSP_SHAPE_NS

// The control-block reset does not depend on the shape: the first build
// alone defines it.
#ifndef MOSP_EXTRA_SHAPE
hipError_t launch_ctl_reset_many(Ctl *ctls, int count, hipStream_t stream, unsigned *abortWord) {
  hipLaunchKernelGGL(ctl_reset, dim3(count), dim3(256), 0, stream, ctls, abortWord);
  return hipGetLastError();
}

hipError_t launch_ctl_reset(Ctl *ctl, hipStream_t stream) {
  return launch_ctl_reset_many(ctl, 1, stream, nullptr);
}
#endif

}  // namespace mosp
