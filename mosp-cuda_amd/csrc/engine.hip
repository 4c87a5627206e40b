// engine.hip -- host side of the gfx950 engine and the C ABI (mosp_hip.h).
//
// Workspace management, descriptor set-up and the cooperative launch of the
// persistent kernel (sosp_kernel.hip); the device-graph build (reverse CSR
// and objective-major weight columns, src/deviceGraph.cu:28-135 in the
// reference) and the GPU path-cost walk (the reference's host
// mospPathCosts, src/combinedGraphGpu.cu:161-208, moved to the device).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"
#include "engine_internal.hpp"
#include "host_api.hpp"
#include "mosp_hip.h"


using namespace mosp;

// ---------------------------------------------------------------------------
// Errors
// ---------------------------------------------------------------------------

namespace mosp {
namespace detail {

thread_local std::string g_error;

int fail(const std::string &message) {
  g_error = message;
  return 0;
}

void clear_error() { g_error.clear(); }

hipStream_t as_stream(void *s) { return static_cast<hipStream_t>(s); }

size_t align256(size_t bytes) { return (bytes + 255) & ~static_cast<size_t>(255); }

}  // namespace detail
}  // namespace mosp
using namespace mosp::detail;

extern "C" const char *mosp_last_error(void) { return g_error.c_str(); }

extern "C" int mosp_set_device(int32_t device) {
  HIP_TRY(hipSetDevice(device));
  return 1;
}

extern "C" int32_t mosp_device_cus(void) {
  int device = 0, cus = 0;
  if (hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) {
    return 0;
  }
  return cus;
}

extern "C" int32_t mosp_device_count(void) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess) {
    return 0;
  }
  return count;
}

extern "C" void *mosp_dev_alloc(uint64_t bytes) {
  void *p = nullptr;
  if (hipMalloc(&p, bytes > 0 ? bytes : 1) != hipSuccess) {
    fail("Error: hipMalloc of " + std::to_string(bytes) + " bytes failed.");
    return nullptr;
  }
  return p;
}

extern "C" int mosp_dev_free(void *ptr) {
  HIP_TRY(hipFree(ptr));
  return 1;
}

extern "C" int mosp_memcpy(void *dst, const void *src, uint64_t bytes, int32_t kind) {
  if (bytes == 0) {
    return 1;
  }
  const hipMemcpyKind k = kind == 1   ? hipMemcpyHostToDevice
                          : kind == 2 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  HIP_TRY(hipMemcpy(dst, src, bytes, k));
  return 1;
}

extern "C" int mosp_device_synchronize(void) {
  HIP_TRY(hipDeviceSynchronize());
  return 1;
}

// ---------------------------------------------------------------------------
// Workspace
// ---------------------------------------------------------------------------

struct mosp_ws {
  int capacity = 0;
  int slots = 0;
  void *slab = nullptr;
  std::vector<ObjState> state;
  ObjDesc *dDesc = nullptr;  // MOSP_MAX_OBJ + 1 descriptors
  ObjDesc *hDesc = nullptr;  // pinned staging
  Ctl *dCtl = nullptr;
  Ctl *hCtl = nullptr;       // pinned readback
  int *genDev = nullptr;
  // combined graph
  int cCapacity = 0;
  int cK = 0;
  int *cSlot = nullptr;     // combined edges' places in their parents' rows (K n)
  int *cDegree = nullptr, *cRowPtr = nullptr, *cColInd = nullptr,
      *cWeights = nullptr, *cPartials = nullptr;
  int4 *cEdge4 = nullptr;   // combined records with row ranges (asynchronous loop)
  // The combined graph as a mask over a device graph's out-edges (fused
  // launches of graphs without hubs with the asynchronous loop): one record
  // per edge, and the combined descriptor naming them; the CSR descriptor
  // is kept for launches that build the combined CSR instead.
  bool combMask = true;     // MOSP_COMB_MASK=0: always build the combined CSR
  int initPush = 0;         // LaunchArgs::initPush (MOSP_INIT_PUSH)
  int4 *cMask = nullptr;
  size_t cMaskCap = 0;      // records
  int cMaskIndex = -1;      // descriptor slot of the combined search, or -1
  ObjDesc cMaskDesc{}, cCsrDesc{};
  double clockKhz = 100000;  // wall_clock64 rate
  int maxBlocks = 0;        // cooperative residency limit (1024-thread shape)
  int maxBlocks512 = 0;     // ... of the 512-thread shape
  int gridBlocks = 0;       // workgroups per launch of the 1024-thread shape
  int gridOverride = 0;
  int gridRequest = 0;      // workgroups asked for (MOSP_GRID or mosp_ws_set_grid; 0: all)
  int shapeThreads = 0;     // workgroup shape: 1024, 512, or 0 (512 for the hub variant)
  int queueCap = 0;         // LDS queue entries of the push loop (0: all)
  int laneBits = 2;         // 4 lanes per frontier vertex
  unsigned *degreeWord = nullptr;  // max-degree reduction of caller CSRs
  int combDeltaScale = 1;   // combined search: default width times this
  int hopsUpdate = kHopsUpdate;  // local relaxation rounds per near-far iteration
  int hopsSearch = kHopsSearch;
  int listCap = 0;          // near / far list entries: 0 = n
  double barrierMs = -1;    // longest grid-barrier wait before a launch gives up (< 0: default)
  bool scrub = false;       // a launch gave up: clear the scratch before the next
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // Asynchronous near-far (light single-update kernel): its counters, the
  // entries of each pool segment of the slab's states, the spill thresholds.
  AsyncCtl *dAsync = nullptr;
  bool async = true;
  unsigned segCap = 0;
  unsigned spillHi = 0, spillLo = 0;  // 0: from the queue capacity
  // Defaults measured on a 1.96M-vertex road grid (K = 3): spills to XCDs
  // with idle workgroups on, listed words re-read (carrying them raised the
  // work by 12 % for no net gain).
  unsigned asyncFlags = ASYNC_HUNGRY;
  bool lastAsync = false;  // the last launch ran the asynchronous loop
  int lastGrid = 0;         // workgroups of the last launch
  unsigned keepHungry = 0;            // 0: one batch (BLOCK >> laneBits)
};

namespace mosp {
namespace detail {

void ws_release_slab(mosp_ws *ws) {
  (void)hipFree(ws->slab);
  ws->slab = nullptr;
  ws->state.clear();
  ws->capacity = 0;
  ws->slots = 0;
  ws->segCap = 0;
}

void ws_release_combined(mosp_ws *ws) {
  (void)hipFree(ws->cDegree);
  (void)hipFree(ws->cRowPtr);
  (void)hipFree(ws->cColInd);  // cWeights lies in the same allocation
  (void)hipFree(ws->cEdge4);
  ws->cEdge4 = nullptr;
  (void)hipFree(ws->cSlot);
  ws->cSlot = nullptr;
  (void)hipFree(ws->cMask);
  ws->cMask = nullptr;
  ws->cMaskCap = 0;
  ws->cDegree = ws->cRowPtr = ws->cColInd = ws->cWeights = nullptr;
  ws->cCapacity = 0;
  ws->cK = 0;
}

int ws_init(mosp_ws *ws) {
  if (ws->dCtl != nullptr) {
    return 1;
  }
  int device = 0, cooperative = 0, cus = 0, perCu = 0;
  HIP_TRY(hipGetDevice(&device));
  HIP_TRY(hipDeviceGetAttribute(&cooperative, hipDeviceAttributeCooperativeLaunch,
                                device));
  if (!cooperative) {
    return fail("Error: the GPU does not support cooperative launches.");
  }
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount,
                                device));
  ws->maxBlocks = shape_max_blocks(kernel_shape(1024), false);
  ws->maxBlocks512 = shape_max_blocks(kernel_shape(512), false);
  if (ws->maxBlocks < 1 || ws->maxBlocks512 < 1) {
    return fail("Error: the persistent kernel does not fit on a CU.");
  }
  (void)perCu;
  int clockKhz = 0;
  if (hipDeviceGetAttribute(&clockKhz, hipDeviceAttributeWallClockRate, device) ==
          hipSuccess &&
      clockKhz > 0) {
    ws->clockKhz = clockKhz;
  }
  // Default: one 1024-thread workgroup per CU -- 262K lanes for the frontier
  // sweeps with only one barrier arrival per CU (measured best on MI355X
  // among 256/512/1024-thread workgroups at 1-2 per CU).
  int blocks = cus;
  if (const char *env = std::getenv("MOSP_GRID")) {
    blocks = std::atoi(env);
    ws->gridRequest = std::max(0, blocks);
  }
  if (ws->gridOverride > 0) {
    blocks = ws->gridOverride;
    ws->gridRequest = blocks;
  }
  ws->gridBlocks = std::max(1, std::min(blocks, ws->maxBlocks));
  if (const char *env = std::getenv("MOSP_WG_THREADS")) {
    const int threads = std::atoi(env);
    ws->shapeThreads = threads == 512 || threads == 1024 ? threads : 0;
  }
  if (const char *env = std::getenv("MOSP_QUEUE_CAP")) {
    ws->queueCap = std::max(0, std::min(MOSP_QUEUE_CAP_MAX, std::atoi(env)));
  }
  if (const char *env = std::getenv("MOSP_LANE_BITS")) {
    ws->laneBits = std::max(0, std::min(5, std::atoi(env)));
  }
  local_hops_override(ws->hopsUpdate, ws->hopsSearch);
  ws->listCap = list_cap_override();
  // MOSP_ASYNC=0: the iteration-synchronous near-far loop everywhere;
  // MOSP_ASYNC_SPILL=hi,lo: the local queue's spill thresholds.
  if (const char *env = std::getenv("MOSP_ASYNC")) {
    ws->async = std::atoi(env) != 0;
  }
  if (const char *env = std::getenv("MOSP_ASYNC_SPILL")) {
    unsigned hi = 0, lo = 0;
    if (std::sscanf(env, "%u,%u", &hi, &lo) == 2 && lo < hi) {
      ws->spillHi = hi;
      ws->spillLo = lo;
    }
  }
  if (const char *env = std::getenv("MOSP_ASYNC_FLAGS")) {
    ws->asyncFlags = static_cast<unsigned>(std::atoi(env)) & ASYNC_FLAG_MASK;
  }
  if (const char *env = std::getenv("MOSP_COMB_MASK")) {
    ws->combMask = std::atoi(env) != 0;
  }
  ws->initPush = init_push_override();
  if (const char *env = std::getenv("MOSP_ASYNC_KEEP")) {
    ws->keepHungry = static_cast<unsigned>(std::max(0, std::atoi(env)));
  }
  HIP_TRY(hipMalloc(&ws->dAsync, sizeof(AsyncCtl)));
  HIP_TRY(hipMalloc(&ws->dDesc, sizeof(ObjDesc) * (MOSP_MAX_OBJ + 1)));
  HIP_TRY(hipHostMalloc(&ws->hDesc, sizeof(ObjDesc) * (MOSP_MAX_OBJ + 1),
                        hipHostMallocDefault));
  HIP_TRY(hipMalloc(&ws->dCtl, sizeof(Ctl)));
  HIP_TRY(hipHostMalloc(&ws->hCtl, sizeof(Ctl), hipHostMallocDefault));
  HIP_TRY(hipMalloc(&ws->genDev, sizeof(int)));
  HIP_TRY(hipMemset(ws->genDev, 0, sizeof(int)));
  HIP_TRY(hipMalloc(&ws->degreeWord, sizeof(unsigned)));
  HIP_TRY(hipMalloc(&ws->cPartials, sizeof(int) * std::max(ws->maxBlocks, ws->maxBlocks512)));
  HIP_TRY(hipEventCreate(&ws->ev0));
  HIP_TRY(hipEventCreate(&ws->ev1));
  return 1;
}

// Reports the input-tree checks of a launch's pack pass (ObjCtl::overflow).
int input_tree_ok(unsigned overflow, const std::string &where) {
  if (overflow & 2u) {
    return fail("Error: " + where + "an input parent is not a vertex id in [-1, n); the "
                "initial tree does not belong to this graph.");
  }
  if (overflow & 1u) {
    return fail("Error: " + where + "an input distance is negative or larger than "
                "(n - 1) * maxWeight; the initial tree does not belong to this graph.");
  }
  return 1;
}

int run_launch(mosp_ws *ws, LaunchArgs &args, int descCount, hipStream_t stream,
               float *kernelMs);

// One empty persistent launch when the workspace is created: the first
// cooperative launch of a process pays one-time set-up costs, which belong
// to workspace creation, not to the first update.
int ws_warm_up(mosp_ws *ws) {
  LaunchArgs args{};
  args.objs = ws->dDesc;
  return run_launch(ws, args, 0, nullptr, nullptr);
}

// Entries of one pool segment for graphs of up to n vertices: the segments
// hold 2n entries in all (a spill that finds its segment full keeps its
// entries in the local queue).  MOSP_POOL_CAP=c: c entries (tests).
unsigned pool_seg_cap(int n) {
  if (const char *env = std::getenv("MOSP_POOL_CAP")) {
    const int cap = std::atoi(env);
    if (cap > 0) {
      return static_cast<unsigned>(cap);
    }
  }
  const size_t entries = 2 * static_cast<size_t>(std::max(n, 1));
  return static_cast<unsigned>(std::max<size_t>(256, (entries + ASYNC_SEGS - 1) / ASYNC_SEGS));
}

size_t state_bytes(int n, unsigned segCap) {
  const size_t count = static_cast<size_t>(std::max(n, 1));
  return align256(count * sizeof(u64)) + 7 * align256(count * sizeof(int)) +
         align256(2 * static_cast<size_t>(HUB_LIST) * sizeof(int)) +
         align256(static_cast<size_t>(ASYNC_SEGS) * segCap * sizeof(int));
}

size_t pool_bytes(unsigned segCap) {
  return static_cast<size_t>(ASYNC_SEGS) * segCap * sizeof(int);
}

// segCap 0: no pools (st.pool stays null).
ObjState carve_state(char *&p, int n, unsigned segCap) {
  const size_t count = static_cast<size_t>(std::max(n, 1));
  const size_t wordBytes = align256(count * sizeof(u64));
  const size_t intBytes = align256(count * sizeof(int));
  ObjState st{};
  st.packed = reinterpret_cast<u64 *>(p);
  p += wordBytes;
  int **ints[7] = {&st.stamp, &st.inFar, &st.ancestor, &st.near[0],
                   &st.near[1], &st.far[0], &st.far[1]};
  for (int i = 0; i < 7; ++i) {
    *ints[i] = reinterpret_cast<int *>(p);
    p += intBytes;
  }
  st.hubs = reinterpret_cast<int *>(p);
  p += align256(2 * static_cast<size_t>(HUB_LIST) * sizeof(int));
  if (segCap > 0) {
    st.pool = reinterpret_cast<int *>(p);
    p += align256(pool_bytes(segCap));
  }
  return st;
}

// Pool slots are -1 when free.
int fill_pools(mosp_ws *ws, hipStream_t stream) {
  for (const ObjState &st : ws->state) {
    if (st.pool != nullptr) {
      HIP_TRY(hipMemsetAsync(st.pool, 0xff, pool_bytes(ws->segCap), stream));
    }
  }
  return 1;
}

int ws_reserve(mosp_ws *ws, int capacity, int slots) {
  const bool fresh = ws->dCtl == nullptr;
  if (!ws_init(ws) || (fresh && !ws_warm_up(ws))) {
    return 0;
  }
  slots = std::max(slots, 1);
  if (capacity <= ws->capacity && slots <= ws->slots) {
    return 1;
  }
  capacity = std::max(capacity, ws->capacity);
  slots = std::max(slots, ws->slots);
  ws_release_slab(ws);
  const int n = std::max(capacity, 1);
  const unsigned segCap = pool_seg_cap(n);
  const size_t perSlot = state_bytes(n, segCap);
  HIP_TRY(hipMalloc(&ws->slab, perSlot * slots));
  // stamps and far marks must start at zero
  HIP_TRY(hipMemset(ws->slab, 0, perSlot * slots));
  char *p = static_cast<char *>(ws->slab);
  ws->state.resize(slots);
  for (int s = 0; s < slots; ++s) {
    ws->state[s] = carve_state(p, n, segCap);
  }
  ws->capacity = capacity;
  ws->slots = slots;
  ws->segCap = segCap;
  if (!fill_pools(ws, nullptr)) {
    return 0;
  }
  HIP_TRY(hipStreamSynchronize(nullptr));
  return 1;
}

int ws_reserve_combined(mosp_ws *ws, int n, int K) {
  if (n <= ws->cCapacity && K <= ws->cK) {
    return 1;
  }
  ws_release_combined(ws);
  const size_t nn = static_cast<size_t>(std::max(n, 1));
  HIP_TRY(hipMalloc(&ws->cDegree, (nn + 1) * sizeof(int)));
  HIP_TRY(hipMalloc(&ws->cRowPtr, (nn + 1) * sizeof(int)));
  // Columns and weights in one allocation: together they also hold the K
  // staged words per vertex of the fused parent recovery (LaunchArgs::stage).
  HIP_TRY(hipMalloc(&ws->cColInd, 2 * nn * K * sizeof(int)));
  ws->cWeights = ws->cColInd + nn * K;
  if (ws->async) {
    HIP_TRY(hipMalloc(&ws->cEdge4, nn * K * sizeof(int4)));
  }
  HIP_TRY(hipMalloc(&ws->cSlot, nn * K * sizeof(int)));
  ws->cCapacity = n;
  ws->cK = K;
  return 1;
}


const KernelShape &kernel_shape(int threads) {
  static const KernelShape shapes[2] = {
      {1024, shape1024::sosp_kernel_symbol, shape1024::sosp_batched_symbol},
      {512, shape512::sosp_kernel_symbol, shape512::sosp_batched_symbol}};
  return threads == 512 ? shapes[1] : shapes[0];
}

int shape_max_blocks(const KernelShape &shape, bool batched) {
  int device = 0, cus = 0, perCu = 0;
  if (hipGetDevice(&device) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) !=
          hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &perCu, batched ? shape.batched(true) : shape.persistent(true), shape.threads, 0) !=
          hipSuccess) {
    return 0;
  }
  return perCu * cus;
}

int choose_packing(int n, long long maxWeight, PackChoice &pc, bool optimistic) {
  const u64 outputMax = static_cast<u64>(DISTANCE_INF / 2 - 1);
  const u64 weight = static_cast<u64>(std::max(maxWeight, 1LL));
  const u64 hops = static_cast<u64>(std::max(n - 1, 1));
  if (weight > outputMax / hops) {
    return fail("Error: distances up to " + std::to_string(weight) + " * " +
                std::to_string(hops) + " do not fit in 62 bits.");
  }
  pc.bound = weight * hops;
  int bits = 1;
  while ((1ULL << bits) - 1 < static_cast<u64>(n)) {
    ++bits;
  }
  u64 maxDistance = (PACKED_INF >> bits) - 1;
  // MOSP_PACK_LIMIT=d: packed words hold distances up to d only, so that
  // small graphs take the optimistic packing and its redo (tests).
  if (const char *env = std::getenv("MOSP_PACK_LIMIT")) {
    const long long limit = std::atoll(env);
    if (limit > 0 && static_cast<u64>(limit) < maxDistance) {
      maxDistance = static_cast<u64>(limit);
    }
  }
  pc.optimistic = false;
  pc.maxDistance = pc.bound;
  if (pc.bound <= maxDistance || optimistic) {
    pc.bits = bits;
    pc.noParent = (1ULL << bits) - 1;
    if (pc.bound > maxDistance) {
      pc.optimistic = true;
      pc.maxDistance = maxDistance;
    }
  } else {
    pc.bits = 0;
    pc.noParent = 0;
  }
  return 1;
}

int jump_rounds(int n) {
  int rounds = 0;
  while ((1LL << rounds) < n) {
    ++rounds;
  }
  return rounds + 1;
}

void fill_stats(const Ctl &c, int k, const ObjDesc &d, float ms,
                mosp_sosp_stats *stats) {
  if (stats == nullptr) {
    return;
  }
  const ObjCtl &o = c.obj[k];
  stats->invalidated = static_cast<int32_t>(o.invalidated);
  stats->jump_rounds = static_cast<int32_t>(c.rounds);
  stats->iterations = static_cast<int32_t>(o.iterations);
  stats->epochs = static_cast<int32_t>(o.epochs);
  stats->pushes = static_cast<int64_t>(o.pushes);
  stats->packed_parents = d.parentBits > 0 ? 1 : 0;
  stats->kernel_ms = ms;
}

// Whether an optimistic launch met a distance past its packing's limit for
// one of objectives [0, count) (it then wrote no result).
bool launch_dropped(const Ctl &c, int count) {
  for (int k = 0; k < count; ++k) {
    if (c.obj[k].dropped != 0) {
      return true;
    }
  }
  return false;
}

// Upload descriptors, reset the control block, run the cooperative launch,
// read the control block back.  Synchronous on `stream`.
double barrier_timeout_ms_default() {
  const char *env = std::getenv("MOSP_BARRIER_TIMEOUT_MS");
  return env != nullptr && *env != '\0' ? std::max(0.0, std::atof(env)) : 2000.0;
}

unsigned long long barrier_ticks(double ms, double clockKhz) {
  const double ticks = ms * clockKhz;
  return ticks >= 1.8e19 ? ~0ULL : static_cast<unsigned long long>(ticks);
}

hipError_t launch_persistent(const void *kernel, dim3 grid, dim3 block, void **params,
                             hipStream_t stream) {
  static const bool plain = [] {
    const char *env = std::getenv("MOSP_LAUNCH");
    return env != nullptr && std::strcmp(env, "plain") == 0;
  }();
  if (plain) {
    return hipLaunchKernel(kernel, grid, block, params, 0, stream);
  }
  const hipError_t err = hipLaunchCooperativeKernel(kernel, grid, block, params, 0, stream);
  if (err != hipErrorIllegalState) {
    return err;
  }
  (void)hipGetLastError();  // the refused cooperative launch
  return hipLaunchKernel(kernel, grid, block, params, 0, stream);
}

// Workgroups of a launch of `shape`: the caller's request (MOSP_GRID or
// mosp_ws_set_grid), else every workgroup of the shape that fits on the
// device at once (one per CU at 1024 threads, two at 512).
int launch_grid(const mosp_ws *ws, const KernelShape &shape) {
  const int most = shape.threads == 512 ? ws->maxBlocks512 : ws->maxBlocks;
  const int blocks = ws->gridRequest > 0 ? ws->gridRequest : most;
  return std::max(1, std::min(blocks, most));
}

int run_launch(mosp_ws *ws, LaunchArgs &args, int descCount,
               hipStream_t stream, float *kernelMs) {
  if (ws->scrub) {
    // The last launch ended early: its stamps, far marks and invalidation
    // marks are not in the state the next launch relies on.
    if (ws->slab != nullptr) {
      HIP_TRY(hipMemsetAsync(ws->slab, 0,
                             state_bytes(std::max(ws->capacity, 1), ws->segCap) * ws->slots,
                             stream));
      if (!fill_pools(ws, stream)) {
        return 0;
      }
    }
    HIP_TRY(hipMemsetAsync(ws->genDev, 0, sizeof(int), stream));
    ws->scrub = false;
  }
  args.ctl = ws->dCtl;
  args.abortWord = &ws->dCtl->gaveUp;
  const double waitMs = ws->barrierMs >= 0 ? ws->barrierMs : barrier_timeout_ms_default();
  args.barrierTicks = barrier_ticks(waitMs, ws->clockKhz);
  args.genDev = ws->genDev;
  args.laneBits = ws->laneBits;
  args.hopsUpdate = ws->hopsUpdate;
  args.hopsSearch = ws->hopsSearch;
  args.listCap = ws->listCap;
  args.queueCap = ws->queueCap;
  args.hubEdges = hub_edges_override();
  args.cPartials = ws->cPartials;
  args.cDeltaScale = ws->combDeltaScale;
  args.initPush = ws->initPush;
  // The asynchronous near-far loop runs in the light single-update kernel
  // when the workspace has pools; spill thresholds from the queue capacity.
  const unsigned qcap = ws->queueCap >= 256 ? static_cast<unsigned>(ws->queueCap)
                                            : static_cast<unsigned>(MOSP_QUEUE_CAP_MAX);
  args.async = nullptr;
  args.segCap = ws->segCap;
  args.asyncFlags = ws->asyncFlags;
  const int forced = hubs_override();
  const bool hubs = forced >= 0 ? forced != 0 : args.hubs != 0;
  // Shape: the caller's, else 512-thread workgroups two per CU for the hub
  // variant (R-MAT scale 22: 9.3 -> 8.4 ms per update) and 1024-thread ones
  // for the light variant (road grid 1400x1400: 4.2 ms against 4.7).
  const KernelShape &shape = kernel_shape(ws->shapeThreads > 0 ? ws->shapeThreads
                                                                : (hubs ? 512 : 1024));
  const int grid = launch_grid(ws, shape);
  if (ws->async && !hubs && ws->segCap > 0 && !ws->state.empty() && grid <= ASYNC_MAX_GRID) {
    args.async = ws->dAsync;
  }
  // Spill above a quarter of the queue capacity (at most half the local
  // ring, so a batch's appends always fit) down to half of that; idle
  // workgroups elsewhere take all but one batch.
  const unsigned ring = std::min(qcap, async_ring_entries(shape.threads));
  args.spillHi = ws->spillHi > 0 ? std::min(ws->spillHi, qcap - 1) : std::min(qcap / 4, ring / 2);
  args.spillLo = ws->spillHi > 0 ? std::min(ws->spillLo, args.spillHi - 1) : args.spillHi / 2;
  args.keepHungry = ws->keepHungry > 0
                        ? ws->keepHungry
                        : static_cast<unsigned>(shape.threads >> std::max(0, ws->laneBits));
  ws->lastAsync = args.async != nullptr;
  ws->lastGrid = grid;
  // The masked combined graph (set up by the fused entry point for this
  // launch) is searched by the asynchronous loop only; otherwise the launch
  // builds the combined CSR and searches it.
  if (args.cMaskRec != nullptr) {
    const bool inRange = ws->cMaskIndex >= 0 && ws->cMaskIndex < descCount;
    const bool mask = args.async != nullptr && inRange;
    if (inRange) {
      ws->hDesc[ws->cMaskIndex] = mask ? ws->cMaskDesc : ws->cCsrDesc;
    }
    if (!mask) {
      args.cMaskRec = nullptr;
      args.cPar4 = nullptr;
    }
  }
  if (descCount > 0) {
    HIP_TRY(hipMemcpyAsync(ws->dDesc, ws->hDesc, sizeof(ObjDesc) * descCount,
                           hipMemcpyHostToDevice, stream));
  }
  HIP_TRY(launch_ctl_reset(ws->dCtl, stream));
  void *params[] = {&args};
  HIP_TRY(hipEventRecord(ws->ev0, stream));
  HIP_TRY(launch_persistent(shape.persistent(hubs), dim3(grid), dim3(shape.threads), params,
                            stream));
  HIP_TRY(hipEventRecord(ws->ev1, stream));
  HIP_TRY(hipMemcpyAsync(ws->hCtl, ws->dCtl, CTL_READBACK,
                         hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, ws->ev0, ws->ev1));
#ifdef MOSP_SUBPROFILE
  {
    // Sums over workgroups of wall_clock64 ticks, per workgroup-iteration;
    // "release" is the barrier time of the last workgroup to arrive.
    const Ctl &h = *ws->hCtl;
    const double tick = 1e3 / static_cast<double>(ws->clockKhz);  // us
    const double G = static_cast<double>(ws->gridBlocks);
    const double itU = std::max(1.0, h.prof[11] / G), itC = std::max(1.0, h.prof[15] / G);
    if (h.clk[PHASE_COUNT] > h.clk[PHASE_INVALIDATE] && h.clk[PHASE_PULL] > h.clk[PHASE_COUNT]) {
      std::fprintf(stderr, "pull split: sweep %.1f us, %llu heavy candidates %.1f us\n",
                   (h.clk[PHASE_COUNT] - h.clk[PHASE_INVALIDATE]) * tick,
                   static_cast<unsigned long long>(h.clk[PHASE_COUNT + 1]),
                   (h.clk[PHASE_PULL] - h.clk[PHASE_COUNT]) * tick);
    }
    if (args.async != nullptr) {
      // Asynchronous near-far: per workgroup and launch (ticks -> us).
      const double ph = std::max(1.0, h.prof[11] / G);
      std::fprintf(stderr,
                   "asyncprof grid=%d barriers=%u phases %.1f: busy %.1f idle %.1f us, phase "
                   "body %.1f us, barrier %.1f us (per wg, summed over phases); per wg: local "
                   "batches %.1f (%.1f entries each), pool batches %.1f (%.1f each), seed "
                   "passes %.1f, idle polls %.1f, spills %.1f (%.1f entries each); lost "
                   "events %llu; per batch: decide %.2f relax %.2f rest %.2f us\n",
                   ws->gridBlocks, h.barriers, ph, h.prof[0] * tick / G, h.prof[1] * tick / G,
                   h.prof[12] * tick / G, h.prof[10] * tick / G, h.prof[4] / G,
                   h.prof[2] / std::max(1.0, static_cast<double>(h.prof[4])), h.prof[5] / G,
                   h.prof[3] / std::max(1.0, static_cast<double>(h.prof[5])), h.prof[15] / G,
                   h.prof[7] / G, h.prof[8] / G,
                   h.prof[9] / std::max(1.0, static_cast<double>(h.prof[8])), h.prof[13],
                   h.prof[6] * tick / std::max(1.0, static_cast<double>(h.prof[4] + h.prof[5] + h.prof[15])),
                   h.prof[14] * tick / std::max(1.0, static_cast<double>(h.prof[4] + h.prof[5] + h.prof[15])),
                   (h.prof[0] - h.prof[6] - h.prof[14]) * tick /
                       std::max(1.0, static_cast<double>(h.prof[4] + h.prof[5] + h.prof[15])));
    }
    std::fprintf(stderr,
                 "subprof grid=%d barriers=%u update (%.0f iters): snapshot %.2f work %.2f "
                 "wait %.2f release %.2f us/iter; push loop %.2f flush %.2f us/call (%.0f calls); "
                 "combined (%.0f iters): snapshot %.2f work %.2f wait %.2f release %.2f\n",
                 ws->gridBlocks, h.barriers, itU, h.prof[8] * tick / G / itU,
                 h.prof[9] * tick / G / itU, h.prof[10] * tick / G / itU, h.prof[6] * tick / itU,
                 h.prof[3] * tick / std::max<double>(1, h.prof[5]),
                 h.prof[4] * tick / std::max<double>(1, h.prof[5]), static_cast<double>(h.prof[5]),
                 itC, h.prof[12] * tick / G / itC, h.prof[13] * tick / G / itC,
                 h.prof[14] * tick / G / itC, h.prof[7] * tick / itC);
  }
#endif
  if (kernelMs != nullptr) {
    *kernelMs = ms;
  }
  static const bool asyncDebug = std::getenv("MOSP_ASYNC_DEBUG") != nullptr;
  if (asyncDebug && args.async != nullptr) {
    std::fprintf(stderr, "async events: lost batches %u, short spills %u, sweep phases %u\n",
                 ws->hCtl->asyncLost, ws->hCtl->asyncShortSpills, ws->hCtl->asyncSweeps);
  }
  if (ws->hCtl->gaveUp != 0) {
    ws->scrub = true;
    return fail("Error: a grid barrier of the persistent launch waited more than " +
                std::to_string(waitMs) +
                " ms and the launch gave up (its workgroups were not all resident: the device "
                "was running other persistent work); no result was written.");
  }
  for (int k = 0; k < args.nobj; ++k) {
    if (!input_tree_ok(ws->hCtl->obj[k].overflow, "")) {
      return 0;
    }
  }
  return 1;
}

ObjDesc make_desc(const mosp_csr &out, const mosp_csr *in,
                  const mosp_changes *ch, const PackChoice &pc, u64 delta,
                  const ObjState &st) {
  ObjDesc d{};
  d.outRow = out.row_ptr;
  d.outCol = out.col_ind;
  d.outW = out.weights;
  if (in != nullptr) {
    d.inRow = in->row_ptr;
    d.inCol = in->col_ind;
    d.inW = in->weights;
  }
  if (ch != nullptr) {
    d.chFrom = ch->changed_from;
    d.chTo = ch->changed_to;
    d.nChanged = ch->n_changed;
    d.heads = ch->insert_heads;
    d.nHeads = ch->n_insert_heads;
  }
  d.parentBits = pc.bits;
  d.noParent = pc.noParent;
  d.maxDistance = pc.maxDistance;
  d.optimistic = pc.optimistic ? 1 : 0;
  d.delta = delta;
  d.s = st;
  return d;
}

void local_hops_override(int &update, int &search) {
  const char *env = std::getenv("MOSP_LOCAL_HOPS");
  int u = 0, c = 0;
  if (env != nullptr && std::sscanf(env, "%d,%d", &u, &c) == 2) {
    update = std::max(0, std::min(8, u));
    search = std::max(0, std::min(8, c));
  }
}

// MOSP_INIT_PUSH=0|1|2: Step 1's candidates for invalidated vertices pushed
// from the valid vertices when more than half are invalidated (0, default),
// always pulled (1), always pushed (2); results are the same (tests).
int init_push_override() {
  const char *env = std::getenv("MOSP_INIT_PUSH");
  const int v = env != nullptr ? std::atoi(env) : 0;
  return v == 1 || v == 2 ? v : 0;
}

// MOSP_LIST_CAP=c: near / far lists of c entries instead of n, so that
// small graphs take the sweep that follows a full list (tests).
int list_cap_override() {
  const char *env = std::getenv("MOSP_LIST_CAP");
  return env == nullptr ? 0 : std::max(0, std::atoi(env));
}

int hub_edges_override() {
  const char *env = std::getenv("MOSP_HUB_EDGES");
  return env == nullptr || *env == '\0' ? HUB_GRID : std::max(64, std::atoi(env));
}

int hubs_override() {
  const char *env = std::getenv("MOSP_HUBS");
  return env == nullptr || *env == '\0' ? -1 : (std::atoi(env) != 0 ? 1 : 0);
}

// Whether some vertex of the caller's CSRs (out and in row arrays of K
// objectives; views of one graph share them) has more than HUB_DEGREE edges.
int caller_hubs(mosp_ws *ws, const mosp_csr *out, const mosp_csr *in, int K, hipStream_t stream,
                int &hubs) {
  hubs = 0;
  std::vector<const int *> seen;
  for (int k = 0; k < K && !hubs; ++k) {
    for (const int *row : {out[k].row_ptr, in != nullptr ? in[k].row_ptr : nullptr}) {
      if (row == nullptr || std::find(seen.begin(), seen.end(), row) != seen.end()) {
        continue;
      }
      seen.push_back(row);
      int degree = 0;
      if (!graph_max_degree(row, nullptr, out[k].n, ws->degreeWord, stream, &degree)) {
        return 0;
      }
      if (degree > HUB_DEGREE) {
        hubs = 1;
        break;
      }
    }
  }
  return 1;
}

}  // namespace detail
}  // namespace mosp

extern "C" mosp_ws *mosp_ws_create(void) { return new mosp_ws(); }

extern "C" int mosp_ws_reserve(mosp_ws *ws, int32_t capacity, int32_t slots) {
  return ws_reserve(ws, capacity, slots);
}

extern "C" void mosp_ws_destroy(mosp_ws *ws) {
  if (ws == nullptr) {
    return;
  }
  ws_release_slab(ws);
  ws_release_combined(ws);
  (void)hipFree(ws->dDesc);
  (void)hipHostFree(ws->hDesc);
  (void)hipFree(ws->dCtl);
  (void)hipHostFree(ws->hCtl);
  (void)hipFree(ws->genDev);
  (void)hipFree(ws->degreeWord);
  (void)hipFree(ws->cPartials);
  (void)hipFree(ws->dAsync);
  if (ws->ev0) (void)hipEventDestroy(ws->ev0);
  if (ws->ev1) (void)hipEventDestroy(ws->ev1);
  delete ws;
}

extern "C" int mosp_ws_set_grid(mosp_ws *ws, int32_t workgroups) {
  ws->gridOverride = workgroups;
  ws->gridRequest = std::max(0, workgroups);
  if (ws->maxBlocks > 0) {
    int blocks = workgroups > 0 ? workgroups : ws->maxBlocks;
    ws->gridBlocks = std::max(1, std::min(blocks, ws->maxBlocks));
  }
  return 1;
}

extern "C" int32_t mosp_ws_grid(const mosp_ws *ws) {
  return ws->lastGrid > 0 ? ws->lastGrid : ws->gridBlocks;
}

extern "C" int mosp_ws_set_shape(mosp_ws *ws, int32_t threads, int32_t queue_cap) {
  clear_error();
  if (ws == nullptr || (threads != 0 && threads != 512 && threads != 1024) || queue_cap < 0 ||
      (queue_cap > 0 && (queue_cap < 256 || queue_cap > MOSP_QUEUE_CAP_MAX))) {
    return fail("Error: workgroups of 0 (automatic), 512 or 1024 threads and an LDS queue of "
                "0 (all) or 256.." + std::to_string(MOSP_QUEUE_CAP_MAX) + " entries.");
  }
  ws->shapeThreads = threads;
  ws->queueCap = queue_cap;
  return 1;
}

extern "C" int mosp_ws_set_barrier_timeout(mosp_ws *ws, double ms) {
  clear_error();
  if (ws == nullptr || !(ms >= 0)) {
    return fail("Error: the barrier timeout must be >= 0 ms.");
  }
  ws->barrierMs = ms;
  return 1;
}

extern "C" int mosp_ws_set_combined_delta_scale(mosp_ws *ws, int32_t scale) {
  clear_error();
  if (ws == nullptr || scale < 1) {
    return fail("Error: the combined width scale must be >= 1.");
  }
  ws->combDeltaScale = scale;
  return 1;
}

extern "C" int mosp_ws_async_events(const mosp_ws *ws, int32_t events[4]) {
  clear_error();
  if (ws == nullptr || events == nullptr) {
    return fail("Error: missing workspace or event array.");
  }
  const Ctl *c = ws->hCtl;
  events[0] = c != nullptr ? static_cast<int32_t>(c->asyncLost) : 0;
  events[1] = c != nullptr ? static_cast<int32_t>(c->asyncShortSpills) : 0;
  events[2] = c != nullptr ? static_cast<int32_t>(c->asyncSweeps) : 0;
  events[3] = ws->lastAsync ? 1 : 0;
  return 1;
}

extern "C" int mosp_ws_last_profile(const mosp_ws *ws, double *ms, int32_t nPhases,
                                    int32_t *barriers) {
  if (ws == nullptr || ws->hCtl == nullptr) {
    return 0;
  }
  const Ctl &c = *ws->hCtl;
  const int count = std::min<int>(nPhases, PHASE_COUNT);
  u64 prev = c.clkStart;
  for (int i = 0; i < count; ++i) {
    if (c.clk[i] != 0 && c.clk[i] >= prev) {
      ms[i] = static_cast<double>(c.clk[i] - prev) / ws->clockKhz;
      prev = c.clk[i];
    } else {
      ms[i] = 0;
    }
  }
  if (barriers != nullptr) {
    *barriers = static_cast<int32_t>(c.barriers);
  }
  return count;
}

extern "C" int64_t mosp_default_delta(int64_t m, int32_t n, int64_t weightSum) {
  if (m <= 0 || n <= 0) {
    return 1;
  }
  const double avgW = static_cast<double>(weightSum) / static_cast<double>(m);
  const double avgDeg = static_cast<double>(m) / static_cast<double>(n);
  return std::max<int64_t>(1, static_cast<int64_t>(32.0 * avgW / avgDeg));
}

extern "C" int mosp_choose_packing(int32_t n, int64_t maxWeight,
                                   int32_t *parentBits, uint64_t *bound) {
  PackChoice pc{};
  if (!choose_packing(n, maxWeight, pc)) {
    return 0;
  }
  *parentBits = pc.bits;
  *bound = pc.bound;
  return 1;
}

// ---------------------------------------------------------------------------
// SOSP update / from scratch
// ---------------------------------------------------------------------------

extern "C" int mosp_sosp_update_multi(
    int32_t K, const mosp_csr *out, const mosp_csr *in,
    const mosp_changes *changes, int32_t source, const int64_t *delta,
    const int64_t *maxWeight, mosp_ws *ws, const int64_t *const *distIn,
    const int32_t *const *parentIn, int64_t *const *distOut,
    int32_t *const *parentOut, mosp_sosp_stats *stats, void *stream) {
  g_error.clear();
  if (K <= 0 || K > MOSP_MAX_OBJ) {
    return fail("Error: 1..32 objectives per launch.");
  }
  const int n = out[0].n;
  for (int k = 0; k < K; ++k) {
    if (stats != nullptr) {
      stats[k] = mosp_sosp_stats{};
      stats[k].packed_parents = 1;
    }
  }
  if (n == 0) {
    return 1;
  }
  if (source < 0 || source >= n) {
    return fail("Error: source vertex out of range.");
  }
  if (!ws_reserve(ws, n, K)) {
    return 0;
  }
  int hubs = 0;
  if (!caller_hubs(ws, out, in, K, as_stream(stream), hubs)) {
    return 0;
  }
  float ms = 0;
  // Packed words first, also where only their own limit bounds the
  // distances; a launch that meets a larger one wrote nothing and is redone
  // with the conservative packing.
  for (int attempt = 0; attempt < 2; ++attempt) {
    int anyFallback = 0, anyOptimistic = 0;
    for (int k = 0; k < K; ++k) {
      if (delta[k] <= 0) {
        return fail("Error: the near-far bucket width must be positive.");
      }
      PackChoice pc{};
      if (!choose_packing(n, maxWeight[k], pc, attempt == 0)) {
        return 0;
      }
      ObjDesc d = make_desc(out[k], &in[k], &changes[k], pc,
                            static_cast<u64>(delta[k]), ws->state[k]);
      d.distIn = distIn[k];
      d.parentIn = parentIn[k];
      d.distOut = distOut[k];
      d.parentOut = parentOut[k];
      ws->hDesc[k] = d;
      anyFallback |= pc.bits == 0 ? 1 : 0;
      anyOptimistic |= pc.optimistic ? 1 : 0;
    }
    LaunchArgs args{};
    args.objs = ws->dDesc;
    args.nobj = K;
    args.n = n;
    args.source = source;
    args.phases = PH_UPDATE;
    args.maxRounds = jump_rounds(n);
    args.anyFallback = anyFallback;
    args.anyOptimistic = anyOptimistic;
    args.hubs = hubs;
    if (!run_launch(ws, args, K, as_stream(stream), &ms)) {
      return 0;
    }
    if (!anyOptimistic || !launch_dropped(*ws->hCtl, K)) {
      break;
    }
  }
  for (int k = 0; k < K; ++k) {
    fill_stats(*ws->hCtl, k, ws->hDesc[k], ms, stats ? &stats[k] : nullptr);
  }
  return 1;
}

extern "C" int mosp_sosp_update(const mosp_csr *out, const mosp_csr *in,
                                const mosp_changes *changes, int32_t source,
                                int64_t delta, int64_t maxWeight, mosp_ws *ws,
                                int64_t *d_dist, int32_t *d_parent,
                                mosp_sosp_stats *stats, void *stream) {
  const int64_t *dIn[1] = {d_dist};
  const int32_t *pIn[1] = {d_parent};
  int64_t *dOut[1] = {d_dist};
  int32_t *pOut[1] = {d_parent};
  return mosp_sosp_update_multi(1, out, in, changes, source, &delta, &maxWeight,
                                ws, dIn, pIn, dOut, pOut, stats, stream);
}

extern "C" int mosp_sosp_from_scratch(const mosp_csr *out, int32_t source,
                                      int64_t delta, int64_t maxWeight,
                                      mosp_ws *ws, int64_t *d_dist,
                                      int32_t *d_parent, mosp_sosp_stats *stats,
                                      void *stream) {
  g_error.clear();
  const int n = out->n;
  if (stats != nullptr) {
    *stats = mosp_sosp_stats{};
    stats->packed_parents = 1;
  }
  if (n == 0) {
    return 1;
  }
  if (delta <= 0 || source < 0 || source >= n) {
    return fail("Error: invalid from-scratch parameters (delta > 0, source in range).");
  }
  if (!ws_reserve(ws, n, 1)) {
    return 0;
  }
  int hubs = 0;
  if (!caller_hubs(ws, out, nullptr, 1, as_stream(stream), hubs)) {
    return 0;
  }
  float ms = 0;
  ObjDesc d{};
  // Optimistic packed words first, as in mosp_sosp_update_multi.
  for (int attempt = 0; attempt < 2; ++attempt) {
    PackChoice pc{};
    if (!choose_packing(n, maxWeight, pc, attempt == 0)) {
      return 0;
    }
    d = make_desc(*out, nullptr, nullptr, pc, static_cast<u64>(delta), ws->state[0]);
    d.distOut = d_dist;
    d.parentOut = d_parent;
    ws->hDesc[0] = d;
    LaunchArgs args{};
    args.objs = ws->dDesc;
    args.nobj = 1;
    args.n = n;
    args.source = source;
    args.phases = PH_SCRATCH;
    args.maxRounds = jump_rounds(n);
    args.anyOptimistic = pc.optimistic ? 1 : 0;
    args.hubs = hubs;
    if (!run_launch(ws, args, 1, as_stream(stream), &ms)) {
      return 0;
    }
    if (!pc.optimistic || !launch_dropped(*ws->hCtl, 1)) {
      break;
    }
  }
  fill_stats(*ws->hCtl, 0, d, ms, stats);
  return 1;
}

// ---------------------------------------------------------------------------
// Combined graph + its search
// ---------------------------------------------------------------------------

extern "C" int64_t mosp_preference_scale(const int32_t *pref, int32_t nPref,
                                         int32_t K) {
  return preference_scale(pref, nPref, K);
}

extern "C" int64_t mosp_combined_edge_weight(uint32_t mask, const int32_t *pref,
                                             int32_t nPref, int32_t K,
                                             int64_t scale) {
  return combined_edge_weight(mask, pref, nPref, K, scale);
}

namespace {

// Fill the combined-step fields of `args` and descriptor slot `descIndex`.
int setup_combined(mosp_ws *ws, LaunchArgs &args, const int32_t *d_parents,
                   int n, int K, const int32_t *pref, int32_t nPref,
                   int64_t delta, int64_t *d_dist, int32_t *d_parent,
                   int descIndex, mosp_combine_stats *stats, PackChoice &pc) {
  if (K <= 0 || K > MOSP_MAX_OBJ || n <= 0) {
    return fail("Error: invalid combined-graph parameters.");
  }
  const int64_t scale = preference_scale(pref, nPref, K);
  if (scale == 0) {
    return fail("Error: invalid preference vector (need K values >= 1).");
  }
  if (!ws_reserve_combined(ws, n, K)) {
    return 0;
  }
  const int base = static_cast<int>(scale * (K + 1));
  args.cParents = d_parents;
  args.cK = K;
  args.cBase = base;
  for (int k = 0; k < K; ++k) {
    args.cTerms[k] = static_cast<int>(nPref == 0 ? scale : scale / pref[k]);
  }
  args.cDelta = delta;
  args.cDegree = ws->cDegree;
  args.cSlot = ws->cSlot;
  args.cRowPtr = ws->cRowPtr;
  args.cColInd = ws->cColInd;
  args.cWeights = ws->cWeights;
  args.cEdge = reinterpret_cast<int2 *>(ws->cColInd);  // K * n records over both arrays
  args.cEdge4 = ws->cEdge4;
  if (!choose_packing(n, base, pc)) {
    return 0;
  }
  mosp_csr view{n, 0, ws->cRowPtr, nullptr, nullptr};
  ObjDesc d = make_desc(view, nullptr, nullptr, pc, 0, ws->state[0]);
  d.outEdge = args.cEdge;
  d.outEdge4 = args.cEdge4;
  d.distOut = d_dist;
  d.parentOut = d_parent;
  ws->hDesc[descIndex] = d;
  args.cDesc = ws->dDesc + descIndex;
  if (stats != nullptr) {
    stats->scale = scale;
  }
  return 1;
}

void fill_combined_stats(const mosp_ws *ws, const ObjDesc &d, float ms,
                         mosp_combine_stats *stats) {
  if (stats == nullptr) {
    return;
  }
  const Ctl &c = *ws->hCtl;
  stats->n_edges = static_cast<int32_t>(c.combEdges);
  stats->delta = static_cast<int64_t>(c.combDelta);
  fill_stats(c, MOSP_MAX_OBJ, d, ms, &stats->search);
  stats->search.jump_rounds = 0;
  stats->search.invalidated = 0;
  if (c.clkSospEnd > c.clkStart && c.clkEnd >= c.clkSospEnd) {
    stats->sosp_phase_ms = static_cast<float>((c.clkSospEnd - c.clkStart) / ws->clockKhz);
    stats->combined_phase_ms = static_cast<float>((c.clkEnd - c.clkSospEnd) / ws->clockKhz);
  } else {
    stats->sosp_phase_ms = 0;
    stats->combined_phase_ms = static_cast<float>((c.clkEnd - c.clkStart) / ws->clockKhz);
  }
}

}  // namespace

extern "C" int mosp_combined_sosp(const int32_t *d_parents, int32_t n,
                                  int32_t K, const int32_t *pref,
                                  int32_t nPref, int32_t source, int64_t delta,
                                  mosp_ws *ws, int64_t *d_dist,
                                  int32_t *d_parent, mosp_combine_stats *stats,
                                  void *stream) {
  g_error.clear();
  if (stats != nullptr) {
    *stats = mosp_combine_stats{};
    stats->scale = 1;
  }
  if (K <= 0 || K > MOSP_MAX_OBJ || n <= 0 || source < 0 || source >= n) {
    return fail("Error: invalid combined-graph parameters.");
  }
  if (!ws_reserve(ws, n, 1)) {
    return 0;
  }
  LaunchArgs args{};
  PackChoice pc{};
  if (!setup_combined(ws, args, d_parents, n, K, pref, nPref, delta, d_dist,
                      d_parent, 0, stats, pc)) {
    return 0;
  }
  args.objs = ws->dDesc;
  args.nobj = 0;
  args.n = n;
  args.source = source;
  args.phases = PH_COMBINED;
  args.maxRounds = jump_rounds(n);
  // The combined graph is built inside the launch: its degrees are unknown
  // here, so the variant that handles any degree runs.
  args.hubs = 1;
  float ms = 0;
  if (!run_launch(ws, args, 1, as_stream(stream), &ms)) {
    return 0;
  }
  fill_combined_stats(ws, ws->hDesc[0], ms, stats);
  return 1;
}

// ---------------------------------------------------------------------------
// MOSP path costs on the device graph (device_graph.hip)
// ---------------------------------------------------------------------------


namespace {

int grid_for(long long work) {
  return static_cast<int>((std::max(work, 1LL) + BLOCK - 1) / BLOCK);
}

// Path costs: every vertex starts with its tree edge's K weights (first (p,v)
// edge of row p) and its parent as ancestor, then pointer jumping adds the
// ancestor's partial sum and jumps (ping-pong buffers), ceil(log2 n)+1
// rounds.
__global__ void path_cost_init(int n, int K, int m, const int *rowPtr,
                               const int *colInd, const int *wcols,
                               const int *parent, int source, int *anc,
                               long long *cost, int *missing) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) {
    return;
  }
  const size_t at = static_cast<size_t>(v) * K;
  if (v == source) {
    anc[v] = -1;
    for (int k = 0; k < K; ++k) cost[at + k] = 0;
    return;
  }
  const int p = parent[v];
  if (p < 0) {
    anc[v] = -1;
    for (int k = 0; k < K; ++k) cost[at + k] = DISTANCE_INF;
    return;
  }
  int edge = -1;
  for (int e = rowPtr[p]; e < rowPtr[p + 1]; ++e) {
    if (colInd[e] == v) {
      edge = e;
      break;
    }
  }
  if (edge < 0) {
    atomicOr(missing, 1);
    anc[v] = -1;
    for (int k = 0; k < K; ++k) cost[at + k] = DISTANCE_INF;
    return;
  }
  anc[v] = p;
  for (int k = 0; k < K; ++k) {
    cost[at + k] = wcols[static_cast<size_t>(k) * m + edge];
  }
}

__global__ void path_cost_jump(int n, int K, const int *ancIn,
                               const long long *costIn, int *ancOut,
                               long long *costOut) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) {
    return;
  }
  const size_t at = static_cast<size_t>(v) * K;
  const int a = ancIn[v];
  if (a < 0) {
    ancOut[v] = -1;
    for (int k = 0; k < K; ++k) costOut[at + k] = costIn[at + k];
    return;
  }
  const size_t aa = static_cast<size_t>(a) * K;
  for (int k = 0; k < K; ++k) {
    costOut[at + k] = costIn[at + k] + costIn[aa + k];
  }
  ancOut[v] = ancIn[a];
}

}  // namespace

extern "C" int mosp_path_costs_gpu(const mosp_device_graph *g,
                                   const int32_t *d_parent, int32_t source,
                                   int64_t *d_costs, void *stream) {
  g_error.clear();
  const int n = g->n, K = g->K;
  hipStream_t s = as_stream(stream);
  int *ancA = nullptr, *ancB = nullptr, *missing = nullptr;
  long long *costB = nullptr;
  HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&ancA), sizeof(int) * n, s));
  HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&ancB), sizeof(int) * n, s));
  HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&costB),
                         sizeof(long long) * static_cast<size_t>(n) * K, s));
  HIP_TRY(hipMallocAsync(reinterpret_cast<void **>(&missing), sizeof(int), s));
  HIP_TRY(hipMemsetAsync(missing, 0, sizeof(int), s));
  long long *costA = reinterpret_cast<long long *>(d_costs);
  hipLaunchKernelGGL(path_cost_init, dim3(grid_for(n)), dim3(BLOCK), 0, s, n, K,
                     static_cast<int>(g->stride), g->outRow, g->outCol, g->outW, d_parent, source, ancA,
                     costA, missing);
  const int rounds = jump_rounds(n);
  for (int r = 0; r < rounds; ++r) {
    hipLaunchKernelGGL(path_cost_jump, dim3(grid_for(n)), dim3(BLOCK), 0, s, n,
                       K, ancA, costA, ancB, costB);
    std::swap(ancA, ancB);
    std::swap(costA, costB);
  }
  if (costA != reinterpret_cast<long long *>(d_costs)) {
    HIP_TRY(hipMemcpyAsync(d_costs, costA,
                           sizeof(long long) * static_cast<size_t>(n) * K,
                           hipMemcpyDeviceToDevice, s));
    std::swap(costA, costB);  // costB is the scratch buffer again
  }
  int hostMissing = 0;
  HIP_TRY(hipMemcpyAsync(&hostMissing, missing, sizeof(int),
                         hipMemcpyDeviceToHost, s));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipFreeAsync(ancA, s));
  HIP_TRY(hipFreeAsync(ancB, s));
  HIP_TRY(hipFreeAsync(costB, s));
  HIP_TRY(hipFreeAsync(missing, s));
  HIP_TRY(hipStreamSynchronize(s));
  if (hostMissing) {
    return fail("Error: a MOSP tree edge is missing from the updated graph.");
  }
  return 1;
}

// ---------------------------------------------------------------------------
// Fused device pipeline: K updates + combined graph + search, one launch.
// ---------------------------------------------------------------------------

extern "C" int mosp_update_fused(const mosp_device_graph *g, int32_t K,
                                 const mosp_changes *changes, int32_t source,
                                 const int64_t *delta, const int64_t *maxWeight,
                                 const int32_t *pref, int32_t nPref,
                                 mosp_ws *ws, const int64_t *d_dist_in,
                                 const int32_t *d_parents_in, int64_t *d_dist,
                                 int32_t *d_parents, int64_t *d_comb_dist,
                                 int32_t *d_comb_parent, mosp_sosp_stats *stats,
                                 mosp_combine_stats *cstats, void *stream) {
  g_error.clear();
  const int n = g->n;
  if (K <= 0 || K > g->K) {
    return fail("Error: invalid number of objectives.");
  }
  if (source < 0 || source >= n) {
    return fail("Error: source vertex out of range.");
  }
  if (cstats != nullptr) {
    *cstats = mosp_combine_stats{};
    cstats->scale = 1;
  }
  if (!ws_reserve(ws, n, K)) {
    return 0;
  }
  const int64_t *distIn = d_dist_in != nullptr ? d_dist_in : d_dist;
  const int32_t *parIn = d_parents_in != nullptr ? d_parents_in : d_parents;
  for (int k = 0; k < K; ++k) {
    if (delta[k] <= 0) {
      return fail("Error: the near-far bucket width must be positive.");
    }
  }
  // The K objectives' descriptors; `optimistic`: packed words also where
  // only their own limit bounds the distances.
  // Graphs without hubs run the asynchronous loop: it reads the records
  // with row ranges.
  const bool rec4 = ws->async && g->maxDegree <= HUB_DEGREE;
  if (rec4 && !graph_records4(g, as_stream(stream))) {
    return 0;
  }
  auto describe = [&](bool optimistic, int &anyFallback, int &anyOptimistic) {
    anyFallback = anyOptimistic = 0;
    for (int k = 0; k < K; ++k) {
      PackChoice pc{};
      if (!choose_packing(n, maxWeight[k], pc, optimistic)) {
        return 0;
      }
      mosp_csr out{}, in{};
      mosp_device_graph_view(g, k, &out, &in);
      ObjDesc d = make_desc(out, &in, &changes[k], pc, static_cast<u64>(delta[k]),
                            ws->state[k]);
      graph_edges(g, k, &d.outEdge, &d.inEdge);
      if (rec4) {
        d.outEdge4 = g->outEdge4 + static_cast<size_t>(k) * g->stride;
      }
      const size_t off = static_cast<size_t>(k) * n;
      d.distIn = distIn + off;
      d.parentIn = parIn + off;
      d.distOut = d_dist + off;
      d.parentOut = d_parents + off;
      ws->hDesc[k] = d;
      anyFallback |= pc.bits == 0 ? 1 : 0;
      anyOptimistic |= pc.optimistic ? 1 : 0;
    }
    return 1;
  };
  int anyFallback = 0, anyOptimistic = 0;
  if (!describe(true, anyFallback, anyOptimistic)) {
    return 0;
  }
  LaunchArgs args{};
  PackChoice cpc{};
  if (!setup_combined(ws, args, d_parents, n, K, pref, nPref, 0, d_comb_dist,
                      d_comb_parent, K, cstats, cpc)) {
    return 0;
  }
  // Graphs without hubs: the combined graph as a mask over the graph's
  // out-edges (every tree edge is one), searched by the asynchronous loop.
  ws->cMaskIndex = -1;
  if (rec4 && ws->combMask && cpc.bits > 0) {
    const size_t m = static_cast<size_t>(std::max(g->m, 1));
    if (m > ws->cMaskCap) {
      (void)hipFree(ws->cMask);
      ws->cMask = nullptr;
      ws->cMaskCap = 0;
      HIP_TRY(hipMalloc(&ws->cMask, m * sizeof(int4)));
      ws->cMaskCap = m;
    }
    mosp_csr out{}, in{};
    mosp_device_graph_view(g, 0, &out, &in);
    args.cMaskRec = ws->cMask;
    args.cMaskSrc = g->outEdge4;
    args.cMaskRow = out.row_ptr;
    // The combined CSR's arrays (8 K n bytes) are unused by a masked launch.
    args.cPar4 = K >= 2 && K <= 4 ? reinterpret_cast<int4 *>(ws->cColInd) : nullptr;
    ws->cCsrDesc = ws->hDesc[K];
    ws->cMaskDesc = ws->hDesc[K];
    ws->cMaskDesc.outRow = out.row_ptr;
    ws->cMaskDesc.outEdge = nullptr;
    ws->cMaskDesc.outEdge4 = ws->cMask;
    ws->cMaskIndex = K;
  }
  args.stage = reinterpret_cast<u64 *>(ws->cColInd);  // 8 * K * n bytes, free until Step 3
  args.objs = ws->dDesc;
  args.nobj = K;
  args.n = n;
  args.source = source;
  args.phases = PH_UPDATE | PH_COMBINED;
  args.maxRounds = jump_rounds(n);
  args.anyFallback = anyFallback;
  args.anyOptimistic = anyOptimistic;
  // A combined-graph vertex has at most its out-degree in the graph.
  args.hubs = g->maxDegree > HUB_DEGREE ? 1 : 0;
  float ms = 0;
  if (!run_launch(ws, args, K + 1, as_stream(stream), &ms)) {
    return 0;
  }
  if (anyOptimistic && launch_dropped(*ws->hCtl, K)) {
    // A distance past the optimistic packing's limit: nothing was written;
    // again with the conservative packing.
    if (!describe(false, anyFallback, anyOptimistic)) {
      return 0;
    }
    args.anyFallback = anyFallback;
    args.anyOptimistic = 0;
    if (!run_launch(ws, args, K + 1, as_stream(stream), &ms)) {
      return 0;
    }
  }
  for (int k = 0; k < K; ++k) {
    fill_stats(*ws->hCtl, k, ws->hDesc[k], ms, stats ? &stats[k] : nullptr);
  }
  fill_combined_stats(ws, ws->hDesc[K], ms, cstats);
  return 1;
}

// ---------------------------------------------------------------------------
// mosp_sosp_update without allocation: the same launch as mosp_sosp_update,
// its scratch in one file-scope device array instead of a workspace.
// ---------------------------------------------------------------------------

namespace {

constexpr int NOALLOC_MAX_VERTICES = MOSP_NOALLOC_MAX_VERTICES;

// Byte layout of the array: control block, one descriptor, the generation
// word, the degree word, then one objective's state for up to
// NOALLOC_MAX_VERTICES vertices.
constexpr size_t NA_CTL = 0;
constexpr size_t NA_DESC = (sizeof(Ctl) + 255) / 256 * 256;
constexpr size_t NA_WORDS = NA_DESC + (sizeof(ObjDesc) + 255) / 256 * 256;
constexpr size_t NA_STATE = NA_WORDS + 256;
constexpr size_t NA_STATE_BYTES =
    (static_cast<size_t>(NOALLOC_MAX_VERTICES) * sizeof(u64) + 255) / 256 * 256 +
    7 * ((static_cast<size_t>(NOALLOC_MAX_VERTICES) * sizeof(int) + 255) / 256 * 256) +
    (2 * static_cast<size_t>(HUB_LIST) * sizeof(int) + 255) / 256 * 256;
constexpr size_t NA_BYTES = NA_STATE + NA_STATE_BYTES;

// The scratch of every mosp_sosp_update_noalloc call on a device (each
// device holds its own copy until the process ends).  Only its address as
// returned by hipGetSymbolAddress is used on the host.
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Per device: the array's device address and the launch shape, taken once
// (the entry point's first call on that device).
struct NoAllocDevice {
  char *ptr = nullptr;
  int gridBlocks = 0;
  double clockKhz = 100000;
};
constexpr int NA_DEVICES = 64;
NoAllocDevice g_noalloc[NA_DEVICES];
std::mutex g_noalloc_mu;

int noalloc_device(NoAllocDevice **out) {
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  if (device < 0 || device >= NA_DEVICES) {
    return fail("Error: device " + std::to_string(device) + " is beyond the no-allocation table.");
  }
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  NoAllocDevice &d = g_noalloc[device];
  if (d.ptr == nullptr) {
    void *ptr = nullptr;
    HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    int cus = 0, perCu = 0, clockKhz = 0;
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &perCu, kernel_shape(1024).persistent(true), BLOCK, 0));
    if (perCu < 1) {
      return fail("Error: the persistent kernel does not fit on a CU.");
    }
    if (hipDeviceGetAttribute(&clockKhz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
        clockKhz > 0) {
      d.clockKhz = clockKhz;
    }
    d.gridBlocks = std::max(1, std::min(cus, perCu * cus));
    d.ptr = static_cast<char *>(ptr);
  }
  *out = &d;
  return 1;
}

// Stamps and far marks of the first n vertices, and the generation word,
// back to zero: the last call on this device left them as it ended (or in
// any state, had it given up).
__global__ void noalloc_clear(int *stamp, int *inFar, int n, int *gen) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    stamp[i] = 0;
    inFar[i] = 0;
  }
  if (i == 0) {
    *gen = 0;
  }
}

}  // namespace

extern "C" int mosp_sosp_update_noalloc(const mosp_csr *out, const mosp_csr *in,
                                        const mosp_changes *changes, int32_t source,
                                        int64_t delta, int64_t maxWeight, int64_t *d_dist,
                                        int32_t *d_parent, mosp_sosp_stats *stats,
                                        void *streamArg) {
  g_error.clear();
  if (stats != nullptr) {
    *stats = mosp_sosp_stats{};
    stats->packed_parents = 1;
  }
  if (out == nullptr || in == nullptr || changes == nullptr) {
    return fail("Error: missing graph or change arguments.");
  }
  const int n = out->n;
  if (n == 0) {
    return 1;
  }
  if (n < 0 || n > NOALLOC_MAX_VERTICES) {
    return fail("Error: the no-allocation update holds at most " +
                std::to_string(NOALLOC_MAX_VERTICES) + " vertices; this graph has " +
                std::to_string(n) + ".");
  }
  if (source < 0 || source >= n) {
    return fail("Error: source vertex out of range.");
  }
  if (delta <= 0) {
    return fail("Error: the near-far bucket width must be positive.");
  }
  PackChoice pc{};
  if (!choose_packing(n, maxWeight, pc, true)) {
    return 0;
  }
  NoAllocDevice *dev = nullptr;
  if (!noalloc_device(&dev)) {
    return 0;
  }
  hipStream_t stream = as_stream(streamArg);
  char *base = dev->ptr;
  Ctl *dCtl = reinterpret_cast<Ctl *>(base + NA_CTL);
  ObjDesc *dDesc = reinterpret_cast<ObjDesc *>(base + NA_DESC);
  int *genDev = reinterpret_cast<int *>(base + NA_WORDS);
  unsigned *degreeWord = reinterpret_cast<unsigned *>(base + NA_WORDS + 128);
  char *p = base + NA_STATE;
  const ObjState st = carve_state(p, NOALLOC_MAX_VERTICES, 0);

  // Variant choice as in mosp_sosp_update: any vertex of more than
  // HUB_DEGREE out- or in-edges.
  int hubs = 0;
  for (const int *row : {out->row_ptr, in->row_ptr}) {
    int degree = 0;
    if (!graph_max_degree(row, nullptr, n, degreeWord, stream, &degree)) {
      return 0;
    }
    hubs |= degree > HUB_DEGREE ? 1 : 0;
  }
  const int forced = hubs_override();
  if (forced >= 0) {
    hubs = forced;
  }

  static thread_local Ctl hCtl;  // host copy of the control block (no allocation per call)
  ObjDesc d{};
  // Optimistic packed words first, as in mosp_sosp_update_multi; again with
  // the conservative packing when the launch met a larger distance.
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (attempt > 0 && !choose_packing(n, maxWeight, pc, false)) {
      return 0;
    }
    d = make_desc(*out, in, changes, pc, static_cast<u64>(delta), st);
    d.distIn = d_dist;
    d.parentIn = d_parent;
    d.distOut = d_dist;
    d.parentOut = d_parent;
    LaunchArgs args{};
    args.objs = dDesc;
    args.nobj = 1;
    args.n = n;
    args.source = source;
    args.phases = PH_UPDATE;
    args.maxRounds = jump_rounds(n);
    args.anyFallback = pc.bits == 0 ? 1 : 0;
    args.anyOptimistic = pc.optimistic ? 1 : 0;
    args.laneBits = 2;
    args.hopsUpdate = kHopsUpdate;
    args.hopsSearch = kHopsSearch;
    local_hops_override(args.hopsUpdate, args.hopsSearch);
    args.listCap = list_cap_override();
    args.initPush = init_push_override();
    args.hubEdges = hub_edges_override();
    args.hubs = hubs;
    args.genDev = genDev;
    args.ctl = dCtl;
    args.abortWord = &dCtl->gaveUp;
    args.barrierTicks = barrier_ticks(barrier_timeout_ms_default(), dev->clockKhz);
    args.cDeltaScale = 1;

    HIP_TRY(hipMemcpyAsync(dDesc, &d, sizeof(ObjDesc), hipMemcpyHostToDevice, stream));
    hipLaunchKernelGGL(noalloc_clear, dim3((n + BLOCK - 1) / BLOCK), dim3(BLOCK), 0, stream,
                       st.stamp, st.inFar, n, genDev);
    HIP_TRY(hipGetLastError());
    HIP_TRY(launch_ctl_reset(dCtl, stream));
    void *params[] = {&args};
    HIP_TRY(launch_persistent(kernel_shape(1024).persistent(hubs != 0), dim3(dev->gridBlocks),
                              dim3(BLOCK), params, stream));
    HIP_TRY(hipMemcpyAsync(&hCtl, dCtl, CTL_READBACK, hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
    if (hCtl.gaveUp != 0) {
      return fail("Error: a grid barrier of the persistent launch gave up waiting (its workgroups "
                  "were not all resident); the tree was not updated.");
    }
    if (!input_tree_ok(hCtl.obj[0].overflow, "")) {
      return 0;
    }
    if (!pc.optimistic || hCtl.obj[0].dropped == 0) {
      break;
    }
  }
  const float ms = hCtl.clkEnd > hCtl.clkStart
                       ? static_cast<float>((hCtl.clkEnd - hCtl.clkStart) / dev->clockKhz)
                       : 0.0f;
  fill_stats(hCtl, 0, d, ms, stats);
  return 1;
}
