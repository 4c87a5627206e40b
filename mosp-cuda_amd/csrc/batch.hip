// batch.hip -- batched launches (C ABI mosp_batch_*): many independent update
// instances, each with its own graph, K, batch, source and preferences, in
// ONE persistent launch of sosp_batched (sosp_kernel.hip).
//
// Each instance computes exactly what mosp_update_fused computes for it (K
// objective updates + combined graph + its search); only the schedule is
// shared.  The launch's workgroups are split into groups: one group per
// instance, sized by a cost model so that all instances finish together
// (model_sizes), or -- with more instances than workgroups -- single-
// workgroup groups that run their instances in turn (longest first).  Groups that fit on one XCD are
// placed there (workgroup b runs on XCD b % 8), so a small instance's
// working set stays in one L2.  Every instance synchronizes its group
// through its own control block, so groups never wait for each other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>
#include <queue>
#include <string>
#include <vector>

#include "engine.hpp"
#include "engine_internal.hpp"
#include "host_api.hpp"
#include "mosp_hip.h"

using namespace mosp;
using namespace mosp::detail;

namespace {

struct Instance {
  mosp_batch_instance spec{};
  int n = 0;
  int K = 0;
  std::vector<mosp_changes> changes;
  std::vector<mosp_csr> out, in;
  std::vector<int64_t> delta, maxWeight;
  std::vector<int32_t> pref;
  std::vector<PackChoice> pack;
  PackChoice combPack{};
  int64_t scale = 1;
  size_t slabOffset = 0;  // K objective states, then the combined workspace
  int descOffset = 0;     // K objective descriptors + 1 combined
  int groupSize = 0;
};

size_t combined_bytes(int n, int K, int partials) {
  const size_t nn = static_cast<size_t>(std::max(n, 1));
  return 2 * align256((nn + 1) * sizeof(int)) +
         2 * align256(nn * static_cast<size_t>(K) * sizeof(int)) +
         align256(static_cast<size_t>(partials) * sizeof(int));
}

}  // namespace

struct mosp_batch {
  std::vector<Instance> inst;
  int capInstances = 0;   // allocation capacities (mosp_batch_reset reuses them)
  int capDescs = 0;
  size_t capSlab = 0;
  int cus = 0;
  int maxBlocks = 0;  // co-resident workgroups of sosp_batched (of the shape)
  int shapeThreads = 1024;  // workgroup shape: 1024 or 512 threads
  int partialsCap = 0;      // combined-scan partials per instance: either shape's largest grid
  int queueCap = 0;         // LDS queue entries of the push loop (0: all)
  int xcds = 1;
  int workgroups = 0;
  int laneBits = 2;
  int hopsUpdate = kHopsUpdate;  // local relaxation rounds (MOSP_LOCAL_HOPS)
  int hopsSearch = kHopsSearch;
  double deltaScale = 1.0;
  int xcdAware = 1;
  double latencyItems = 400;  // schedule cost model: a / b (see model_sizes)
  double clockKhz = 100000;
  char *slab = nullptr;
  Ctl *dCtl = nullptr;  // one control block per instance
  int *dGen = nullptr;  // stamp generation per instance
  ObjDesc *dDesc = nullptr;
  LaunchArgs *dArgs = nullptr;  // schedule order
  BatchGroup *dGroups = nullptr;
  int *dGroupOf = nullptr;
  int *dRankOf = nullptr;
  Ctl *hCtl = nullptr;  // pinned read-back of one control block
  std::vector<ObjDesc> hDesc;
  std::vector<LaunchArgs> hArgs;
  std::vector<BatchGroup> hGroups;
  std::vector<int> groupOf, rankOf;
  std::vector<int> argsIndex;  // instance -> position in hArgs
  int descCount = 0;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graphMode = 0;
  int hubs = 0;  // some instance's graph has a vertex of degree > HUB_DEGREE
  bool dirty = true;
  unsigned *dAbort = nullptr;  // set by a workgroup whose barrier wait ran out
  double barrierMs = -1;       // longest barrier wait (< 0: the default)
  bool scrub = false;          // a run gave up: clear the scratch before the next
  size_t usedSlab = 0;         // scratch bytes of the current instances
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  float lastMs = 0;
  // Asynchronous near-far for graphs without hubs: one control block per
  // instance; pool segments in the instances' states, whose slots must hold
  // -1 (refilled after a run gave up).
  unsigned asyncFlags = ASYNC_HUNGRY;
  AsyncCtl *dAsync = nullptr;
  std::vector<std::pair<int *, size_t>> pools;
};

namespace {

void drop_graph(mosp_batch *b) {
  if (b->exec != nullptr) (void)hipGraphExecDestroy(b->exec);
  if (b->graph != nullptr) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  b->graphMode = 0;
}

void release(mosp_batch *b) {
  drop_graph(b);
  (void)hipFree(b->slab);
  (void)hipFree(b->dCtl);
  (void)hipFree(b->dGen);
  (void)hipFree(b->dAbort);
  (void)hipFree(b->dDesc);
  (void)hipFree(b->dArgs);
  (void)hipFree(b->dGroups);
  (void)hipFree(b->dGroupOf);
  (void)hipFree(b->dRankOf);
  (void)hipFree(b->dAsync);
  if (b->hCtl != nullptr) (void)hipHostFree(b->hCtl);
  if (b->stream != nullptr) (void)hipStreamDestroy(b->stream);
  if (b->ev0 != nullptr) (void)hipEventDestroy(b->ev0);
  if (b->ev1 != nullptr) (void)hipEventDestroy(b->ev1);
}

int describe(Instance &I, const mosp_batch_instance &spec, int index) {
  const std::string where = "instance " + std::to_string(index) + ": ";
  I.spec = spec;
  if (spec.graph == nullptr || spec.changes == nullptr || spec.delta == nullptr ||
      spec.max_weight == nullptr) {
    return fail("Error: " + where + "missing graph, changes, delta or max_weight.");
  }
  const int kg = mosp_device_graph_objectives(spec.graph);
  I.K = spec.k;
  if (I.K <= 0 || I.K > kg || I.K > MOSP_MAX_OBJ) {
    return fail("Error: " + where + "invalid number of objectives.");
  }
  I.out.resize(I.K);
  I.in.resize(I.K);
  for (int k = 0; k < I.K; ++k) {
    if (!mosp_device_graph_view(spec.graph, k, &I.out[k], &I.in[k])) {
      return 0;
    }
  }
  I.n = I.out[0].n;
  if (I.n <= 0 || spec.source < 0 || spec.source >= I.n) {
    return fail("Error: " + where + "source vertex out of range.");
  }
  if (spec.d_dist_in == nullptr || spec.d_parents_in == nullptr || spec.d_dist == nullptr ||
      spec.d_parents == nullptr || spec.d_comb_dist == nullptr ||
      spec.d_comb_parent == nullptr) {
    return fail("Error: " + where + "missing tree buffers.");
  }
  I.changes.assign(spec.changes, spec.changes + I.K);
  I.delta.assign(spec.delta, spec.delta + I.K);
  I.maxWeight.assign(spec.max_weight, spec.max_weight + I.K);
  I.pref.assign(spec.pref, spec.pref + std::max(spec.n_pref, 0));
  I.pack.resize(I.K);
  for (int k = 0; k < I.K; ++k) {
    if (I.delta[k] <= 0) {
      return fail("Error: " + where + "the near-far bucket width must be positive.");
    }
    if (!choose_packing(I.n, I.maxWeight[k], I.pack[k])) {
      return 0;
    }
  }
  I.scale = preference_scale(I.pref.data(), static_cast<int32_t>(I.pref.size()), I.K);
  if (I.scale == 0) {
    return fail("Error: " + where + "invalid preference vector (need K values >= 1).");
  }
  if (!choose_packing(I.n, I.scale * (I.K + 1), I.combPack)) {
    return 0;
  }
  return 1;
}

// Device attributes, and allocations for up to maxInstances instances, maxDescs
// descriptors and slabBytes of scratch.
int setup(mosp_batch *b, int maxInstances, int maxDescs, size_t slabBytes) {
  int device = 0, cooperative = 0, perCu = 0, clockKhz = 0;
  HIP_TRY(hipGetDevice(&device));
  HIP_TRY(hipDeviceGetAttribute(&cooperative, hipDeviceAttributeCooperativeLaunch, device));
  if (!cooperative) {
    return fail("Error: the GPU does not support cooperative launches.");
  }
  HIP_TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, device));
  if (const char *env = std::getenv("MOSP_WG_THREADS")) {
    b->shapeThreads = std::atoi(env) == 512 ? 512 : 1024;
  }
  if (const char *env = std::getenv("MOSP_QUEUE_CAP")) {
    b->queueCap = std::max(0, std::min(MOSP_QUEUE_CAP_MAX, std::atoi(env)));
  }
  b->maxBlocks = shape_max_blocks(kernel_shape(b->shapeThreads), true);
  b->partialsCap = std::max(shape_max_blocks(kernel_shape(1024), true),
                            shape_max_blocks(kernel_shape(512), true));
  (void)perCu;
  if (b->maxBlocks < 1) {
    return fail("Error: the batched kernel does not fit on a CU.");
  }
  if (hipDeviceGetAttribute(&clockKhz, hipDeviceAttributeWallClockRate, device) == hipSuccess &&
      clockKhz > 0) {
    b->clockKhz = clockKhz;
  }
  // MI300-class parts dispatch workgroups round-robin over 8 XCDs.
  b->xcds = b->cus % 8 == 0 && b->cus >= 64 ? 8 : 1;
  if (const char *env = std::getenv("MOSP_XCDS")) {
    b->xcds = std::max(1, std::atoi(env));
  }
  local_hops_override(b->hopsUpdate, b->hopsSearch);
  if (const char *env = std::getenv("MOSP_ASYNC_FLAGS")) {
    b->asyncFlags = static_cast<unsigned>(std::atoi(env)) & ASYNC_FLAG_MASK;
  }
  b->workgroups = b->maxBlocks;  // every co-resident workgroup of the shape
  maxInstances = std::max(maxInstances, 1);
  maxDescs = std::max(maxDescs, 1);
  b->capInstances = maxInstances;
  b->capDescs = maxDescs;
  b->capSlab = std::max<size_t>(slabBytes, 256);
  HIP_TRY(hipMalloc(&b->slab, b->capSlab));
  HIP_TRY(hipMalloc(&b->dCtl, sizeof(Ctl) * maxInstances));
  HIP_TRY(hipMalloc(&b->dAsync, sizeof(AsyncCtl) * maxInstances));
  HIP_TRY(hipMalloc(&b->dGen, sizeof(int) * maxInstances));
  HIP_TRY(hipMalloc(&b->dAbort, sizeof(unsigned)));
  HIP_TRY(hipMemset(b->dAbort, 0, sizeof(unsigned)));
  HIP_TRY(hipMalloc(&b->dDesc, sizeof(ObjDesc) * maxDescs));
  HIP_TRY(hipMalloc(&b->dArgs, sizeof(LaunchArgs) * maxInstances));
  HIP_TRY(hipMalloc(&b->dGroups, sizeof(BatchGroup) * std::max(maxInstances, b->partialsCap)));
  HIP_TRY(hipMalloc(&b->dGroupOf, sizeof(int) * b->partialsCap));
  HIP_TRY(hipMalloc(&b->dRankOf, sizeof(int) * b->partialsCap));
  HIP_TRY(hipHostMalloc(&b->hCtl, sizeof(Ctl), hipHostMallocDefault));
  HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreate(&b->ev0));
  HIP_TRY(hipEventCreate(&b->ev1));
  return 1;
}

// Slab bytes of one instance: K objective states + the combined workspace.
size_t instance_bytes(int n, int K, int partials) {
  return static_cast<size_t>(K) * state_bytes(n, pool_seg_cap(n)) +
         combined_bytes(n, K, partials);
}

// Describe `count` instances into b (within its capacity) and clear the
// scratch they use (stamps and far marks start at zero;
// generations restart).
int load(mosp_batch *b, const mosp_batch_instance *spec, int count) {
  std::vector<Instance> inst(count);
  size_t slabBytes = 0;
  int descs = 0;
  for (int i = 0; i < count; ++i) {
    Instance &I = inst[i];
    if (!describe(I, spec[i], i)) {
      return 0;
    }
    I.slabOffset = slabBytes;
    slabBytes += instance_bytes(I.n, I.K, b->partialsCap);
    I.descOffset = descs;
    descs += I.K + 1;
  }
  if (count > b->capInstances || descs > b->capDescs || slabBytes > b->capSlab) {
    return fail("Error: the batch needs " + std::to_string(slabBytes) + " bytes of scratch for " +
                std::to_string(count) + " instances; this batch holds " +
                std::to_string(b->capSlab) + " bytes for " + std::to_string(b->capInstances) +
                ".");
  }
  b->inst = std::move(inst);
  b->descCount = descs;
  b->hubs = 0;
  for (int i = 0; i < count; ++i) {
    b->hubs |= spec[i].graph->maxDegree > HUB_DEGREE ? 1 : 0;
  }
  if (hubs_override() >= 0) b->hubs = hubs_override();
  HIP_TRY(hipMemsetAsync(b->slab, 0, std::max<size_t>(slabBytes, 256), b->stream));
  HIP_TRY(hipMemsetAsync(b->dGen, 0, sizeof(int) * count, b->stream));
  b->usedSlab = std::max<size_t>(slabBytes, 256);
  b->scrub = false;
  HIP_TRY(hipStreamSynchronize(b->stream));
  b->dirty = true;
  return 1;
}

// Cost model of one instance on g workgroups: T(g) = a * I + b * W / g with
// I ~ 0.2 sqrt(n) + 10 near-far iterations (a grid's hop diameter, measured
// ~0.17 sqrt(n) over the K searches and the combined one) and W ~ K n
// relaxed vertices; latencyItems = a / b (vertices a workgroup relaxes in
// one iteration's latency).  Sizes for equal finishing times: g = W / (T' -
// I * latencyItems) with T' found by bisection so that the sizes fill the
// launch.
std::vector<int> model_sizes(const std::vector<double> &iters, const std::vector<double> &work,
                             int G, double latencyItems) {
  const int count = static_cast<int>(work.size());
  auto sizes_at = [&](double T, std::vector<double> &g) {
    double sum = 0;
    for (int i = 0; i < count; ++i) {
      const double slack = T - iters[i] * latencyItems;
      g[i] = slack > 0 ? std::max(1.0, work[i] / slack) : 1e30;
      sum += g[i];
    }
    return sum;
  };
  std::vector<double> g(count);
  double lo = 0, hi = 1;
  for (int i = 0; i < count; ++i) {
    lo = std::max(lo, iters[i] * latencyItems);
    hi = std::max(hi, iters[i] * latencyItems + work[i]);
  }
  while (sizes_at(hi, g) > G) hi *= 2;
  for (int step = 0; step < 100; ++step) {
    const double mid = 0.5 * (lo + hi);
    (sizes_at(mid, g) > G ? lo : hi) = mid;
  }
  sizes_at(hi, g);
  std::vector<int> sizes(count);
  int sum = 0;
  for (int i = 0; i < count; ++i) {
    sizes[i] = std::max(1, static_cast<int>(g[i]));
    sum += sizes[i];
  }
  auto finish = [&](int i, int size) { return iters[i] * latencyItems + work[i] / size; };
  while (sum > G) {  // take from the group that loses least
    int best = -1;
    for (int i = 0; i < count; ++i) {
      if (sizes[i] > 1 && (best < 0 || finish(i, sizes[i] - 1) < finish(best, sizes[best] - 1))) {
        best = i;
      }
    }
    --sizes[best];
    --sum;
  }
  while (sum < G) {  // give to the group that finishes last
    int best = 0;
    for (int i = 1; i < count; ++i) {
      if (finish(i, sizes[i]) > finish(best, sizes[best])) best = i;
    }
    ++sizes[best];
    ++sum;
  }
  return sizes;
}

// Groups and their placement on the launch's workgroups.
void build_schedule(mosp_batch *b) {
  const int G = b->workgroups;
  const int count = static_cast<int>(b->inst.size());
  std::vector<double> iters(count), work(count), total(count);
  for (int i = 0; i < count; ++i) {
    const Instance &I = b->inst[i];
    iters[i] = 0.2 * std::sqrt(static_cast<double>(I.n)) + 10.0;
    work[i] = static_cast<double>(I.K) * I.n;
    total[i] = iters[i] * b->latencyItems + work[i];  // time on one workgroup
  }
  std::vector<std::vector<int>> members;
  std::vector<int> sizes;
  if (count <= G) {
    sizes = model_sizes(iters, work, G, b->latencyItems);
    members.resize(count);
    for (int i = 0; i < count; ++i) members[i] = {i};
  } else {
    // More instances than workgroups: G single-workgroup groups, longest
    // processing time first onto the least loaded group.
    sizes.assign(G, 1);
    members.resize(G);
    std::vector<int> order(count);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return total[x] > total[y]; });
    using Load = std::pair<double, int>;
    std::priority_queue<Load, std::vector<Load>, std::greater<Load>> heap;
    for (int g = 0; g < G; ++g) heap.push({0.0, g});
    for (int i : order) {
      Load l = heap.top();
      heap.pop();
      members[l.second].push_back(i);
      heap.push({l.first + total[i], l.second});
    }
  }
  const int groups = static_cast<int>(sizes.size());
  // Instance arguments in group order; BatchGroup ranges over them.
  b->argsIndex.assign(count, -1);
  b->hGroups.resize(groups);
  int next = 0;
  for (int g = 0; g < groups; ++g) {
    b->hGroups[g].size = sizes[g];
    b->hGroups[g].first = next;
    for (int i : members[g]) {
      b->argsIndex[i] = next++;
      b->inst[i].groupSize = sizes[g];
    }
    b->hGroups[g].last = next;
  }
  // Placement: workgroup b runs on XCD b % xcds; slot j of XCD x is
  // workgroup x + xcds * j.
  b->groupOf.assign(G, -1);
  b->rankOf.assign(G, 0);
  std::vector<int> order(groups);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sizes[x] > sizes[y]; });
  const int X = b->xcdAware && b->xcds > 1 && G % b->xcds == 0 ? b->xcds : 1;
  const int perX = G / X;
  std::vector<int> used(X, 0);
  auto place = [&](int g, int x, int rank) {
    const int wg = x + X * used[x]++;
    b->groupOf[wg] = g;
    b->rankOf[wg] = rank;
  };
  for (int g : order) {
    const int s = sizes[g];
    const int roomiest = static_cast<int>(std::min_element(used.begin(), used.end()) - used.begin());
    if (perX - used[roomiest] >= s) {
      for (int r = 0; r < s; ++r) place(g, roomiest, r);
    } else {
      int r = 0;
      while (r < s) {  // spread: always the XCD with the most free slots
        const int x = static_cast<int>(std::min_element(used.begin(), used.end()) - used.begin());
        place(g, x, r++);
      }
    }
  }
}

int prepare(mosp_batch *b) {
  drop_graph(b);
  const int count = static_cast<int>(b->inst.size());
  b->hDesc.assign(b->descCount, ObjDesc{});
  build_schedule(b);
  b->hArgs.assign(count, LaunchArgs{});
  // Light instances always run the asynchronous loop: the batched kernel
  // compiles no other for them.
  const bool async = b->hubs == 0;
  const unsigned qcap = b->queueCap >= 256 ? static_cast<unsigned>(b->queueCap)
                                           : static_cast<unsigned>(MOSP_QUEUE_CAP_MAX);
  b->pools.clear();
  for (int i = 0; i < count; ++i) {
    Instance &I = b->inst[i];
    const int n = I.n, K = I.K;
    char *p = b->slab + I.slabOffset;
    std::vector<ObjState> states(K);
    const unsigned segCap = pool_seg_cap(n);
    for (int k = 0; k < K; ++k) {
      states[k] = carve_state(p, n, segCap);
      b->pools.emplace_back(states[k].pool, pool_bytes(segCap));
    }
    if (async && !graph_records4(I.spec.graph, b->stream)) {
      return 0;
    }
    const size_t nn = static_cast<size_t>(n);
    int *cDegree = reinterpret_cast<int *>(p);
    p += align256((nn + 1) * sizeof(int));
    int *cRowPtr = reinterpret_cast<int *>(p);
    p += align256((nn + 1) * sizeof(int));
    int *cColInd = reinterpret_cast<int *>(p);
    p += align256(nn * K * sizeof(int));
    int *cWeights = reinterpret_cast<int *>(p);
    p += align256(nn * K * sizeof(int));
    int *cPartials = reinterpret_cast<int *>(p);
    int anyFallback = 0;
    for (int k = 0; k < K; ++k) {
      const double scaled = std::llround(static_cast<double>(I.delta[k]) * b->deltaScale);
      const u64 delta = static_cast<u64>(std::max(1.0, scaled));
      ObjDesc d = make_desc(I.out[k], &I.in[k], &I.changes[k], I.pack[k], delta, states[k]);
      graph_edges(I.spec.graph, k, &d.outEdge, &d.inEdge);
      if (async) {
        d.outEdge4 = I.spec.graph->outEdge4 + static_cast<size_t>(k) * I.spec.graph->stride;
      }
      const size_t off = static_cast<size_t>(k) * nn;
      d.distIn = I.spec.d_dist_in + off;
      d.parentIn = I.spec.d_parents_in + off;
      d.distOut = I.spec.d_dist + off;
      d.parentOut = I.spec.d_parents + off;
      b->hDesc[I.descOffset + k] = d;
      anyFallback |= I.pack[k].bits == 0 ? 1 : 0;
    }
    mosp_csr view{n, 0, cRowPtr, nullptr, nullptr};
    ObjDesc c = make_desc(view, nullptr, nullptr, I.combPack, 0, states[0]);
    c.outEdge = reinterpret_cast<int2 *>(cColInd);  // K * n records over cColInd + cWeights
    c.distOut = I.spec.d_comb_dist;
    c.parentOut = I.spec.d_comb_parent;
    b->hDesc[I.descOffset + K] = c;

    LaunchArgs a{};
    a.objs = b->dDesc + I.descOffset;
    a.nobj = K;
    a.n = n;
    a.source = I.spec.source;
    a.phases = PH_UPDATE | PH_COMBINED;
    a.maxRounds = jump_rounds(n);
    a.anyFallback = anyFallback;
    a.laneBits = b->laneBits;
    a.listCap = list_cap_override();
    a.initPush = init_push_override();
    a.queueCap = b->queueCap;
    a.hubEdges = hub_edges_override();
    a.hopsUpdate = b->hopsUpdate;
    a.hopsSearch = b->hopsSearch;
    a.genDev = b->dGen + i;
    a.ctl = b->dCtl + i;
    a.abortWord = b->dAbort;
    a.barrierTicks =
        barrier_ticks(b->barrierMs >= 0 ? b->barrierMs : barrier_timeout_ms_default(), b->clockKhz);
    a.cParents = I.spec.d_parents;
    a.cK = K;
    a.cBase = static_cast<int>(I.scale * (K + 1));
    for (int k = 0; k < K; ++k) {
      a.cTerms[k] = static_cast<int>(I.pref.empty() ? I.scale : I.scale / I.pref[k]);
    }
    a.cDelta = 0;  // the device rule of the combined step
    a.cDegree = cDegree;
    a.cRowPtr = cRowPtr;
    a.cColInd = cColInd;
    a.cWeights = cWeights;
    a.cEdge = reinterpret_cast<int2 *>(cColInd);
    a.cPartials = cPartials;
    a.stage = reinterpret_cast<u64 *>(cColInd);  // cColInd, cWeights: 8 * K * n bytes in a row
    a.cDesc = b->dDesc + I.descOffset + K;
    if (async) {
      a.async = b->dAsync + i;
      a.segCap = segCap;
      a.spillHi = std::min(qcap / 4, async_ring_entries(b->shapeThreads) / 2);
      a.spillLo = a.spillHi / 2;
      a.asyncFlags = b->asyncFlags;
      a.keepHungry = static_cast<unsigned>(b->shapeThreads >> std::max(0, b->laneBits));
    }
    b->hArgs[b->argsIndex[i]] = a;
  }
  // Pool slots hold -1 when free.
  for (const auto &pool : b->pools) {
    HIP_TRY(hipMemsetAsync(pool.first, 0xff, pool.second, b->stream));
  }
  HIP_TRY(hipStreamSynchronize(b->stream));
  const size_t groups = b->hGroups.size();
  HIP_TRY(hipMemcpy(b->dDesc, b->hDesc.data(), sizeof(ObjDesc) * b->descCount,
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->dArgs, b->hArgs.data(), sizeof(LaunchArgs) * count,
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->dGroups, b->hGroups.data(), sizeof(BatchGroup) * groups,
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->dGroupOf, b->groupOf.data(), sizeof(int) * b->workgroups,
                    hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(b->dRankOf, b->rankOf.data(), sizeof(int) * b->workgroups,
                    hipMemcpyHostToDevice));
  b->dirty = false;
  return 1;
}

// Captures reset + persistent launch into a graph.  A cooperative launch is
// tried first; if the runtime refuses it under capture, a plain launch of the
// same grid is captured instead: every workgroup is still co-resident because
// the grid never exceeds the occupancy limit (maxBlocks) and the stream runs
// nothing else meanwhile.
int capture(mosp_batch *b) {
  const int count = static_cast<int>(b->inst.size());
  void *params[] = {&b->dArgs, &b->dGroups, &b->dGroupOf, &b->dRankOf};
  for (int mode = 1; mode <= 2; ++mode) {
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t err = launch_ctl_reset_many(b->dCtl, count, b->stream, b->dAbort);
    if (err == hipSuccess) {
      const KernelShape &shape = kernel_shape(b->shapeThreads);
      const void *kernel = shape.batched(b->hubs != 0);
      err = mode == 1 ? hipLaunchCooperativeKernel(kernel, dim3(b->workgroups),
                                                   dim3(shape.threads), params, 0, b->stream)
                      : hipLaunchKernel(kernel, dim3(b->workgroups), dim3(shape.threads), params,
                                        0, b->stream);
    }
    hipGraph_t graph = nullptr;
    const hipError_t endErr = hipStreamEndCapture(b->stream, &graph);
    if (err == hipSuccess && endErr == hipSuccess && graph != nullptr) {
      hipGraphExec_t exec = nullptr;
      if (hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
        b->graph = graph;
        b->exec = exec;
        b->graphMode = mode;
        return 1;
      }
    }
    if (graph != nullptr) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();  // clear the refused launch before the next try
  }
  return fail("Error: could not capture the batched launch in a HIP graph.");
}

}  // namespace

extern "C" mosp_batch *mosp_batch_create(const mosp_batch_instance *instances, int32_t count) {
  clear_error();
  if (instances == nullptr || count <= 0) {
    fail("Error: a batch needs at least one instance.");
    return nullptr;
  }
  std::unique_ptr<mosp_batch> b(new mosp_batch());
  // Sizes first (they need the device's workgroup count), then allocations.
  int descs = 0;
  for (int i = 0; i < count; ++i) descs += std::max(instances[i].k, 0) + 1;
  if (!setup(b.get(), count, descs, 256)) {
    release(b.get());
    return nullptr;
  }
  size_t slab = 0;
  for (int i = 0; i < count; ++i) {
    Instance probe;
    if (!describe(probe, instances[i], i)) {
      release(b.get());
      return nullptr;
    }
    slab += instance_bytes(probe.n, probe.K, b->partialsCap);
  }
  (void)hipFree(b->slab);
  b->slab = nullptr;
  b->capSlab = std::max<size_t>(slab, 256);
  if (hipMalloc(&b->slab, b->capSlab) != hipSuccess || !load(b.get(), instances, count)) {
    if (b->slab == nullptr) fail("Error: could not allocate the batch scratch.");
    release(b.get());
    return nullptr;
  }
  return b.release();
}

extern "C" mosp_batch *mosp_batch_create_capacity(int32_t max_instances, int32_t max_objectives,
                                                  int64_t slab_bytes) {
  clear_error();
  if (max_instances <= 0 || max_objectives <= 0 || slab_bytes < 0) {
    fail("Error: invalid batch capacity.");
    return nullptr;
  }
  std::unique_ptr<mosp_batch> b(new mosp_batch());
  if (!setup(b.get(), max_instances, max_objectives + max_instances,
             static_cast<size_t>(slab_bytes))) {
    release(b.get());
    return nullptr;
  }
  return b.release();
}

extern "C" int mosp_batch_reset(mosp_batch *b, const mosp_batch_instance *instances,
                                int32_t count) {
  clear_error();
  if (b == nullptr || instances == nullptr || count <= 0) {
    return fail("Error: a batch needs at least one instance.");
  }
  drop_graph(b);
  return load(b, instances, count);
}

extern "C" void *mosp_batch_stream(const mosp_batch *b) {
  return b == nullptr ? nullptr : static_cast<void *>(b->stream);
}

extern "C" int64_t mosp_batch_instance_bytes(int32_t n, int32_t k) {
  // Room for the combined scan's partials of either shape's largest grid.
  const int most = std::max(shape_max_blocks(kernel_shape(1024), true),
                            shape_max_blocks(kernel_shape(512), true));
  if (n <= 0 || k <= 0 || most < 1) {
    return -1;
  }
  return static_cast<int64_t>(instance_bytes(n, k, most));
}

extern "C" void mosp_batch_destroy(mosp_batch *b) {
  if (b == nullptr) return;
  (void)hipStreamSynchronize(b->stream);
  release(b);
  delete b;
}

extern "C" int mosp_batch_configure(mosp_batch *b, int32_t workgroups, int32_t lane_bits,
                                    double delta_scale, int32_t xcd_aware) {
  clear_error();
  if (b == nullptr || !(delta_scale > 0)) {
    return fail("Error: invalid batch configuration.");
  }
  const int wg = workgroups > 0 ? workgroups : b->maxBlocks;
  if (wg > b->maxBlocks) {
    return fail("Error: " + std::to_string(wg) + " workgroups exceed the " +
                std::to_string(b->maxBlocks) + " that can be co-resident.");
  }
  b->workgroups = wg;
  if (lane_bits >= 0) b->laneBits = std::min(5, static_cast<int>(lane_bits));
  b->deltaScale = delta_scale;
  b->xcdAware = xcd_aware;
  b->dirty = true;
  return 1;
}

extern "C" int mosp_batch_set_shape(mosp_batch *b, int32_t threads, int32_t queue_cap) {
  clear_error();
  if (b == nullptr || (threads != 512 && threads != 1024) || queue_cap < 0 ||
      (queue_cap > 0 && (queue_cap < 256 || queue_cap > MOSP_QUEUE_CAP_MAX))) {
    return fail("Error: workgroups of 512 or 1024 threads and an LDS queue of 0 (all) or "
                "256.." + std::to_string(MOSP_QUEUE_CAP_MAX) + " entries.");
  }
  if (threads != b->shapeThreads) {
    const int most = shape_max_blocks(kernel_shape(threads), true);
    if (most < 1) {
      return fail("Error: the batched kernel does not fit on a CU.");
    }
    // The same share of the device in the new shape.
    const double share = static_cast<double>(b->workgroups) / std::max(b->maxBlocks, 1);
    b->shapeThreads = threads;
    b->maxBlocks = most;
    b->workgroups = std::max(1, std::min(most, static_cast<int>(share * most + 0.5)));
    drop_graph(b);
  }
  b->queueCap = queue_cap;
  b->dirty = true;
  return 1;
}

extern "C" int mosp_batch_set_cost_model(mosp_batch *b, double latency_items) {
  clear_error();
  if (b == nullptr || !(latency_items >= 0)) {
    return fail("Error: invalid cost model.");
  }
  b->latencyItems = latency_items;
  b->dirty = true;
  return 1;
}

extern "C" int mosp_batch_run(mosp_batch *b, int32_t use_graph, float *kernel_ms) {
  clear_error();
  if (b == nullptr) return fail("Error: no batch.");
  if (b->dirty && !prepare(b)) return 0;
  const int count = static_cast<int>(b->inst.size());
  if (b->scrub) {
    // The last run ended early: its stamps and far marks
    // are not in the state the next run relies on.
    HIP_TRY(hipMemsetAsync(b->slab, 0, b->usedSlab, b->stream));
    HIP_TRY(hipMemsetAsync(b->dGen, 0, sizeof(int) * std::max(count, 1), b->stream));
    for (const auto &pool : b->pools) {
      HIP_TRY(hipMemsetAsync(pool.first, 0xff, pool.second, b->stream));
    }
    b->scrub = false;
  }
  if (use_graph == 1) {
    if (b->exec == nullptr && !capture(b)) return 0;
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    HIP_TRY(hipGraphLaunch(b->exec, b->stream));
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
  } else {
    // use_graph 0: cooperative launch; 2: plain launch of the same
    // occupancy-bounded grid (co-resident as long as the launches running
    // together fit the device, e.g. slots of a server splitting the CUs).
    void *params[] = {&b->dArgs, &b->dGroups, &b->dGroupOf, &b->dRankOf};
    HIP_TRY(launch_ctl_reset_many(b->dCtl, count, b->stream, b->dAbort));
    HIP_TRY(hipEventRecord(b->ev0, b->stream));
    const KernelShape &shape = kernel_shape(b->shapeThreads);
    const void *kernel = shape.batched(b->hubs != 0);
    if (use_graph == 2) {
      HIP_TRY(hipLaunchKernel(kernel, dim3(b->workgroups), dim3(shape.threads), params, 0,
                              b->stream));
    } else {
      HIP_TRY(launch_persistent(kernel, dim3(b->workgroups), dim3(shape.threads), params,
                                b->stream));
    }
    HIP_TRY(hipEventRecord(b->ev1, b->stream));
  }
  unsigned gaveUp = 0;
  HIP_TRY(hipMemcpyAsync(&gaveUp, b->dAbort, sizeof(unsigned), hipMemcpyDeviceToHost, b->stream));
  HIP_TRY(hipStreamSynchronize(b->stream));
  HIP_TRY(hipEventElapsedTime(&b->lastMs, b->ev0, b->ev1));
  if (kernel_ms != nullptr) *kernel_ms = b->lastMs;
  if (gaveUp != 0) {
    b->scrub = true;
    return fail("Error: a grid barrier of the batched launch waited longer than its limit and "
                "the launch gave up (its workgroups were not all resident); no result was "
                "written.");
  }
  return 1;
}

extern "C" int mosp_batch_set_barrier_timeout(mosp_batch *b, double ms) {
  clear_error();
  if (b == nullptr || !(ms >= 0)) return fail("Error: the barrier timeout must be >= 0 ms.");
  b->barrierMs = ms;
  b->dirty = true;  // the limit is part of every instance's launch arguments
  return 1;
}

extern "C" int mosp_batch_check(mosp_batch *b) {
  clear_error();
  if (b == nullptr) return fail("Error: no batch.");
  for (size_t i = 0; i < b->inst.size(); ++i) {
    HIP_TRY(hipMemcpy(b->hCtl, b->dCtl + i, CTL_READBACK, hipMemcpyDeviceToHost));
    for (int k = 0; k < b->inst[i].K; ++k) {
      if (!input_tree_ok(b->hCtl->obj[k].overflow, "instance " + std::to_string(i) + ": ")) {
        return 0;
      }
    }
  }
  return 1;
}

extern "C" int mosp_batch_stats(mosp_batch *b, int32_t i, mosp_sosp_stats *stats,
                                mosp_combine_stats *cstats) {
  clear_error();
  if (b == nullptr || i < 0 || i >= static_cast<int32_t>(b->inst.size())) {
    return fail("Error: no such batch instance.");
  }
  const Instance &I = b->inst[i];
  HIP_TRY(hipMemcpy(b->hCtl, b->dCtl + i, CTL_READBACK, hipMemcpyDeviceToHost));
  const Ctl &c = *b->hCtl;
  for (int k = 0; k < I.K && stats != nullptr; ++k) {
    fill_stats(c, k, b->hDesc[I.descOffset + k], b->lastMs, &stats[k]);
  }
  if (cstats != nullptr) {
    *cstats = mosp_combine_stats{};
    cstats->scale = I.scale;
    cstats->n_edges = static_cast<int32_t>(c.combEdges);
    cstats->delta = static_cast<int64_t>(c.combDelta);
    fill_stats(c, MOSP_MAX_OBJ, b->hDesc[I.descOffset + I.K], b->lastMs, &cstats->search);
    cstats->search.jump_rounds = 0;
    cstats->search.invalidated = 0;
    if (c.clkSospEnd > c.clkStart && c.clkEnd >= c.clkSospEnd) {
      cstats->sosp_phase_ms = static_cast<float>((c.clkSospEnd - c.clkStart) / b->clockKhz);
      cstats->combined_phase_ms = static_cast<float>((c.clkEnd - c.clkSospEnd) / b->clockKhz);
    }
  }
  return 1;
}

extern "C" int mosp_batch_schedule(const mosp_batch *b, int32_t *group_size,
                                   int32_t *graph_mode) {
  if (b == nullptr) return fail("Error: no batch.");
  for (size_t i = 0; i < b->inst.size() && group_size != nullptr; ++i) {
    group_size[i] = b->inst[i].groupSize;
  }
  if (graph_mode != nullptr) *graph_mode = b->graphMode;
  return 1;
}
