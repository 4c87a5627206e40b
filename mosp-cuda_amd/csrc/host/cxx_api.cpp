// cxx_api.cpp -- the reference-named C++ API (sospUpdateGpu.h,
// combinedGraphGpu.h, parallelCombinedGraph.h, deviceGraph.h, mospUpdate.h)
// over the C ABI, and the host pipeline entry mosp_update_host.
//
// mospUpdate computes what the reference's src/mospUpdate.cu:57-225 computes
// (same checks, per-objective bounds, changed lists, outputs and timed
// regions) as a device pipeline of this library: the original graph goes
// to HBM once, the batch is applied there (mosp_device_graph_apply, equal to
// the host applyChangeBatch), the weight statistics are reduced there, the K
// updates and Steps 2-3 run as one persistent launch, and the updated graph
// comes back with the trees.  Helpers follow src/combinedGraphGpu.cu:129-208
// (Pref scale, edge weight, path costs) and src/mospUpdate.cu:40-55
// (canonical tie rule).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <iostream>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "combinedGraphGpu.h"
#include "csrGraph.h"
#include "deviceArray.h"
#include "deviceGraph.h"
#include "host_api.hpp"
#include "mospUpdate.h"
#include "mosp_hip.h"
#include "parallelCombinedGraph.h"
#include "sospUpdateGpu.h"
#include "stageTimer.h"

namespace {

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
      .count();
}

void report_error() {
  const char *e = mosp_last_error();
  if (e != nullptr && *e != '\0') {
    std::cerr << e << "\n";
  }
}

mosp_csr to_c(const DeviceCsr &d) {
  return mosp_csr{d.numberOfNodes, d.numberOfEdges, d.rowPtr, d.colInd, d.weights};
}

void from_c(const mosp_sosp_stats &s, SospStats &out) {
  out.invalidated = s.invalidated;
  out.jumpRounds = s.jump_rounds;
  out.iterations = s.iterations;
  out.epochs = s.epochs;
  out.pushes = s.pushes;
  out.packedParents = s.packed_parents != 0;
  out.kernelMs = s.kernel_ms;
}

}  // namespace

// ------------------------------------------------------------------ helpers

namespace mosp {

int64_t preference_scale(const int32_t *pref, int32_t nPref, int32_t K) {
  if (nPref == 0) {
    return 1;
  }
  if (nPref != K) {
    return 0;
  }
  int64_t L = 1;
  for (int i = 0; i < nPref; ++i) {
    if (pref[i] < 1) {
      return 0;
    }
    L = L / std::gcd(L, static_cast<int64_t>(pref[i])) * pref[i];
    if (L > (int64_t{1} << 20)) {
      return 0;
    }
  }
  return L;
}

int64_t combined_edge_weight(uint32_t mask, const int32_t *pref, int32_t nPref,
                             int32_t K, int64_t scale) {
  int64_t w = scale * (K + 1);
  for (int i = 0; i < K; ++i) {
    if (mask & (1u << i)) {
      w -= nPref == 0 ? scale : scale / pref[i];
    }
  }
  return w;
}

}  // namespace mosp

long long preferenceScale(const std::vector<int> &pref, int K) {
  return mosp::preference_scale(pref.data(), static_cast<int32_t>(pref.size()), K);
}

long long combinedEdgeWeight(unsigned int mask, const std::vector<int> &pref, int K,
                             long long scale) {
  return mosp::combined_edge_weight(mask, pref.data(), static_cast<int32_t>(pref.size()),
                                    K, scale);
}

// Costs by memoised ancestor walks: a vertex's chain is followed up to the
// source or a vertex already resolved, then resolved downwards.  A chain
// that ends at another root, or runs into itself, is not reachable (INF),
// and its edges are not needed -- the vertices whose tree edge is looked up
// are exactly those reachable from the source, as in a walk of the tree
// from the source down.
bool mospPathCosts(const CsrGraph &graph, const std::vector<int> &parent, int source,
                   std::vector<long long> &costs) {
  const int n = graph.numberOfNodes, K = graph.numberOfObjectives;
  const size_t kk = static_cast<size_t>(K);
  costs.assign(static_cast<size_t>(n) * kk, DISTANCE_INF);
  enum : char { UNSEEN = 0, ON_CHAIN = 1, REACHED = 2, CUT = 3 };
  std::vector<char> state(static_cast<size_t>(n), UNSEEN);
  state[source] = REACHED;
  std::fill(costs.begin() + static_cast<std::ptrdiff_t>(source * kk),
            costs.begin() + static_cast<std::ptrdiff_t>((source + 1) * kk), 0LL);
  std::vector<int> chain;
  for (int start = 0; start < n; ++start) {
    int x = start;
    chain.clear();
    while (state[x] == UNSEEN) {
      state[x] = ON_CHAIN;
      chain.push_back(x);
      if (parent[x] < 0) break;
      x = parent[x];
    }
    // The walk stopped at a resolved vertex (REACHED or CUT), at a root other
    // than the source (on the chain), or on the chain itself (a cycle).
    const bool reached = state[x] == REACHED;
    for (auto it = chain.rbegin(); it != chain.rend(); ++it) {
      const int v = *it;
      if (!reached) {
        state[v] = CUT;
        continue;
      }
      const int p = parent[v];
      int edge = graph.rowPtr[p];
      while (edge < graph.rowPtr[p + 1] && graph.colInd[edge] != v) ++edge;
      if (edge == graph.rowPtr[p + 1]) return false;  // the tree edge p -> v is missing
      for (size_t k = 0; k < kk; ++k) {
        costs[v * kk + k] = costs[p * kk + k] + graph.weight(edge, static_cast<int>(k));
      }
      state[v] = REACHED;
    }
  }
  return true;
}

// ---------------------------------------------------------------- workspace

SospWorkspace::~SospWorkspace() { release(); }

bool SospWorkspace::reserve(int requested, int slots) {
  if (handle == nullptr) {
    handle = mosp_ws_create();
  }
  if (!mosp_ws_reserve(handle, requested, slots)) {
    report_error();
    return false;
  }
  capacity = std::max(capacity, requested);
  gridBlocks = mosp_ws_grid(handle);
  return true;
}

void SospWorkspace::release() {
  mosp_ws_destroy(handle);
  handle = nullptr;
  capacity = 0;
  gridBlocks = 0;
}

long long defaultDelta(long long m, int n, long long weightSum) {
  return mosp_default_delta(m, n, weightSum);
}

bool sospUpdateGpu(const DeviceCsr &out, const DeviceCsr &in, const DeviceChanges &ch,
                   int source, long long delta, long long maxWeight, SospWorkspace &ws,
                   long long *d_distances, int *d_parent, SospStats *stats) {
  SospStats local;
  SospStats &s = stats != nullptr ? *stats : local;
  s = SospStats();
  if (out.numberOfNodes == 0) return true;
  if (!ws.reserve(out.numberOfNodes) || delta <= 0) return false;
  const mosp_csr o = to_c(out), i = to_c(in);
  const mosp_changes c{ch.changedFrom, ch.changedTo, ch.numberOfChanged, ch.insertHeads,
                       ch.numberOfInsertHeads};
  mosp_sosp_stats cs{};
  if (!mosp_sosp_update(&o, &i, &c, source, delta, maxWeight, ws.handle,
                        reinterpret_cast<int64_t *>(d_distances), d_parent, &cs, nullptr)) {
    report_error();
    return false;
  }
  from_c(cs, s);
  return true;
}

bool sospFromScratchGpu(const DeviceCsr &out, int source, long long delta,
                        long long maxWeight, SospWorkspace &ws, long long *d_distances,
                        int *d_parent, SospStats *stats) {
  SospStats local;
  SospStats &s = stats != nullptr ? *stats : local;
  s = SospStats();
  if (out.numberOfNodes == 0) return true;
  if (!ws.reserve(out.numberOfNodes) || delta <= 0 || source < 0 ||
      source >= out.numberOfNodes) {
    return false;
  }
  const mosp_csr o = to_c(out);
  mosp_sosp_stats cs{};
  if (!mosp_sosp_from_scratch(&o, source, delta, maxWeight, ws.handle,
                              reinterpret_cast<int64_t *>(d_distances), d_parent, &cs,
                              nullptr)) {
    report_error();
    return false;
  }
  from_c(cs, s);
  return true;
}

bool combinedGraphSospGpu(const int *d_parents, int n, int K, const std::vector<int> &pref,
                          int source, long long delta, CombineWorkspace &cws,
                          SospWorkspace &ws, long long *d_distances, int *d_parent,
                          CombineStats *stats) {
  CombineStats local;
  CombineStats &s = stats != nullptr ? *stats : local;
  s = CombineStats();
  if (K <= 0 || K > 32 || n <= 0 || source < 0 || source >= n) {
    std::cerr << "Error: invalid combined-graph parameters.\n";
    return false;
  }
  if (preferenceScale(pref, K) == 0) {
    std::cerr << "Error: invalid preference vector (need K values >= 1).\n";
    return false;
  }
  if (!cws.reserve(n, K) || !ws.reserve(n)) return false;
  mosp_combine_stats cs{};
  if (!mosp_combined_sosp(d_parents, n, K, pref.data(), static_cast<int32_t>(pref.size()),
                          source, delta, ws.handle, reinterpret_cast<int64_t *>(d_distances),
                          d_parent, &cs, nullptr)) {
    report_error();
    return false;
  }
  s.numberOfEdges = cs.n_edges;
  s.scale = cs.scale;
  s.delta = cs.delta;
  from_c(cs.search, s.search);
  return true;
}

// ------------------------------------------------------------- device graph

DeviceGraph::~DeviceGraph() { mosp_device_graph_free(handle); }

DeviceCsr DeviceGraph::out(int k) const {
  mosp_csr o{};
  mosp_device_graph_view(handle, k, &o, nullptr);
  DeviceCsr d;
  d.numberOfNodes = o.n;
  d.numberOfEdges = o.m;
  d.rowPtr = o.row_ptr;
  d.colInd = o.col_ind;
  d.weights = o.weights;
  return d;
}

DeviceCsr DeviceGraph::in(int k) const {
  mosp_csr i{};
  mosp_device_graph_view(handle, k, nullptr, &i);
  DeviceCsr d;
  d.numberOfNodes = i.n;
  d.numberOfEdges = i.m;
  d.rowPtr = i.row_ptr;
  d.colInd = i.col_ind;
  d.weights = i.weights;
  return d;
}

bool uploadDeviceGraph(const CsrGraph &graph, DeviceGraph &device) {
  mosp_device_graph_free(device.handle);
  device.handle = mosp_device_graph_upload(graph.numberOfNodes, graph.numberOfObjectives,
                                           graph.rowPtr.data(), graph.colInd.data(),
                                           graph.weights.data());
  if (device.handle == nullptr) {
    report_error();
    return false;
  }
  device.numberOfNodes = graph.numberOfNodes;
  device.numberOfEdges = graph.numberOfEdges();
  device.numberOfObjectives = graph.numberOfObjectives;
  return true;
}

// ------------------------------------------------------------ MOSP pipeline

double MospTimings::gpuCompute() const {
  return std::accumulate(objectives.begin(), objectives.end(), combined);
}

// The lowest-id tie rule as a per-vertex minimum: parent[v] becomes the
// smallest u over every tight edge u -> v (d[u] + w = d[v]) and its current
// parent, for every vertex with a parent; vertices without one (the source,
// unreachable vertices) keep -1.  Rows are split over threads; the minimum
// is taken with atomic compare-exchange, so the order does not matter --
// which is why one sequential pass over the out-edges
// (src/mospUpdate.cu:40-55) gives the same tree.
void canonicalizeTree(const CsrGraph &graph, int objective, int source,
                      const std::vector<long long> &distances, int *parent) {
  const int n = graph.numberOfNodes;
  std::vector<std::atomic<int>> low(static_cast<size_t>(n));
  for (int v = 0; v < n; ++v) low[v].store(parent[v], std::memory_order_relaxed);
  const unsigned workers = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int parts = n >= (1 << 16) ? static_cast<int>(workers) : 1;
  auto rows = [&](int part) {
    const int lo = static_cast<int>(static_cast<long long>(n) * part / parts);
    const int hi = static_cast<int>(static_cast<long long>(n) * (part + 1) / parts);
    for (int u = lo; u < hi; ++u) {
      const long long du = distances[u];
      if (du >= DISTANCE_INF / 2) continue;
      for (int e = graph.rowPtr[u]; e < graph.rowPtr[u + 1]; ++e) {
        const int v = graph.colInd[e];
        if (v == source || du + graph.weight(e, objective) != distances[v]) continue;
        int cur = low[v].load(std::memory_order_relaxed);
        while (cur >= 0 && u < cur &&
               !low[v].compare_exchange_weak(cur, u, std::memory_order_relaxed)) {
        }
      }
    }
  };
  std::vector<std::thread> threads;
  for (int t = 1; t < parts; ++t) threads.emplace_back(rows, t);
  rows(0);
  for (std::thread &t : threads) t.join();
  for (int v = 0; v < n; ++v) parent[v] = low[v].load(std::memory_order_relaxed);
}

namespace {

// Device state of one mospUpdate call, released on every return path.
struct UpdatePipeline {
  mosp_device_graph *original = nullptr;
  mosp_device_graph *updated = nullptr;
  SospWorkspace ws;
  DeviceArray<long long> dist, combDist;
  DeviceArray<int> parents, combParent;
  DeviceArray<int> lists;  // per objective: changed tails, changed heads; then the insert heads
  ~UpdatePipeline() {
    mosp_device_graph_free(updated);
    mosp_device_graph_free(original);
  }
};

// The check of src/mospUpdate.cu:77-92: every initial tree is rooted at the
// source (distance 0, no parent); the message says what was found.
bool trees_rooted_at(int source, int n, int K, const std::vector<long long> &dist,
                     const std::vector<int> &par) {
  for (int k = 0; k < K; ++k) {
    const size_t s = static_cast<size_t>(k) * n + source;
    if (dist[s] == 0 && par[s] == -1) continue;
    const std::string d = dist[s] >= DISTANCE_INF / 2 ? "INF" : std::to_string(dist[s]);
    std::cerr << "Error: the initial tree of objective " << k
              << " is not rooted at the source " << source << " (its distance is " << d
              << ", its parent " << par[s] << "); compute the trees for this source.\n";
    return false;
  }
  return true;
}

// Per objective k < K: the changed edges (every deletion, then the
// insertions whose weight-increase bit k is set; src/mospUpdate.cu:124-136),
// as tails then heads, followed by the insert heads shared by all.  Returns
// the host image and each objective's [offset, count).
std::vector<int> changed_lists(const ChangeBatch &b, int K, std::vector<size_t> &at,
                               std::vector<int> &count) {
  std::vector<int> img;
  at.assign(K, 0);
  count.assign(K, 0);
  for (int k = 0; k < K; ++k) {
    std::vector<int> tails(b.deleteFrom), heads(b.deleteTo);
    for (int i = 0; i < b.numberOfInserts(); ++i) {
      if ((b.weightIncreaseMask[i] >> k) & 1u) {
        tails.push_back(b.insertFrom[i]);
        heads.push_back(b.insertTo[i]);
      }
    }
    at[k] = img.size();
    count[k] = static_cast<int>(tails.size());
    img.insert(img.end(), tails.begin(), tails.end());
    img.insert(img.end(), heads.begin(), heads.end());
  }
  img.insert(img.end(), b.insertTo.begin(), b.insertTo.end());
  return img;
}

void record_counters(const MospResult &r) {
  if (!instrumentationEnabled()) return;
  for (int k = 0; k < r.numberOfObjectives; ++k) {
    const SospStats &s = r.objectiveStats[k];
    const std::string o = "obj" + std::to_string(k) + "/";
    recordCounter(o + "invalidated", s.invalidated);
    recordCounter(o + "jump_rounds", s.jumpRounds);
    recordCounter(o + "iterations", s.iterations);
    recordCounter(o + "epochs", s.epochs);
    recordCounter(o + "pushes", static_cast<double>(s.pushes));
    recordCounter(o + "packed_parents", s.packedParents ? 1 : 0);
  }
  const CombineStats &c = r.combineStats;
  recordCounter("comb/edges", c.numberOfEdges);
  recordCounter("comb/L", static_cast<double>(c.scale));
  recordCounter("comb/delta", static_cast<double>(c.delta));
  recordCounter("comb/iterations", c.search.iterations);
  recordCounter("comb/pushes", static_cast<double>(c.search.pushes));
}

}  // namespace

bool mospUpdate(const CsrGraph &original, ChangeBatch &batch,
                const std::vector<long long> &initialDistances,
                const std::vector<int> &initialParents, const MospOptions &options,
                CsrGraph &updated, MospResult &result) {
  using clock = std::chrono::steady_clock;
  const int n = original.numberOfNodes;
  const int KG = original.numberOfObjectives;
  const int K = options.numberOfObjectives > 0 ? std::min(options.numberOfObjectives, KG) : KG;
  const int s = options.source;
  const size_t kn = static_cast<size_t>(K) * n;
  result = MospResult();
  result.numberOfObjectives = K;
  MospTimings &tm = result.timings;
  const std::vector<int> &pref = options.preferences;

  if (n <= 0 || K <= 0 || s < 0 || s >= n || initialDistances.size() < kn ||
      initialParents.size() < kn) {
    std::cerr << "Error: invalid MOSP update input (graph, objectives, source or tree sizes).\n";
    return false;
  }
  if (!trees_rooted_at(s, n, K, initialDistances, initialParents)) return false;
  if (preferenceScale(pref, K) == 0) {
    std::cerr << "Error: invalid preference vector (need K values >= 1).\n";
    return false;
  }
  if (batch.numberOfObjectives != KG) {
    std::cout << "Error: change batch has " << batch.numberOfObjectives
              << " objectives, graph has " << KG << ".\n";
    return false;
  }
  if (KG > 32) {
    std::cout << "Error: at most 32 objectives are supported.\n";
    return false;
  }
  auto gpu_failed = [](const char *what) {
    report_error();
    std::cerr << "Error: " << what << " failed on the GPU.\n";
    return false;
  };

  UpdatePipeline dev;
  // Upload: the original graph (reverse CSR and weight columns built in HBM),
  // the initial trees, the engine's workspace.
  auto t0 = clock::now();
  {
    ScopedStage stage("upload", true);
    dev.original = mosp_device_graph_upload(n, KG, original.rowPtr.data(),
                                            original.colInd.data(), original.weights.data());
    if (dev.original == nullptr) return gpu_failed("uploading the graph");
    dev.updated = mosp_device_graph_create();
    const std::vector<long long> d0(initialDistances.begin(), initialDistances.begin() + kn);
    const std::vector<int> p0(initialParents.begin(), initialParents.begin() + kn);
    if (!dev.dist.upload(d0) || !dev.parents.upload(p0) || !dev.combDist.allocate(n) ||
        !dev.combParent.allocate(n) || !dev.ws.reserve(n, options.fused ? K : 1) ||
        !mosp_ws_set_combined_delta_scale(dev.ws.handle, std::max(1, options.combinedDeltaScale))) {
      return gpu_failed("uploading the trees");
    }
  }
  tm.upload = std::chrono::duration<double, std::milli>(clock::now() - t0).count();

  // The batch, applied to the graph in HBM; the weight-increase mask comes
  // back to the batch.  Bounds from the original graph and the inserted
  // weights (src/mospUpdate.cu:104-122), changed lists from the mask.
  t0 = clock::now();
  std::vector<long long> maxWeight(KG, 1), weightSum(KG, 0);
  std::vector<size_t> listAt;
  std::vector<int> listCount;
  {
    ScopedStage stage("apply_batch", true);
    batch.weightIncreaseMask.assign(static_cast<size_t>(batch.numberOfInserts()), 0u);
    if (!mosp_device_graph_apply(dev.original, batch.numberOfInserts(), batch.insertFrom.data(),
                                 batch.insertTo.data(), batch.insertWeights.data(),
                                 batch.numberOfDeletes(), batch.deleteFrom.data(),
                                 batch.deleteTo.data(), dev.updated,
                                 batch.weightIncreaseMask.data(), nullptr)) {
      return gpu_failed("applying the change batch");
    }
    std::vector<int64_t> mx(KG, 0), sum(KG, 0);
    if (!mosp_device_graph_weight_stats(dev.original, mx.data(), sum.data(), nullptr)) {
      return gpu_failed("the weight statistics");
    }
    for (int k = 0; k < KG; ++k) {
      maxWeight[k] = std::max<long long>(1, mx[k]);
      weightSum[k] = sum[k];
    }
    for (size_t i = 0; i < batch.insertWeights.size(); ++i) {
      const int k = static_cast<int>(i % KG);
      maxWeight[k] = std::max<long long>(maxWeight[k], batch.insertWeights[i]);
    }
    if (!dev.lists.upload(changed_lists(batch, K, listAt, listCount))) {
      return gpu_failed("uploading the changed edges");
    }
  }
  tm.applyBatch = std::chrono::duration<double, std::milli>(clock::now() - t0).count();

  // Width per objective: the reference's rule on the original graph
  // (src/mospUpdate.cu:177-181), or the caller's.
  const long long m0 = original.numberOfEdges();
  const long long m = m0 > 0 ? m0 : std::max<int64_t>(1, mosp_device_graph_edges(dev.updated));
  std::vector<long long> delta(K);
  std::vector<mosp_changes> changes(K);
  const int *heads = dev.lists.data() + (listAt.empty() ? 0 : listAt.back() + 2 * listCount.back());
  for (int k = 0; k < K; ++k) {
    const double rule = static_cast<double>(defaultDelta(m, n, weightSum[k]));
    delta[k] = options.delta > 0 ? options.delta
                                 : std::max(1LL, static_cast<long long>(
                                                     rule * std::max(options.deltaScale, 0.0)));
    const int *tails = dev.lists.data() + listAt[k];
    changes[k] = mosp_changes{tails, tails + listCount[k], listCount[k], heads,
                              batch.numberOfInserts()};
  }

  result.objectiveStats.resize(K);
  CombineStats &cst = result.combineStats;
  if (options.fused) {
    // One launch: K updates, combined graph, its search.  The host wall time
    // is split between the K updates and Steps 2-3 by the kernel's own clock.
    std::vector<mosp_sosp_stats> st(K);
    mosp_combine_stats cs{};
    t0 = clock::now();
    const bool ok = mosp_update_fused(
        dev.updated, K, changes.data(), s, reinterpret_cast<const int64_t *>(delta.data()),
        reinterpret_cast<const int64_t *>(maxWeight.data()), pref.data(),
        static_cast<int32_t>(pref.size()), dev.ws.handle, nullptr, nullptr,
        reinterpret_cast<int64_t *>(dev.dist.data()), dev.parents.data(),
        reinterpret_cast<int64_t *>(dev.combDist.data()), dev.combParent.data(), st.data(), &cs,
        nullptr);
    const double wall = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
    if (!ok) return gpu_failed("the MOSP update");
    for (int k = 0; k < K; ++k) from_c(st[k], result.objectiveStats[k]);
    cst.numberOfEdges = cs.n_edges;
    cst.scale = cs.scale;
    cst.delta = cs.delta;
    from_c(cs.search, cst.search);
    const double device = static_cast<double>(cs.sosp_phase_ms) + cs.combined_phase_ms;
    const double share = device > 0 ? cs.sosp_phase_ms / device : 1.0;
    tm.objectives.assign(K, wall * share / K);
    tm.combined = wall * (1.0 - share);
    for (int k = 0; k < K; ++k) recordStage("obj" + std::to_string(k) + "/sosp_update_gpu", tm.objectives[k]);
    recordStage("combined_graph_gpu", tm.combined);
  } else {
    // The reference's schedule: K launches, then the combined step.
    for (int k = 0; k < K; ++k) {
      mosp_csr out{}, in{};
      mosp_sosp_stats st{};
      mosp_device_graph_view(dev.updated, k, &out, &in);
      ScopedStage stage("obj" + std::to_string(k) + "/sosp_update_gpu", true);
      t0 = clock::now();
      if (!mosp_sosp_update(&out, &in, &changes[k], s, delta[k], maxWeight[k], dev.ws.handle,
                            reinterpret_cast<int64_t *>(dev.dist.data()) + static_cast<size_t>(k) * n,
                            dev.parents.data() + static_cast<size_t>(k) * n, &st, nullptr)) {
        return gpu_failed(("the update of objective " + std::to_string(k)).c_str());
      }
      tm.objectives.push_back(std::chrono::duration<double, std::milli>(clock::now() - t0).count());
      from_c(st, result.objectiveStats[k]);
    }
    ScopedStage stage("combined_graph_gpu", true);
    t0 = clock::now();
    CombineWorkspace cws;
    if (!combinedGraphSospGpu(dev.parents.data(), n, K, pref, s, 0, cws, dev.ws,
                              dev.combDist.data(), dev.combParent.data(), &cst)) {
      return gpu_failed("the combined graph step");
    }
    tm.combined = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  }

  // Download: trees, combined search, and the updated graph.
  t0 = clock::now();
  {
    ScopedStage stage("download", true);
    const int64_t mu = mosp_device_graph_edges(dev.updated);
    updated = CsrGraph();
    updated.numberOfNodes = n;
    updated.numberOfObjectives = KG;
    updated.rowPtr.resize(static_cast<size_t>(n) + 1);
    updated.colInd.resize(static_cast<size_t>(mu));
    updated.weights.resize(static_cast<size_t>(mu) * KG);
    if (!dev.dist.download(result.distances) || !dev.parents.download(result.parents) ||
        !dev.combDist.download(result.combinedDistances) ||
        !dev.combParent.download(result.combinedParent) ||
        !mosp_device_graph_download(dev.updated, updated.rowPtr.data(), updated.colInd.data(),
                                    updated.weights.data())) {
      return gpu_failed("copying the results back");
    }
  }
  tm.download = std::chrono::duration<double, std::milli>(clock::now() - t0).count();
  record_counters(result);
  return true;
}

// ------------------------------------------------------------------ C ABI

extern "C" int64_t mosp_apply_batch(int32_t n, int32_t K, const int32_t *rowPtr,
                                    const int32_t *colInd, const int32_t *weights,
                                    int32_t nIns, const int32_t *insFrom, const int32_t *insTo,
                                    const int32_t *insW, int32_t nDel, const int32_t *delFrom,
                                    const int32_t *delTo, int32_t *outRow, int32_t *outCol,
                                    int32_t *outW, uint32_t *incMask) {
  CsrGraph g, u;
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  const int m = rowPtr[n];
  g.rowPtr.assign(rowPtr, rowPtr + n + 1);
  g.colInd.assign(colInd, colInd + m);
  g.weights.assign(weights, weights + static_cast<size_t>(m) * K);
  ChangeBatch b;
  b.numberOfObjectives = K;
  b.insertFrom.assign(insFrom, insFrom + nIns);
  b.insertTo.assign(insTo, insTo + nIns);
  b.insertWeights.assign(insW, insW + static_cast<size_t>(nIns) * K);
  b.deleteFrom.assign(delFrom, delFrom + nDel);
  b.deleteTo.assign(delTo, delTo + nDel);
  if (!applyChangeBatch(g, b, u)) return -1;
  const size_t mu = static_cast<size_t>(u.numberOfEdges());
  std::copy(u.rowPtr.begin(), u.rowPtr.end(), outRow);
  std::copy(u.colInd.begin(), u.colInd.end(), outCol);
  std::copy(u.weights.begin(), u.weights.end(), outW);
  std::copy(b.weightIncreaseMask.begin(), b.weightIncreaseMask.end(), incMask);
  return static_cast<int64_t>(mu);
}

namespace {

CsrGraph graph_of(int32_t n, int32_t K, const int32_t *rowPtr, const int32_t *colInd,
                  const int32_t *weights) {
  CsrGraph g;
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  const size_t m = static_cast<size_t>(rowPtr[n]);
  g.rowPtr.assign(rowPtr, rowPtr + n + 1);
  g.colInd.assign(colInd, colInd + m);
  g.weights.assign(weights, weights + m * K);
  return g;
}

}  // namespace

extern "C" void mosp_canonicalize_tree(int32_t n, int32_t K, const int32_t *rowPtr,
                                       const int32_t *colInd, const int32_t *weights,
                                       int32_t objective, int32_t source, const int64_t *dist,
                                       int32_t *parent) {
  const CsrGraph g = graph_of(n, K, rowPtr, colInd, weights);
  const std::vector<long long> d(dist, dist + n);
  canonicalizeTree(g, objective, source, d, parent);
}

extern "C" int mosp_path_costs_host(int32_t n, int32_t K, const int32_t *rowPtr,
                                    const int32_t *colInd, const int32_t *weights,
                                    const int32_t *parent, int32_t source, int64_t *costs) {
  const CsrGraph g = graph_of(n, K, rowPtr, colInd, weights);
  std::vector<long long> c;
  if (!mospPathCosts(g, std::vector<int>(parent, parent + n), source, c)) return 0;
  std::copy(c.begin(), c.end(), costs);
  return 1;
}

extern "C" int mosp_update_host(
    int32_t n, int32_t kg, const int32_t *rowPtr, const int32_t *colInd,
    const int32_t *weights, int32_t nIns, const int32_t *insFrom, const int32_t *insTo,
    const int32_t *insW, int32_t nDel, const int32_t *delFrom, const int32_t *delTo,
    int32_t kUse, int32_t source, const int32_t *pref, int32_t nPref, int64_t delta,
    int32_t fused, const int64_t *initDist, const int32_t *initParents, int64_t *outM,
    int32_t *updRow, int32_t *updCol, int32_t *updW, int64_t *dist, int32_t *parents,
    int64_t *combDist, int32_t *combParent, mosp_sosp_stats *stats,
    mosp_combine_stats *cstats, double *timings) {
  CsrGraph g, u;
  g.numberOfNodes = n;
  g.numberOfObjectives = kg;
  const int m = rowPtr[n];
  g.rowPtr.assign(rowPtr, rowPtr + n + 1);
  g.colInd.assign(colInd, colInd + m);
  g.weights.assign(weights, weights + static_cast<size_t>(m) * kg);
  ChangeBatch b;
  b.numberOfObjectives = kg;
  b.insertFrom.assign(insFrom, insFrom + nIns);
  b.insertTo.assign(insTo, insTo + nIns);
  b.insertWeights.assign(insW, insW + static_cast<size_t>(nIns) * kg);
  b.deleteFrom.assign(delFrom, delFrom + nDel);
  b.deleteTo.assign(delTo, delTo + nDel);
  const int K = kUse > 0 ? std::min(kUse, kg) : kg;
  std::vector<long long> d0(initDist, initDist + static_cast<size_t>(K) * n);
  std::vector<int> p0(initParents, initParents + static_cast<size_t>(K) * n);
  MospOptions opt;
  opt.source = source;
  opt.numberOfObjectives = K;
  opt.preferences.assign(pref, pref + nPref);
  opt.delta = delta;
  opt.fused = fused != 0;
  MospResult r;
  if (!mospUpdate(g, b, d0, p0, opt, u, r)) return 0;
  *outM = u.numberOfEdges();
  std::copy(u.rowPtr.begin(), u.rowPtr.end(), updRow);
  std::copy(u.colInd.begin(), u.colInd.end(), updCol);
  std::copy(u.weights.begin(), u.weights.end(), updW);
  std::copy(r.distances.begin(), r.distances.end(), dist);
  std::copy(r.parents.begin(), r.parents.end(), parents);
  std::copy(r.combinedDistances.begin(), r.combinedDistances.end(), combDist);
  std::copy(r.combinedParent.begin(), r.combinedParent.end(), combParent);
  for (int k = 0; k < K && stats != nullptr; ++k) {
    const SospStats &s = r.objectiveStats[k];
    stats[k] = mosp_sosp_stats{s.invalidated, s.jumpRounds, s.iterations, s.epochs,
                               s.pushes, s.packedParents ? 1 : 0, s.kernelMs};
  }
  if (cstats != nullptr) {
    const CombineStats &c = r.combineStats;
    cstats->n_edges = c.numberOfEdges;
    cstats->scale = c.scale;
    cstats->delta = c.delta;
    cstats->search = mosp_sosp_stats{c.search.invalidated, c.search.jumpRounds,
                                     c.search.iterations, c.search.epochs, c.search.pushes,
                                     c.search.packedParents ? 1 : 0, c.search.kernelMs};
  }
  if (timings != nullptr) {
    const MospTimings &t = r.timings;
    double sosp = 0;
    for (double x : t.objectives) sosp += x;
    timings[0] = t.applyBatch;
    timings[1] = t.upload;
    timings[2] = sosp;
    timings[3] = t.combined;
    timings[4] = t.download;
  }
  return 1;
}
