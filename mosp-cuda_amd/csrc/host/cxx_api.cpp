// cxx_api.cpp -- the reference-named C++ API (sospUpdateGpu.h,
// combinedGraphGpu.h, parallelCombinedGraph.h, deviceGraph.h, mospUpdate.h)
// over the C ABI, and the host pipeline entry mosp_update_host.
//
// Semantics follow the reference's src/mospUpdate.cu:40-225 (root checks,
// per-objective maxWeight / weightSum / changed lists, timing regions) and
// src/combinedGraphGpu.cu:129-208 (Pref scale, edge weight, path costs).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <iostream>
#include <numeric>
#include <string>
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

bool mospPathCosts(const CsrGraph &graph, const std::vector<int> &parent, int source,
                   std::vector<long long> &costs) {
  const int n = graph.numberOfNodes, K = graph.numberOfObjectives;
  costs.assign(static_cast<size_t>(n) * K, DISTANCE_INF);
  // Children lists of the tree, then a breadth-first walk from the source.
  std::vector<int> first(n + 1, 0), kids(n);
  for (int v = 0; v < n; ++v) {
    if (v != source && parent[v] >= 0) ++first[parent[v] + 1];
  }
  std::partial_sum(first.begin(), first.end(), first.begin());
  std::vector<int> fill(first.begin(), first.end() - 1);
  for (int v = 0; v < n; ++v) {
    if (v != source && parent[v] >= 0) kids[fill[parent[v]]++] = v;
  }
  for (int k = 0; k < K; ++k) costs[static_cast<size_t>(source) * K + k] = 0;
  std::vector<int> queue{source};
  for (size_t i = 0; i < queue.size(); ++i) {
    const int p = queue[i];
    for (int c = first[p]; c < first[p + 1]; ++c) {
      const int v = kids[c];
      int edge = -1;
      for (int e = graph.rowPtr[p]; e < graph.rowPtr[p + 1]; ++e) {
        if (graph.colInd[e] == v) {
          edge = e;
          break;
        }
      }
      if (edge < 0) return false;
      for (int k = 0; k < K; ++k) {
        costs[static_cast<size_t>(v) * K + k] =
            costs[static_cast<size_t>(p) * K + k] + graph.weight(edge, k);
      }
      queue.push_back(v);
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
  double total = combined;
  for (double t : objectives) total += t;
  return total;
}

void canonicalizeTree(const CsrGraph &graph, int objective, int source,
                      const std::vector<long long> &distances, int *parent) {
  for (int u = 0; u < graph.numberOfNodes; ++u) {
    const long long du = distances[u];
    if (du >= DISTANCE_INF / 2) continue;
    for (int e = graph.rowPtr[u]; e < graph.rowPtr[u + 1]; ++e) {
      const int v = graph.colInd[e];
      if (v != source && u < parent[v] && du + graph.weight(e, objective) == distances[v]) {
        parent[v] = u;
      }
    }
  }
}

bool mospUpdate(const CsrGraph &original, ChangeBatch &batch,
                const std::vector<long long> &initialDistances,
                const std::vector<int> &initialParents, const MospOptions &options,
                CsrGraph &updated, MospResult &result) {
  const int n = original.numberOfNodes;
  const int KG = original.numberOfObjectives;
  const int K = options.numberOfObjectives > 0 ? std::min(options.numberOfObjectives, KG) : KG;
  const size_t treeSize = static_cast<size_t>(K) * n;
  result = MospResult();
  result.numberOfObjectives = K;
  MospTimings &tm = result.timings;
  if (n <= 0 || K <= 0 || options.source < 0 || options.source >= n ||
      initialDistances.size() < treeSize || initialParents.size() < treeSize) {
    std::cerr << "Error: invalid MOSP update input.\n";
    return false;
  }
  for (int k = 0; k < K; ++k) {  // trees rooted at the source (mospUpdate.cu:77-92)
    const size_t at = static_cast<size_t>(k) * n + options.source;
    if (initialDistances[at] != 0 || initialParents[at] != -1) {
      std::cerr << "Error: the initial tree of objective " << k
                << " is not rooted at the source " << options.source << " (its distance is ";
      if (initialDistances[at] >= DISTANCE_INF / 2) {
        std::cerr << "INF";
      } else {
        std::cerr << initialDistances[at];
      }
      std::cerr << ", its parent " << initialParents[at]
                << "); compute the trees for this source.\n";
      return false;
    }
  }
  if (preferenceScale(options.preferences, K) == 0) {
    std::cerr << "Error: invalid preference vector (need K values >= 1).\n";
    return false;
  }

  auto t0 = std::chrono::steady_clock::now();
  if (!applyChangeBatch(original, batch, updated)) return false;
  tm.applyBatch = ms_since(t0);

  // Per objective: largest weight before and after the batch, weight sum of
  // the original edges, changed edges (mospUpdate.cu:104-136).
  const size_t m = static_cast<size_t>(original.numberOfEdges());
  std::vector<long long> maxWeight(K, 1), weightSum(K, 0);
  for (size_t e = 0; e < m; ++e) {
    for (int k = 0; k < K; ++k) {
      const int w = original.weights[e * KG + k];
      maxWeight[k] = std::max(maxWeight[k], static_cast<long long>(w));
      weightSum[k] += w;
    }
  }
  for (int i = 0; i < batch.numberOfInserts(); ++i) {
    for (int k = 0; k < K; ++k) {
      maxWeight[k] = std::max(
          maxWeight[k], static_cast<long long>(batch.insertWeights[static_cast<size_t>(i) * KG + k]));
    }
  }
  std::vector<std::vector<int>> chFrom(K), chTo(K);
  for (int k = 0; k < K; ++k) {
    chFrom[k] = batch.deleteFrom;
    chTo[k] = batch.deleteTo;
    for (int i = 0; i < batch.numberOfInserts(); ++i) {
      if (batch.weightIncreaseMask[i] & (1u << k)) {
        chFrom[k].push_back(batch.insertFrom[i]);
        chTo[k].push_back(batch.insertTo[i]);
      }
    }
  }

  t0 = std::chrono::steady_clock::now();
  DeviceGraph graph;
  DeviceArray<long long> dDist, dCombDist(n);
  DeviceArray<int> dPar, dCombPar(n), dHeads;
  std::vector<DeviceArray<int>> dFrom(K), dTo(K);
  SospWorkspace ws;
  {
    std::vector<long long> dist(initialDistances.begin(), initialDistances.begin() + treeSize);
    std::vector<int> par(initialParents.begin(), initialParents.begin() + treeSize);
    bool ok = uploadDeviceGraph(updated, graph) && dDist.upload(dist) && dPar.upload(par) &&
              dHeads.upload(batch.insertTo) && ws.reserve(n, options.fused ? K : 1);
    for (int k = 0; ok && k < K; ++k) {
      ok = dFrom[k].upload(chFrom[k]) && dTo[k].upload(chTo[k]);
    }
    if (!ok || hipDeviceSynchronize() != hipSuccess) {
      std::cerr << "Error: could not copy the MOSP inputs to the GPU.\n";
      return false;
    }
    if (!mosp_ws_set_combined_delta_scale(ws.handle, std::max(1, options.combinedDeltaScale))) {
      std::cerr << mosp_last_error() << "\n";
      return false;
    }
  }
  tm.upload = ms_since(t0);

  const long long edges = std::max<long long>(updated.numberOfEdges(), 1);
  std::vector<long long> delta(K);
  std::vector<mosp_changes> changes(K);
  for (int k = 0; k < K; ++k) {
    const long long rule =
        defaultDelta(m > 0 ? static_cast<long long>(m) : edges, n, weightSum[k]);
    delta[k] = options.delta > 0
                   ? options.delta
                   : std::max(1LL, static_cast<long long>(static_cast<double>(rule) *
                                                          std::max(options.deltaScale, 0.0)));
    changes[k] = mosp_changes{dFrom[k].data(), dTo[k].data(), static_cast<int32_t>(chFrom[k].size()),
                              dHeads.data(), batch.numberOfInserts()};
  }
  result.objectiveStats.resize(K);
  const std::vector<int> &pref = options.preferences;
  if (options.fused) {
    std::vector<mosp_sosp_stats> st(K);
    mosp_combine_stats cs{};
    t0 = std::chrono::steady_clock::now();
    const bool ok = mosp_update_fused(
        graph.handle, K, changes.data(), options.source,
        reinterpret_cast<const int64_t *>(delta.data()),
        reinterpret_cast<const int64_t *>(maxWeight.data()), pref.data(),
        static_cast<int32_t>(pref.size()), ws.handle, nullptr, nullptr,
        reinterpret_cast<int64_t *>(dDist.data()), dPar.data(),
        reinterpret_cast<int64_t *>(dCombDist.data()), dCombPar.data(), st.data(), &cs,
        nullptr);
    const double total = ms_since(t0);
    if (!ok) {
      report_error();
      std::cerr << "Error: MOSP update on the GPU failed.\n";
      return false;
    }
    for (int k = 0; k < K; ++k) from_c(st[k], result.objectiveStats[k]);
    CombineStats &c = result.combineStats;
    c.numberOfEdges = cs.n_edges;
    c.scale = cs.scale;
    c.delta = cs.delta;
    from_c(cs.search, c.search);
    // Split the host wall time of the one launch by the kernel's own clock
    // stamps (end of the K updates); gpuCompute() is the total either way.
    const double phases = static_cast<double>(cs.sosp_phase_ms) + cs.combined_phase_ms;
    const double sospFrac = phases > 0 ? cs.sosp_phase_ms / phases : 1.0;
    tm.objectives.assign(K, total * sospFrac / K);
    tm.combined = total * (1.0 - sospFrac);
  } else {
    for (int k = 0; k < K; ++k) {
      t0 = std::chrono::steady_clock::now();
      mosp_sosp_stats st{};
      mosp_csr out{}, in{};
      mosp_device_graph_view(graph.handle, k, &out, &in);
      if (!mosp_sosp_update(&out, &in, &changes[k], options.source, delta[k], maxWeight[k],
                            ws.handle,
                            reinterpret_cast<int64_t *>(dDist.data()) + static_cast<size_t>(k) * n,
                            dPar.data() + static_cast<size_t>(k) * n, &st, nullptr)) {
        report_error();
        std::cerr << "Error: SOSP update of objective " << k << " failed.\n";
        return false;
      }
      tm.objectives.push_back(ms_since(t0));
      from_c(st, result.objectiveStats[k]);
    }
    t0 = std::chrono::steady_clock::now();
    CombineWorkspace cws;
    if (!combinedGraphSospGpu(dPar.data(), n, K, pref, options.source, 0, cws, ws,
                              dCombDist.data(), dCombPar.data(), &result.combineStats)) {
      std::cerr << "Error: combined graph step failed.\n";
      return false;
    }
    tm.combined = ms_since(t0);
  }

  t0 = std::chrono::steady_clock::now();
  if (!dDist.download(result.distances) || !dPar.download(result.parents) ||
      !dCombDist.download(result.combinedDistances) ||
      !dCombPar.download(result.combinedParent)) {
    std::cerr << "Error: could not copy the results from the GPU.\n";
    return false;
  }
  tm.download = ms_since(t0);
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
