// mosp -- command-line driver of the MOSP update on MI355X.
//
// Drop-in for the reference's bin/mosp (src/mosp.cu:1-420): same options,
// input files, output files and the stable summary line
//   RESULT gpu_compute_ms=<a> end_to_end_ms=<b>
// (a) = the K SOSP updates + Steps 2-3 (host wall clock around the
// synchronous GPU calls), (b) = reading inputs to writing outputs.
//
// usage: mosp --graph <csrPrefix> --changes <dir> --init <dir>
//             [-k K] [--source s] [--pref p1,..,pK] [--delta D]
//             [--cache file] [--canonicalize] [--out dir] [--no-output]
//             [--validate] [--timing file.csv] [--quiet] [--serial]
// --serial runs K launches + 1 like the reference instead of one fused
// launch.  --validate checks every tree against a from-scratch GPU search
// on the updated graph and the MOSP tree against a host-built combined
// graph searched from scratch (distances and lowest-id parents).
#include <hip/hip_runtime_api.h>

#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "csrGraph.h"
#include "deviceArray.h"
#include "deviceGraph.h"
#include "mospUpdate.h"
#include "mosp_hip.h"
#include "parallelCombinedGraph.h"
#include "sospUpdateGpu.h"

namespace {

struct Options {
  std::string graph, changes, init, out = "mosp-output", timingCsv, cache;
  std::vector<int> pref;
  int K = 0, source = 0;
  long long delta = 0;
  double deltaScale = 1.0;
  int combDeltaScale = 1;
  bool canonicalize = false, writeOutput = true, validate = false, quiet = false;
  bool serial = false;
};

void usage() {
  std::cerr << "usage: mosp --graph <csrPrefix> --changes <dir> --init <dir>\n"
               "            [-k K] [--source s] [--pref p1,..,pK] [--delta D]\n"
               "            [--cache file] [--canonicalize] [--out dir]\n"
               "            [--no-output] [--validate] [--timing file.csv]\n"
               "            [--quiet] [--serial] [--comb-delta-scale S] [--delta-scale F]\n";
}

bool parse(int argc, char **argv, Options &o) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](std::string &dst) {
      if (i + 1 >= argc) return false;
      dst = argv[++i];
      return true;
    };
    std::string v;
    long long x = 0;
    auto integer = [&](long long lo, long long hi) {
      if (!next(v)) return false;
      if (parseInteger(v, lo, hi, x)) return true;
      std::cerr << a << " needs an integer in [" << lo << ", " << hi << "], got '" << v << "'\n";
      return false;
    };
    if (a == "--graph") {
      if (!next(o.graph)) return false;
    } else if (a == "--changes") {
      if (!next(o.changes)) return false;
    } else if (a == "--init") {
      if (!next(o.init)) return false;
    } else if (a == "--out") {
      if (!next(o.out)) return false;
    } else if (a == "--timing") {
      if (!next(o.timingCsv)) return false;
    } else if (a == "--cache") {
      if (!next(o.cache)) return false;
    } else if (a == "-k") {
      if (!integer(0, INT_MAX)) return false;
      o.K = static_cast<int>(x);
    } else if (a == "--source") {
      if (!integer(0, INT_MAX)) return false;
      o.source = static_cast<int>(x);
    } else if (a == "--delta") {
      if (!integer(1, LLONG_MAX)) return false;
      o.delta = x;
    } else if (a == "--delta-scale") {
      if (!next(v)) return false;
      try {
        o.deltaScale = std::stod(v);
      } catch (...) {
        o.deltaScale = -1;
      }
      if (!(o.deltaScale > 0)) {
        std::cerr << "--delta-scale needs a positive number, got '" << v << "'\n";
        return false;
      }
    } else if (a == "--comb-delta-scale") {
      if (!integer(1, INT_MAX)) return false;
      o.combDeltaScale = static_cast<int>(x);
    } else if (a == "--pref") {
      if (!next(v)) return false;
      std::istringstream list(v);
      std::string item;
      while (std::getline(list, item, ',')) {
        if (!parseInteger(item, 1, INT_MAX, x)) {
          std::cerr << "--pref needs integers >= 1, got '" << v << "'\n";
          return false;
        }
        o.pref.push_back(static_cast<int>(x));
      }
    } else if (a == "--canonicalize") {
      o.canonicalize = true;
    } else if (a == "--no-output") {
      o.writeOutput = false;
    } else if (a == "--validate") {
      o.validate = true;
    } else if (a == "--quiet") {
      o.quiet = true;
    } else if (a == "--serial") {
      o.serial = true;
    } else {
      std::cerr << "unknown option: " << a << "\n";
      return false;
    }
  }
  return !o.graph.empty() && !o.changes.empty() && !o.init.empty();
}

double ms_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

template <typename T>
std::vector<T> slice(const std::vector<T> &all, int k, int n) {
  return std::vector<T>(all.begin() + static_cast<size_t>(k) * n,
                        all.begin() + static_cast<size_t>(k + 1) * n);
}

bool write_costs(const std::string &path, const std::vector<long long> &costs, int K) {
  FILE *f = std::fopen(path.c_str(), "w");
  if (f == nullptr) return false;
  const size_t n = costs.size() / K;
  std::string buf;
  buf.reserve(1 << 22);
  bool ok = true;
  for (size_t v = 0; v < n; ++v) {
    buf += std::to_string(v);
    for (int k = 0; k < K; ++k) {
      const long long c = costs[v * K + k];
      buf.push_back(' ');
      buf += c >= DISTANCE_INF / 2 ? std::string("INF") : std::to_string(c);
    }
    buf.push_back('\n');
    if (buf.size() > (1 << 22)) {
      ok = ok && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
      buf.clear();
    }
  }
  ok = ok && std::fwrite(buf.data(), 1, buf.size(), f) == buf.size();
  return std::fclose(f) == 0 && ok;
}

// Lowest-id tight in-neighbour check of a tree (distances equal to `ref`).
struct Check {
  long long distance = 0, parent = 0;
};

Check compare(const std::vector<long long> &d, const std::vector<int> &p,
              const std::vector<long long> &rd, const std::vector<int> &rp) {
  Check c;
  for (size_t v = 0; v < d.size(); ++v) {
    const bool fin = d[v] < DISTANCE_INF / 2, rfin = rd[v] < DISTANCE_INF / 2;
    if (fin != rfin || (fin && d[v] != rd[v])) ++c.distance;
    if (p[v] != rp[v]) ++c.parent;
  }
  return c;
}

// Validation by an independent GPU path: from-scratch searches of the
// updated graph (no incremental state) and of a host-built combined graph.
bool validate(const CsrGraph &updated, const MospResult &r, int source,
              const std::vector<int> &pref) {
  const int n = updated.numberOfNodes, K = r.numberOfObjectives;
  DeviceGraph g;
  SospWorkspace ws;
  DeviceArray<long long> d(n);
  DeviceArray<int> p(n);
  if (!uploadDeviceGraph(updated, g) || !ws.reserve(n)) return false;
  bool all = true;
  for (int k = 0; k < K; ++k) {
    long long maxW = 1, sum = 0;
    for (int e = 0; e < updated.numberOfEdges(); ++e) {
      maxW = std::max<long long>(maxW, updated.weight(e, k));
      sum += updated.weight(e, k);
    }
    std::vector<long long> rd;
    std::vector<int> rp;
    if (!sospFromScratchGpu(g.out(k), source, defaultDelta(updated.numberOfEdges(), n, sum),
                            maxW, ws, d.data(), p.data()) ||
        !d.download(rd) || !p.download(rp)) {
      return false;
    }
    const Check c = compare(slice(r.distances, k, n), slice(r.parents, k, n), rd, rp);
    const bool ok = c.distance == 0 && c.parent == 0;
    all = all && ok;
    std::cout << "VALIDATE obj" << k << " " << (ok ? "PASS" : "FAIL")
              << " (distance mismatches=" << c.distance << " parent mismatches=" << c.parent
              << ")\n";
  }
  // Combined graph of the K new trees, built on the host.
  const long long L = preferenceScale(pref, K);
  CsrGraph comb;
  comb.numberOfNodes = n;
  comb.numberOfObjectives = 1;
  comb.rowPtr.assign(n + 1, 0);
  std::vector<std::vector<std::pair<int, int>>> rows(n);
  for (int v = 0; v < n; ++v) {
    if (v == source) continue;
    for (int k = 0; k < K; ++k) {
      const int par = r.parents[static_cast<size_t>(k) * n + v];
      bool seen = par < 0;
      for (int j = 0; j < k && !seen; ++j) seen = r.parents[static_cast<size_t>(j) * n + v] == par;
      if (seen) continue;
      unsigned mask = 0;
      for (int j = 0; j < K; ++j) {
        if (r.parents[static_cast<size_t>(j) * n + v] == par) mask |= 1u << j;
      }
      rows[par].push_back({v, static_cast<int>(combinedEdgeWeight(mask, pref, K, L))});
    }
  }
  long long sum = 0;
  for (int u = 0; u < n; ++u) {
    comb.rowPtr[u + 1] = comb.rowPtr[u] + static_cast<int>(rows[u].size());
    for (auto &e : rows[u]) {
      comb.colInd.push_back(e.first);
      comb.weights.push_back(e.second);
      sum += e.second;
    }
  }
  DeviceGraph cg;
  std::vector<long long> rd;
  std::vector<int> rp;
  if (!uploadDeviceGraph(comb, cg) ||
      !sospFromScratchGpu(cg.out(0), source, defaultDelta(comb.numberOfEdges(), n, sum),
                          L * (K + 1), ws, d.data(), p.data()) ||
      !d.download(rd) || !p.download(rp)) {
    return false;
  }
  const Check c = compare(r.combinedDistances, r.combinedParent, rd, rp);
  const bool ok = c.distance == 0 && c.parent == 0;
  std::cout << "VALIDATE combined " << (ok ? "PASS" : "FAIL") << " (distance mismatches="
            << c.distance << " parent mismatches=" << c.parent << ")\n";
  return all && ok;
}

}  // namespace

int main(int argc, char **argv) {
  Options opt;
  if (!parse(argc, argv, opt)) {
    usage();
    return 2;
  }
  const auto start = std::chrono::steady_clock::now();
  double contextMs = 0, loadMs = 0, canonMs = 0, writeMs = 0;
  {
    const auto t = std::chrono::steady_clock::now();
    (void)hipFree(nullptr);  // create the context outside the timed region
    contextMs = ms_since(t);
  }
  auto t = std::chrono::steady_clock::now();
  CsrGraph original;
  ChangeBatch batch;
  if (!loadCsrGraph(opt.graph, original, opt.cache)) return 1;
  const int n = original.numberOfNodes, KG = original.numberOfObjectives;
  const int K = opt.K > 0 ? std::min(opt.K, KG) : KG;
  std::vector<long long> distances(static_cast<size_t>(K) * n);
  std::vector<int> parents(static_cast<size_t>(K) * n);
  {
    std::vector<std::vector<long long>> dist(K);
    std::vector<std::vector<int>> par(K);
    std::vector<std::function<bool()>> reads{[&] {
      return readChangeBatch(opt.changes + "/insert.txt", opt.changes + "/delete.txt", KG, n,
                             batch);
    }};
    for (int k = 0; k < K; ++k) {
      const std::string dir = opt.init + "/obj" + std::to_string(k);
      reads.push_back([&, k, dir] { return readDistances(dir + "/distancesOriginal.txt", n, dist[k]); });
      reads.push_back([&, k, dir] { return readParents(dir + "/SSSPTreeOriginal.txt", n, par[k]); });
    }
    if (!runConcurrently(reads)) return 1;
    for (int k = 0; k < K; ++k) {
      std::copy(dist[k].begin(), dist[k].end(), distances.begin() + static_cast<size_t>(k) * n);
      std::copy(par[k].begin(), par[k].end(), parents.begin() + static_cast<size_t>(k) * n);
    }
  }
  loadMs = ms_since(t);
  if (opt.canonicalize) {
    t = std::chrono::steady_clock::now();
    for (int k = 0; k < K; ++k) {
      canonicalizeTree(original, k, opt.source, slice(distances, k, n),
                       parents.data() + static_cast<size_t>(k) * n);
    }
    canonMs = ms_since(t);
  }
  MospOptions options;
  options.source = opt.source;
  options.numberOfObjectives = K;
  options.preferences = opt.pref;
  options.delta = opt.delta;
  options.deltaScale = opt.deltaScale;
  options.fused = !opt.serial;
  options.combinedDeltaScale = opt.combDeltaScale;
  CsrGraph updated;
  MospResult result;
  if (!mospUpdate(original, batch, distances, parents, options, updated, result)) return 1;

  if (opt.writeOutput) {
    const auto tw = std::chrono::steady_clock::now();
    namespace fs = std::filesystem;
    for (int k = 0; k < K; ++k) fs::create_directories(opt.out + "/obj" + std::to_string(k));
    fs::create_directories(opt.out + "/combinedGraph");
    // MOSP path costs on the GPU (all KG objectives of the graph).
    std::vector<long long> costs;
    bool costsOk = false;
    {
      DeviceGraph g;
      DeviceArray<int> dp;
      DeviceArray<long long> dc(static_cast<size_t>(n) * KG);
      if (uploadDeviceGraph(updated, g) && dp.upload(result.combinedParent) &&
          mosp_path_costs_gpu(g.handle, dp.data(), opt.source,
                              reinterpret_cast<int64_t *>(dc.data()), nullptr)) {
        costsOk = dc.download(costs);
      } else {
        std::cerr << mosp_last_error() << "\n";
      }
    }
    const std::string cdir = opt.out + "/combinedGraph";
    std::vector<std::function<bool()>> writes;
    for (int k = 0; k < K; ++k) {
      const std::string dir = opt.out + "/obj" + std::to_string(k);
      writes.push_back([&, k, dir] {
        return writeDistances(dir + "/distancesUpdated.txt", slice(result.distances, k, n));
      });
      writes.push_back([&, k, dir] {
        return writeParents(dir + "/SSSPTreeUpdated.txt", slice(result.parents, k, n));
      });
    }
    writes.push_back([&] { return writeDistances(cdir + "/distancesCsr.txt", result.combinedDistances); });
    writes.push_back([&] { return writeParents(cdir + "/SSSPTreeCsr.txt", result.combinedParent); });
    writes.push_back([&] { return costsOk && write_costs(cdir + "/mospCosts.txt", costs, KG); });
    if (!runConcurrently(writes)) {
      std::cerr << "cannot write the outputs to " << opt.out << "\n";
      return 1;
    }
    writeMs = ms_since(tw);
  }
  const double endToEnd = ms_since(start);
  const MospTimings &tm = result.timings;
  if (!opt.quiet) {
    std::printf("graph  n=%d m=%d K=%d; batch %d inserts, %d deletes\n", n,
                updated.numberOfEdges(), K, batch.numberOfInserts(), batch.numberOfDeletes());
    std::printf("host   context %.1f ms, read inputs %.1f ms, canonicalize %.1f ms, apply batch "
                "%.1f ms, upload %.1f ms, download %.1f ms, write %.1f ms\n",
                contextMs, loadMs, canonMs, tm.applyBatch, tm.upload, tm.download, writeMs);
    for (int k = 0; k < K; ++k) {
      const SospStats &s = result.objectiveStats[k];
      std::printf("obj%d   SOSP update %.3f ms (invalidated %d, jump rounds %d, iterations %d, "
                  "epochs %d, pushes %lld)\n",
                  k, tm.objectives[k], s.invalidated, s.jumpRounds, s.iterations, s.epochs,
                  s.pushes);
    }
    const CombineStats &c = result.combineStats;
    std::printf("comb   combined graph + SOSP %.3f ms (%d edges, L=%lld, delta %lld, iterations "
                "%d, pushes %lld)\n",
                tm.combined, c.numberOfEdges, c.scale, c.delta, c.search.iterations,
                c.search.pushes);
  }
  if (!opt.timingCsv.empty()) {
    std::ofstream csv(opt.timingCsv);
    csv << "kind,name,value\n";
    csv << "stage,read_inputs," << loadMs << "\n";
    csv << "stage,apply_batch_host," << tm.applyBatch << "\n";
    csv << "stage,upload," << tm.upload << "\n";
    for (int k = 0; k < K; ++k) csv << "stage,obj" << k << "/sosp_update_gpu," << tm.objectives[k] << "\n";
    csv << "stage,combined_graph_gpu," << tm.combined << "\n";
    csv << "stage,download," << tm.download << "\n";
    csv << "stage,write_outputs," << writeMs << "\n";
    csv << "stage,TOTAL_end_to_end," << endToEnd << "\n";
  }
  std::printf("RESULT gpu_compute_ms=%.3f end_to_end_ms=%.3f\n", tm.gpuCompute(), endToEnd);
  std::fflush(stdout);
  if (opt.validate && !validate(updated, result, opt.source, opt.pref)) return 1;
  return 0;
}
