// mosp_prep -- synthetic inputs and initial trees for bin/mosp.
//
//   mosp_prep road  <prefix> W H K seed [drop=0.3] [diag=0.02] [wmin=1] [wmax=100]
//   mosp_prep rmat  <prefix> scale edgeFactor K seed [wmin=1] [wmax=100]
//   mosp_prep changes <prefix> <dir> changes insPct safe(0/1) seed [wmin] [wmax]
//   mosp_prep init  <prefix> <dir> [source=0]
//   mosp_prep expected <prefix> <changesDir> <dir> [source=0]
//                              ground truth of an update: the batch applied
//                              to the graph (in HBM), then a from-scratch
//                              search per objective of the updated graph ->
//                              <dir>/obj<k>/distancesUpdated.txt +
//                              SSSPTreeUpdated.txt (the reference's
//                              `mospPrep expected`, src/mospPrep.cu:360-375,
//                              with the GPU search instead of host Dijkstra;
//                              both give the canonical lowest-id tree)
//   mosp_prep cache <prefix> <binaryPath>
//                              write the binary cache `mosp --cache` reads
//   mosp_prep mtx2csr <in.mtx> <prefix> K wmin wmax seed
//                              Matrix Market coordinate file -> CSR text:
//                              symmetric matrices get both directions,
//                              self-loops and duplicates are dropped, each
//                              edge gets K seeded weights in [wmin, wmax]
//   mosp_prep check <prefix> [cacheFile]
//                              read the CSR text (host only; through the binary
//                              cache when one is named); print n, m, K and
//                              order-sensitive checksums of the arrays
//   mosp_prep copy  <MiB>      one device-to-device copy of MiB mebibytes
//                              (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)
// `init` writes <dir>/obj<k>/distancesOriginal.txt + SSSPTreeOriginal.txt:
// the canonical tree of every objective, computed by the GPU from-scratch
// search (the role of the reference's `mospPrep init`, src/mospPrep.cu:
// 289-304, 350-359, done on the device instead of host Dijkstra).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "csrGraph.h"
#include "deviceArray.h"
#include "deviceGraph.h"
#include "mosp_hip.h"
#include "sospUpdateGpu.h"

namespace {

double argd(int argc, char **argv, int i, double dflt) {
  return i < argc ? std::atof(argv[i]) : dflt;
}

bool make_graph(int n, int64_t m, int K, std::vector<int> &row, std::vector<int> &col,
                std::vector<int> &w, CsrGraph &g) {
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  g.rowPtr = row;
  g.colInd = col;
  g.weights = w;
  return static_cast<int64_t>(g.numberOfEdges()) == m;
}

// Canonical tree of objective k of a device graph from `source`: the
// from-scratch search with the default width and the graph's own bound.
bool scratch_tree(mosp_device_graph *g, int k, int source, SospWorkspace &ws,
                  std::vector<long long> &dist, std::vector<int> &par) {
  const int K = mosp_device_graph_objectives(g);
  std::vector<int64_t> maxW(K), sum(K);
  mosp_csr out{};
  if (!mosp_device_graph_weight_stats(g, maxW.data(), sum.data(), nullptr) ||
      !mosp_device_graph_view(g, k, &out, nullptr)) {
    return false;
  }
  const int n = out.n;
  DeviceArray<long long> d(n);
  DeviceArray<int> p(n);
  mosp_sosp_stats st{};
  if (!mosp_sosp_from_scratch(&out, source, mosp_default_delta(out.m, n, sum[k]),
                              std::max<int64_t>(1, maxW[k]), ws.handle,
                              reinterpret_cast<int64_t *>(d.data()), p.data(), &st, nullptr)) {
    return false;
  }
  return d.download(dist) && p.download(par);
}

// Unbiased integer in [lo, hi] (rejection on the top of the 64-bit range).
int bounded(std::mt19937_64 &rng, int lo, int hi) {
  const uint64_t span = static_cast<uint64_t>(hi) - static_cast<uint64_t>(lo) + 1;
  const uint64_t limit = UINT64_MAX - UINT64_MAX % span;
  uint64_t x = rng();
  while (x >= limit) x = rng();
  return lo + static_cast<int>(x % span);
}

// Matrix Market coordinate -> CSR (rows sorted by column, no self-loops or
// duplicates; symmetric entries in both directions).
bool mtx_to_csr(const std::string &path, int K, int wmin, int wmax, uint64_t seed, CsrGraph &g) {
  std::ifstream in(path);
  std::string line;
  if (!std::getline(in, line) || line.rfind("%%MatrixMarket", 0) != 0) {
    std::cerr << "Error: " << path << " is not a Matrix Market file\n";
    return false;
  }
  const bool symmetric = line.find("symmetric") != std::string::npos;
  while (std::getline(in, line) && (line.empty() || line[0] == '%')) {
  }
  long long rows = 0, cols = 0, entries = 0;
  if (!(std::istringstream(line) >> rows >> cols >> entries) || rows <= 0 || rows != cols ||
      rows > 2147483646LL) {
    std::cerr << "Error: bad size line in " << path << "\n";
    return false;
  }
  const int n = static_cast<int>(rows);
  std::vector<std::pair<int, int>> arcs;
  arcs.reserve(static_cast<size_t>(entries) * (symmetric ? 2 : 1));
  for (long long i = 0; i < entries; ++i) {
    long long a = 0, b = 0;
    if (!std::getline(in, line) || !(std::istringstream(line) >> a >> b) || a < 1 || a > n ||
        b < 1 || b > n) {
      std::cerr << "Error: bad entry " << i + 1 << " in " << path << "\n";
      return false;
    }
    if (a == b) continue;
    arcs.emplace_back(static_cast<int>(a - 1), static_cast<int>(b - 1));
    if (symmetric) arcs.emplace_back(static_cast<int>(b - 1), static_cast<int>(a - 1));
  }
  std::sort(arcs.begin(), arcs.end());
  arcs.erase(std::unique(arcs.begin(), arcs.end()), arcs.end());
  g = CsrGraph();
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  g.rowPtr.assign(static_cast<size_t>(n) + 1, 0);
  for (const auto &e : arcs) ++g.rowPtr[static_cast<size_t>(e.first) + 1];
  for (int v = 0; v < n; ++v) g.rowPtr[v + 1] += g.rowPtr[v];
  std::mt19937_64 rng(seed);
  for (const auto &e : arcs) {
    g.colInd.push_back(e.second);
    for (int k = 0; k < K; ++k) g.weights.push_back(bounded(rng, wmin, wmax));
  }
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 3) {
    std::cerr << "usage: mosp_prep road|rmat|changes|init ...\n";
    return 2;
  }
  const std::string cmd = argv[1], prefix = argv[2];
  if (cmd == "copy") {
    const size_t bytes = static_cast<size_t>(std::atoll(argv[2])) << 20;
    DeviceArray<char> a(bytes), b(bytes);
    if (hipMemset(a.data(), 1, bytes) != hipSuccess ||
        hipMemcpy(b.data(), a.data(), bytes, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
      std::cerr << "copy failed\n";
      return 1;
    }
    std::cout << "copied " << bytes << " bytes device to device\n";
    return 0;
  }
  if (cmd == "road" || cmd == "rmat") {
    int32_t n = 0;
    int64_t m = 0;
    if (argc < 7) {
      std::cerr << "usage: mosp_prep road <prefix> W H K seed | rmat <prefix> scale ef K seed\n";
      return 2;
    }
    const int K = std::atoi(argv[5]);
    const uint64_t seed = std::strtoull(argv[6], nullptr, 10);
    auto gen = [&](int32_t *row, int32_t *col, int32_t *w, int64_t cap) {
      if (cmd == "road") {
        return mosp_gen_road_grid(std::atoi(argv[3]), std::atoi(argv[4]), K,
                                  static_cast<int>(argd(argc, argv, 9, 1)),
                                  static_cast<int>(argd(argc, argv, 10, 100)),
                                  argd(argc, argv, 7, 0.3), argd(argc, argv, 8, 0.02), seed, &n,
                                  &m, row, col, w, cap);
      }
      return mosp_gen_rmat(std::atoi(argv[3]), std::atoi(argv[4]), 0.57, 0.19, 0.19, K,
                           static_cast<int>(argd(argc, argv, 7, 1)),
                           static_cast<int>(argd(argc, argv, 8, 100)), seed, &n, &m, row, col, w,
                           cap);
    };
    gen(nullptr, nullptr, nullptr, 0);
    std::vector<int> row(n + 1), col(m), w(static_cast<size_t>(m) * K);
    gen(row.data(), col.data(), w.data(), m);
    CsrGraph g;
    make_graph(n, m, K, row, col, w, g);
    if (!writeCsrGraph(prefix, g)) return 1;
    std::cout << "wrote " << prefix << " n=" << n << " m=" << m << " K=" << K << "\n";
    return 0;
  }
  if (cmd == "mtx2csr") {
    if (argc != 8) {
      std::cerr << "usage: mosp_prep mtx2csr <in.mtx> <prefix> K wmin wmax seed\n";
      return 2;
    }
    const int K = std::atoi(argv[4]), wmin = std::atoi(argv[5]), wmax = std::atoi(argv[6]);
    if (K < 1 || K > 32 || wmin < 1 || wmax < wmin) {
      std::cerr << "Error: need 1 <= K <= 32 and 1 <= wmin <= wmax\n";
      return 2;
    }
    CsrGraph g;
    if (!mtx_to_csr(argv[2], K, wmin, wmax, std::strtoull(argv[7], nullptr, 10), g) ||
        !writeCsrGraph(argv[3], g)) {
      return 1;
    }
    std::cout << "wrote " << argv[3] << " n=" << g.numberOfNodes << " m=" << g.numberOfEdges()
              << " K=" << K << "\n";
    return 0;
  }
  CsrGraph g;
  // `check <prefix> <cacheFile>` reads through the binary cache (loadCsrGraph:
  // used when it matches the text files, else rebuilt with a note).
  if (cmd == "check" && argc >= 4 ? !loadCsrGraph(prefix, g, argv[3]) : !readCsrGraph(prefix, g)) {
    return 1;
  }
  const int n = g.numberOfNodes, K = g.numberOfObjectives;
  if (cmd == "cache" && argc == 4) {
    if (!saveCsrGraphBinary(argv[3], g, prefix)) {
      std::cerr << "Error: could not write " << argv[3] << "\n";
      return 1;
    }
    std::cout << "wrote the binary cache " << argv[3] << "\n";
    return 0;
  }
  if (cmd == "expected" && argc >= 5) {
    const std::string changes = argv[3], dir = argv[4];
    const int source = static_cast<int>(argd(argc, argv, 5, 0));
    if (source < 0 || source >= n) {
      std::cerr << "Error: source vertex out of range\n";
      return 1;
    }
    ChangeBatch b;
    if (!readChangeBatch(changes + "/insert.txt", changes + "/delete.txt", K, n, b)) return 1;
    mosp_device_graph *dg = mosp_device_graph_upload(n, K, g.rowPtr.data(), g.colInd.data(),
                                                     g.weights.data());
    mosp_device_graph *du = mosp_device_graph_create();
    SospWorkspace ws;
    bool ok = dg != nullptr && ws.reserve(n) &&
              mosp_device_graph_apply(dg, b.numberOfInserts(), b.insertFrom.data(),
                                      b.insertTo.data(), b.insertWeights.data(),
                                      b.numberOfDeletes(), b.deleteFrom.data(), b.deleteTo.data(),
                                      du, nullptr, nullptr);
    for (int k = 0; ok && k < K; ++k) {
      std::vector<long long> dist;
      std::vector<int> par;
      const std::string obj = dir + "/obj" + std::to_string(k);
      ok = scratch_tree(du, k, source, ws, dist, par) &&
           writeDistances(obj + "/distancesUpdated.txt", dist) &&
           writeParents(obj + "/SSSPTreeUpdated.txt", par);
    }
    mosp_device_graph_free(du);
    mosp_device_graph_free(dg);
    if (!ok) {
      std::cerr << mosp_last_error() << "\n";
      return 1;
    }
    std::cout << "wrote the expected trees of " << K << " objectives to " << dir << "\n";
    return 0;
  }
  if (cmd == "check") {
    // Position-weighted sums mod 2^64 (sum of (i + 1) * a[i]) of the row
    // pointers, columns and edge-major weights: order-sensitive checksums.
    auto sum = [](const std::vector<int> &a) {
      unsigned long long s = 0;
      for (size_t i = 0; i < a.size(); ++i) {
        s += (i + 1) * static_cast<unsigned long long>(static_cast<unsigned>(a[i]));
      }
      return s;
    };
    std::cout << "n=" << n << " m=" << g.numberOfEdges() << " K=" << K
              << " rows=" << sum(g.rowPtr) << " cols=" << sum(g.colInd)
              << " weights=" << sum(g.weights) << "\n";
    return 0;
  }
  if (cmd == "changes" && argc >= 8) {
    const std::string dir = argv[3];
    const int changes = std::atoi(argv[4]);
    std::vector<int> inF(changes), inT(changes), inW(static_cast<size_t>(changes) * K),
        dF(changes), dT(changes);
    int32_t nIns = 0, nDel = 0;
    mosp_gen_batch(n, K, g.rowPtr.data(), g.colInd.data(), g.weights.data(), changes,
                   std::atof(argv[5]), std::atoi(argv[6]),
                   static_cast<int>(argd(argc, argv, 8, 1)),
                   static_cast<int>(argd(argc, argv, 9, 100)),
                   std::strtoull(argv[7], nullptr, 10), 0, &nIns, inF.data(), inT.data(),
                   inW.data(), &nDel, dF.data(), dT.data());
    ChangeBatch b;
    b.numberOfObjectives = K;
    b.insertFrom.assign(inF.begin(), inF.begin() + nIns);
    b.insertTo.assign(inT.begin(), inT.begin() + nIns);
    b.insertWeights.assign(inW.begin(), inW.begin() + static_cast<size_t>(nIns) * K);
    b.deleteFrom.assign(dF.begin(), dF.begin() + nDel);
    b.deleteTo.assign(dT.begin(), dT.begin() + nDel);
    if (!writeChangeBatch(b, dir + "/insert.txt", dir + "/delete.txt")) return 1;
    std::cout << "wrote " << dir << ": " << nIns << " inserts, " << nDel << " deletes\n";
    return 0;
  }
  if (cmd == "init" && argc >= 4) {
    const std::string dir = argv[3];
    const int source = static_cast<int>(argd(argc, argv, 4, 0));
    DeviceGraph dg;
    SospWorkspace ws;
    DeviceArray<long long> d(n);
    DeviceArray<int> p(n);
    if (!uploadDeviceGraph(g, dg) || !ws.reserve(n)) return 1;
    for (int k = 0; k < K; ++k) {
      long long maxW = 1, sum = 0;
      for (int e = 0; e < g.numberOfEdges(); ++e) {
        maxW = std::max<long long>(maxW, g.weight(e, k));
        sum += g.weight(e, k);
      }
      std::vector<long long> dist;
      std::vector<int> par;
      if (!sospFromScratchGpu(dg.out(k), source, defaultDelta(g.numberOfEdges(), n, sum), maxW,
                              ws, d.data(), p.data()) ||
          !d.download(dist) || !p.download(par)) {
        return 1;
      }
      const std::string obj = dir + "/obj" + std::to_string(k);
      if (!writeDistances(obj + "/distancesOriginal.txt", dist) ||
          !writeParents(obj + "/SSSPTreeOriginal.txt", par)) {
        return 1;
      }
    }
    std::cout << "wrote " << K << " trees to " << dir << "\n";
    return 0;
  }
  std::cerr << "usage: mosp_prep road|rmat|mtx2csr|changes|init|expected|cache|check|copy ...\n";
  return 2;
}
