// mosp_prep -- synthetic inputs and initial trees for bin/mosp.
//
//   mosp_prep road  <prefix> W H K seed [drop=0.3] [diag=0.02] [wmin=1] [wmax=100]
//   mosp_prep rmat  <prefix> scale edgeFactor K seed [wmin=1] [wmax=100]
//   mosp_prep changes <prefix> <dir> changes insPct safe(0/1) seed [wmin] [wmax]
//   mosp_prep init  <prefix> <dir> [source=0]
//   mosp_prep check <prefix>   read the CSR text (host only); print n, m, K and
//                              order-sensitive checksums of the arrays
//   mosp_prep copy  <MiB>      one device-to-device copy of MiB mebibytes
//                              (calibrates rocprofv3 FETCH_SIZE / WRITE_SIZE)
// `init` writes <dir>/obj<k>/distancesOriginal.txt + SSSPTreeOriginal.txt:
// the canonical tree of every objective, computed by the GPU from-scratch
// search (the role of the reference's `mospPrep init`, src/mospPrep.cu:
// 289-304, 350-359, done on the device instead of host Dijkstra).
#include <algorithm>
#include <cstdlib>
#include <iostream>
#include <string>
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
  CsrGraph g;
  if (!readCsrGraph(prefix, g)) return 1;
  const int n = g.numberOfNodes, K = g.numberOfObjectives;
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
  std::cerr << "usage: mosp_prep road|rmat|changes|init ...\n";
  return 2;
}
