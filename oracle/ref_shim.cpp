// ref_shim.cpp -- C entry points over the reference's own host C++ sources.
//
// TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile together with the
// reference's host files (compiled from where they lie under the reference
// checkout) into oracle/_ref/libmospref.so; used by tests/ to pin the C
// restatement in mosp_oracle.c against the reference's own code on the same
// inputs.  No reference source is copied: this file only converts flat arrays
// to the reference's CsrGraph / ChangeBatch types and calls:
//   dijkstraCsrGraph      (src/Dijkstra.cu:239-272)
//   applyChangeBatch      (src/csrGraph.cu:727-894)
//   transposeCsrGraph     (src/csrGraph.cu:896-922)
//   sequentialSOSPUpdate  (src/sequentialSOSPUpdate.cu:293-658, file based)
//   generateChangeBatch   (src/changeGenerator.cu:133-344)
//   generateGraphCSR      (src/generateGraphCSR.cu)
//   generateChangedEdges  (src/generateChangedEdges.cu)
//   readCsrGraph/readChangeBatch/readDistances/readParents (src/csrGraph.cu)
//   saveCsrGraphBinary/loadCsrGraphBinary (src/csrGraph.cu:503-596)
#include "changeGenerator.cuh"
#include "csrGraph.cuh"
#include "dijkstra.cuh"
#include "generateChangedEdges.cuh"
#include "generateGraphCSR.cuh"
#include "sequentialSOSPUpdate.cuh"

#include <cstring>
#include <string>
#include <vector>

namespace {

CsrGraph makeGraph(int n, int K, const int *rowPtr, const int *colInd,
                   const int *w) {
  CsrGraph g;
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  const int m = rowPtr[n];
  g.rowPtr.assign(rowPtr, rowPtr + n + 1);
  g.colInd.assign(colInd, colInd + m);
  g.weights.assign(w, w + static_cast<size_t>(m) * K);
  return g;
}

} // namespace

extern "C" {

void ref_dijkstra(int n, const int *rowPtr, const int *colInd, const int *w,
                  int K, int objective, int source, long long *dist,
                  int *parent) {
  CsrGraph g = makeGraph(n, K, rowPtr, colInd, w);
  std::vector<long long> d;
  std::vector<int> p;
  dijkstraCsrGraph(g, objective, source, d, p);
  std::memcpy(dist, d.data(), sizeof(long long) * n);
  std::memcpy(parent, p.data(), sizeof(int) * n);
}

int ref_apply_batch(int n, int K, const int *rowPtr, const int *colInd,
                    const int *w, int nIns, const int *insFrom,
                    const int *insTo, const int *insW, int nDel,
                    const int *delFrom, const int *delTo, int *outRowPtr,
                    int *outColInd, int *outW, unsigned *incMask) {
  CsrGraph g = makeGraph(n, K, rowPtr, colInd, w), updated;
  ChangeBatch b;
  b.numberOfObjectives = K;
  b.insertFrom.assign(insFrom, insFrom + nIns);
  b.insertTo.assign(insTo, insTo + nIns);
  b.insertWeights.assign(insW, insW + static_cast<size_t>(nIns) * K);
  b.deleteFrom.assign(delFrom, delFrom + nDel);
  b.deleteTo.assign(delTo, delTo + nDel);
  if (!applyChangeBatch(g, b, updated)) {
    return -1;
  }
  const int m = updated.numberOfEdges();
  std::memcpy(outRowPtr, updated.rowPtr.data(), sizeof(int) * (n + 1));
  std::memcpy(outColInd, updated.colInd.data(), sizeof(int) * m);
  std::memcpy(outW, updated.weights.data(), sizeof(int) * m * K);
  std::memcpy(incMask, b.weightIncreaseMask.data(), sizeof(unsigned) * nIns);
  return m;
}

int ref_sequential_update_files(const char *csrPrefix, const char *distIn,
                                const char *treeIn, const char *insertPath,
                                const char *deletePath, int objective,
                                int source, const char *distOut,
                                const char *treeOut) {
  return sequentialSOSPUpdate(csrPrefix, distIn, treeIn, insertPath,
                              deletePath, objective, source, distOut, treeOut)
             ? 1
             : 0;
}

// Reference batch generator; writes insert/delete files (writeChangeBatch).
int ref_generate_change_batch_files(const char *csrPrefix, int numberOfChanges,
                                    double insertionPercentage, int mode,
                                    int weightMin, int weightMax,
                                    unsigned seed, int localHops, int safe,
                                    int source, const char *insertPath,
                                    const char *deletePath) {
  CsrGraph g;
  if (!readCsrGraph(csrPrefix, g)) {
    return 0;
  }
  ChangeGeneratorOptions o;
  o.numberOfChanges = numberOfChanges;
  o.insertionPercentage = insertionPercentage;
  o.mode = static_cast<ChangeMode>(mode);
  o.weightMin = weightMin;
  o.weightMax = weightMax;
  o.seed = seed;
  o.localHops = localHops;
  o.safeDeletions = safe != 0;
  o.source = source;
  ChangeBatch b;
  if (!generateChangeBatch(g, o, b)) {
    return 0;
  }
  return writeChangeBatch(b, insertPath, deletePath) ? 1 : 0;
}

int ref_generate_graph_csr(int n, int m, int directed, const char *prefix,
                           int K, int wMin, int wMax, unsigned seed) {
  return generateGraphCSR(n, m, directed != 0, prefix, K, wMin, wMax, seed)
             ? 1
             : 0;
}

int ref_generate_changed_edges(int wMin, int wMax, int K, int n, int changes,
                               double insPct, double delPct, int directed,
                               int exist, int duplicate, int selfLoop,
                               const char *csrPrefix, const char *insertPath,
                               const char *deletePath, unsigned seed) {
  return generateChangedEdges(wMin, wMax, K, n, changes, insPct, delPct,
                              directed != 0, exist != 0, duplicate != 0,
                              selfLoop != 0, csrPrefix, insertPath, deletePath,
                              seed)
             ? 1
             : 0;
}

// Reader checks (input-validation group): 1 accepted, 0 rejected.
int ref_read_csr_ok(const char *prefix) {
  CsrGraph g;
  return readCsrGraph(prefix, g) ? 1 : 0;
}

int ref_read_batch_ok(const char *insertPath, const char *deletePath, int K,
                      int n) {
  ChangeBatch b;
  return readChangeBatch(insertPath, deletePath, K, n, b) ? 1 : 0;
}

int ref_read_distances_ok(const char *path, int n) {
  std::vector<long long> d;
  return readDistances(path, n, d) ? 1 : 0;
}

int ref_read_parents_ok(const char *path, int n) {
  std::vector<int> p;
  return readParents(path, n, p) ? 1 : 0;
}

// Binary cache of the text graph at `prefix`, written by the reference's
// own saveCsrGraphBinary: 1 written, 0 failed.
int ref_save_cache(const char *prefix, const char *path) {
  CsrGraph g;
  return readCsrGraph(prefix, g) && saveCsrGraphBinary(path, g, prefix) ? 1 : 0;
}

// 1 if the reference's loadCsrGraphBinary accepts `path` as the cache of
// `prefix` and the loaded arrays equal the text graph's; 0 otherwise.
int ref_load_cache_ok(const char *path, const char *prefix) {
  CsrGraph text, cached;
  if (!readCsrGraph(prefix, text) || !loadCsrGraphBinary(path, cached, prefix)) {
    return 0;
  }
  return text.numberOfNodes == cached.numberOfNodes &&
                 text.numberOfObjectives == cached.numberOfObjectives &&
                 text.rowPtr == cached.rowPtr && text.colInd == cached.colInd &&
                 text.weights == cached.weights
             ? 1
             : 0;
}

} // extern "C"
