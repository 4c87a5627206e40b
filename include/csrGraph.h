// csrGraph.h -- host graph, change batches and the text / binary file
// layouts of the DynaMOSP update.
//
// Drop-in for the reference's headers/csrGraph.cuh:18-187: same type names,
// members, function names, signatures and file formats, so host code written
// against the reference compiles unchanged against libmosp_hip.  The
// implementation (mosp-cuda_amd/csrc/host/csr_graph.cpp) is this project's.
#ifndef MOSP_CSR_GRAPH_H
#define MOSP_CSR_GRAPH_H

#include <functional>
#include <string>
#include <vector>

// Multi-objective directed graph in CSR form: row u lists the out-edges of u;
// objective k of edge e is weights[e * numberOfObjectives + k] (edge-major,
// the layout of the Values text file).  (csrGraph.cuh:18-29)
struct CsrGraph {
  int numberOfNodes = 0;
  int numberOfObjectives = 0;
  std::vector<int> rowPtr;   // numberOfNodes + 1 offsets
  std::vector<int> colInd;   // neighbour of each edge
  std::vector<int> weights;  // numberOfObjectives weights per edge

  int numberOfEdges() const { return rowPtr.empty() ? 0 : rowPtr.back(); }
  int weight(int edge, int objective) const {
    return weights[static_cast<size_t>(edge) * numberOfObjectives + objective];
  }
};

// A batch of edge changes (insert.txt / delete.txt).  weightIncreaseMask is
// filled by applyChangeBatch: bit k of entry i is set when insertion i
// overwrote an edge of the original graph and its final objective-k weight is
// larger than the original one.  (csrGraph.cuh:31-51)
struct ChangeBatch {
  int numberOfObjectives = 0;
  std::vector<int> insertFrom, insertTo;
  std::vector<int> insertWeights;  // insertion-major, K per insertion
  std::vector<int> deleteFrom, deleteTo;
  std::vector<unsigned int> weightIncreaseMask;

  int numberOfInserts() const { return static_cast<int>(insertFrom.size()); }
  int numberOfDeletes() const { return static_cast<int>(deleteFrom.size()); }
};

// <prefix>RowPtr.txt / ColInd.txt / Values.txt; rejects (message with file
// and line) decreasing or non-zero-based row pointers, column indices
// outside [0, n), ids above INT_MAX and weights outside [1, INT_MAX].
bool readCsrGraph(const std::string &prefix, CsrGraph &graph);
bool writeCsrGraph(const std::string &prefix, const CsrGraph &graph);

// Binary cache (magic MOSPCSR2) recording the identity of its text source:
// canonical prefix and size + mtime (ns) of the three files.
bool saveCsrGraphBinary(const std::string &path, const CsrGraph &graph,
                        const std::string &sourcePrefix);
bool loadCsrGraphBinary(const std::string &path, CsrGraph &graph,
                        const std::string &sourcePrefix);
bool loadCsrGraph(const std::string &prefix, CsrGraph &graph,
                  const std::string &cachePath = "");

// insert.txt ("u v w1 .. wK") and delete.txt ("u v").
bool readChangeBatch(const std::string &insertPath,
                     const std::string &deletePath, int numberOfObjectives,
                     int numberOfNodes, ChangeBatch &batch);
bool writeChangeBatch(const ChangeBatch &batch, const std::string &insertPath,
                      const std::string &deletePath);

// Deletions first (each removes the first remaining (u,v)), then insertions
// in batch order (overwrite the first alive (u,v) or append); survivors keep
// their order, new edges go last in their row.  Fills weightIncreaseMask.
bool applyChangeBatch(const CsrGraph &original, ChangeBatch &batch,
                      CsrGraph &updated);

// Reverse graph: row v lists u for every edge u -> v.
void transposeCsrGraph(const CsrGraph &graph, CsrGraph &reverse);

// "v d" / "v INF" and "v p" files; every vertex exactly once.
bool readDistances(const std::string &path, int numberOfNodes,
                   std::vector<long long> &distances);
bool readParents(const std::string &path, int numberOfNodes,
                 std::vector<int> &parent);
bool writeDistances(const std::string &path,
                    const std::vector<long long> &distances);
bool writeParents(const std::string &path, const std::vector<int> &parent);

// Run independent jobs concurrently; true if all returned true.
bool runConcurrently(const std::vector<std::function<bool()>> &jobs);

// Whole decimal integer in [minimum, maximum] (no sign '+', no blanks).
bool parseInteger(const std::string &text, long long minimum,
                  long long maximum, long long &value);

// Distance sentinel (values >= INF/2 print as "INF").
constexpr long long DISTANCE_INF = 0x7fffffffffffffffLL / 4;

#endif  // MOSP_CSR_GRAPH_H
