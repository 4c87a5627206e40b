// parallelCombinedGraph.h -- the file-based combined graph and its host
// helpers.  Drop-in for the reference's headers/parallelCombinedGraph.cuh:
// 46-89 (parallelCombinedGraph, preferenceScale, combinedEdgeWeight,
// mospPathCosts) with the same signatures and defaults.
#ifndef MOSP_PARALLEL_COMBINED_GRAPH_H
#define MOSP_PARALLEL_COMBINED_GRAPH_H

#include <string>
#include <vector>

struct CsrGraph;

// Reads the vertex count from <prefix>RowPtr.txt and the K parent files,
// builds the combined graph of the K trees and its SOSP tree on the GPU
// (combinedGraphSospGpu) and writes its distances (units of 1/L) and tree in
// the tree file format.  workDir is unused (kept for the signature).
bool parallelCombinedGraph(
    const std::string &originalCsrPrefix,
    const std::vector<std::string> &treeInputPaths, int K, int source = 0,
    const std::string &workDir = "output/combinedGraph",
    const std::string &distancesOutputPath = "output/combinedGraph/distancesCsr.txt",
    const std::string &treeOutputPath = "output/combinedGraph/SSSPTreeCsr.txt",
    const std::vector<int> &preferences = std::vector<int>());

// L = lcm(Pref); 0 if Pref is not K values >= 1 or L > 2^20; 1 if empty.
long long preferenceScale(const std::vector<int> &preferences, int K);

// L * (K + 1) - sum over the trees in treeMask of L / Pref_i.
long long combinedEdgeWeight(unsigned int treeMask,
                             const std::vector<int> &preferences, int K,
                             long long scale);

// costs[v * K + k]: objective-k length of the MOSP tree path to v (first
// (p,v) edge of row p of `graph`), DISTANCE_INF if unreachable; false if a
// tree edge is missing.  Host version; mosp_path_costs_gpu is the device one.
bool mospPathCosts(const CsrGraph &graph, const std::vector<int> &parent,
                   int source, std::vector<long long> &costs);

#endif  // MOSP_PARALLEL_COMBINED_GRAPH_H
