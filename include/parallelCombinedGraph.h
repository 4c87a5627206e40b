// parallelCombinedGraph.h -- host helpers of the combined graph.  Drop-in for
// the reference's headers/parallelCombinedGraph.cuh:62-89 (preferenceScale,
// combinedEdgeWeight, mospPathCosts) with the same signatures.
#ifndef MOSP_PARALLEL_COMBINED_GRAPH_H
#define MOSP_PARALLEL_COMBINED_GRAPH_H

#include <vector>

struct CsrGraph;

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
