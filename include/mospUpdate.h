// mospUpdate.h -- the in-memory MOSP update (K SOSP updates + combined graph
// + its SOSP tree on the GPU).  Drop-in for the reference's
// headers/mospUpdate.cuh:19-85 (same names and signatures); MospOptions has
// one extra field, `fused`.
#ifndef MOSP_MOSP_UPDATE_H
#define MOSP_MOSP_UPDATE_H

#include <vector>

#include "combinedGraphGpu.h"

struct CsrGraph;
struct ChangeBatch;

struct MospOptions {
  int source = 0;
  int numberOfObjectives = 0;    // 0: every objective of the graph
  std::vector<int> preferences;  // Pref; empty = all 1s
  long long delta = 0;           // near-far width; <= 0: default per objective
  // true: the K updates and Steps 2-3 run in ONE persistent launch (the
  // objectives advance together); false: K launches + 1, as the reference.
  bool fused = true;
  // Step 3's near-far width = the default rule times this (>= 1).  Tuning
  // only (no reference counterpart); outputs do not depend on it.
  int combinedDeltaScale = 1;
  // With delta <= 0: each objective's width = max(1, trunc(default rule *
  // deltaScale)).  Tuning only (no reference counterpart), as above.
  double deltaScale = 1.0;
};

struct MospTimings {
  double applyBatch = 0;
  double upload = 0;
  // GPU: SOSP update of each objective.  With fused launches the K updates
  // run together; each entry then holds an equal share of their time.
  std::vector<double> objectives;
  double combined = 0;  // GPU: combined graph + its SOSP tree
  double download = 0;
  double gpuCompute() const;  // K updates + Steps 2-3 (the timed region)
};

struct MospResult {
  int numberOfObjectives = 0;
  std::vector<long long> distances;          // K * n, objective-major
  std::vector<int> parents;                  // K * n, objective-major
  std::vector<long long> combinedDistances;  // units of 1/L
  std::vector<int> combinedParent;           // the MOSP tree
  std::vector<SospStats> objectiveStats;
  CombineStats combineStats;
  MospTimings timings;
};

bool mospUpdate(const CsrGraph &original, ChangeBatch &batch,
                const std::vector<long long> &initialDistances,
                const std::vector<int> &initialParents,
                const MospOptions &options, CsrGraph &updated,
                MospResult &result);

// Lowest-id tie rule on a tree: parent[v] <- u for every tight edge with
// u < parent[v] (one pass over the out-edges).
void canonicalizeTree(const CsrGraph &graph, int objective, int source,
                      const std::vector<long long> &distances, int *parent);

#endif  // MOSP_MOSP_UPDATE_H
