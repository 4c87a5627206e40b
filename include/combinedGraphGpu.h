// combinedGraphGpu.h -- Steps 2-3 of the MOSP update on the GPU (combined
// graph of the K trees + its SOSP tree).  Drop-in for the reference's
// headers/combinedGraphGpu.cuh:15-69 (same names and signatures).
#ifndef MOSP_COMBINED_GRAPH_GPU_H
#define MOSP_COMBINED_GRAPH_GPU_H

#include <vector>

#include "sospUpdateGpu.h"

struct CombineStats {
  int numberOfEdges = 0;
  long long scale = 1;   // L: combined distances are in units of 1/L
  long long delta = 0;   // near-far width used for Step 3
  SospStats search;
};

// The combined-graph buffers live in the SospWorkspace passed alongside;
// this type keeps the reference's call signature and records the size.
struct CombineWorkspace {
  CombineWorkspace() = default;
  CombineWorkspace(const CombineWorkspace &) = delete;
  CombineWorkspace &operator=(const CombineWorkspace &) = delete;
  bool reserve(int numberOfNodes, int K) {
    capacity = numberOfNodes > capacity ? numberOfNodes : capacity;
    trees = K > trees ? K : trees;
    return true;
  }
  void release() { capacity = trees = 0; }
  int capacity = 0;
  int trees = 0;
};

bool combinedGraphSospGpu(const int *d_parents, int numberOfNodes, int K,
                          const std::vector<int> &preferences, int source,
                          long long delta, CombineWorkspace &workspace,
                          SospWorkspace &sospWorkspace, long long *d_distances,
                          int *d_parent, CombineStats *stats = nullptr);

#endif  // MOSP_COMBINED_GRAPH_GPU_H
