// sospUpdateGpu.h -- device-resident SOSP update engine, reference-named C++
// API over the C ABI of mosp_hip.h.
//
// Drop-in for the reference's headers/sospUpdateGpu.cuh:14-110: same type
// names, members used by callers, and function signatures.  The workspace
// wraps an opaque mosp_ws (the scratch layout is this engine's own).
#ifndef MOSP_SOSP_UPDATE_GPU_H
#define MOSP_SOSP_UPDATE_GPU_H

#include "mosp_hip.h"

// A CSR graph with one weight per edge, in device memory (.cuh:14-20).
struct DeviceCsr {
  int numberOfNodes = 0;
  int numberOfEdges = 0;
  const int *rowPtr = nullptr;
  const int *colInd = nullptr;
  const int *weights = nullptr;
};

// The change batch as seen by one objective, device memory (.cuh:22-32).
struct DeviceChanges {
  const int *changedFrom = nullptr;
  const int *changedTo = nullptr;
  int numberOfChanged = 0;
  const int *insertHeads = nullptr;
  int numberOfInsertHeads = 0;
};

// Counters of one update (.cuh:35-45); kernelMs is the HIP-event time of the
// persistent launch.
struct SospStats {
  int invalidated = 0;
  int jumpRounds = 0;
  int iterations = 0;
  int epochs = 0;
  long long pushes = 0;
  bool packedParents = true;
  float kernelMs = 0;
};

// Device scratch reused by every objective and the combined step; reserve()
// also sizes the persistent grid (.cuh:53-78).  `slots` objectives can be
// updated in one launch.
struct SospWorkspace {
  SospWorkspace() = default;
  ~SospWorkspace();
  SospWorkspace(const SospWorkspace &) = delete;
  SospWorkspace &operator=(const SospWorkspace &) = delete;

  bool reserve(int capacity, int slots = 1);
  void release();

  int capacity = 0;
  int gridBlocks = 0;
  mosp_ws *handle = nullptr;
};

long long defaultDelta(long long numberOfEdges, int numberOfNodes,
                       long long weightSum);

bool sospUpdateGpu(const DeviceCsr &out, const DeviceCsr &in,
                   const DeviceChanges &changes, int source, long long delta,
                   long long maxWeight, SospWorkspace &workspace,
                   long long *d_distances, int *d_parent,
                   SospStats *stats = nullptr);

bool sospFromScratchGpu(const DeviceCsr &out, int source, long long delta,
                        long long maxWeight, SospWorkspace &workspace,
                        long long *d_distances, int *d_parent,
                        SospStats *stats = nullptr);

#endif  // MOSP_SOSP_UPDATE_GPU_H
