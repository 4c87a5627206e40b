// deviceGraph.h -- a multi-objective graph resident in HBM (out- and in-CSR,
// one weight column per objective).  Drop-in for the reference's
// headers/deviceGraph.cuh:16-37 (same names, out(k)/in(k) views).
#ifndef MOSP_DEVICE_GRAPH_H
#define MOSP_DEVICE_GRAPH_H

#include "mosp_hip.h"
#include "sospUpdateGpu.h"

struct CsrGraph;

struct DeviceGraph {
  DeviceGraph() = default;
  ~DeviceGraph();
  DeviceGraph(const DeviceGraph &) = delete;
  DeviceGraph &operator=(const DeviceGraph &) = delete;

  int numberOfNodes = 0;
  int numberOfEdges = 0;
  int numberOfObjectives = 0;
  mosp_device_graph *handle = nullptr;

  DeviceCsr out(int k) const;  // out-edge view of objective k
  DeviceCsr in(int k) const;   // in-edge view of objective k
};

// Copy the out-CSR once; the reverse CSR and the objective-major weight
// columns are built on the GPU.
bool uploadDeviceGraph(const CsrGraph &graph, DeviceGraph &device);

#endif  // MOSP_DEVICE_GRAPH_H
