// engine_internal.hpp -- host-side helpers shared by the engine's sources
// (engine.hip: single launches and the device graph; batch.hip: batched
// launches).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "engine.hpp"
#include "mosp_hip.h"

// A graph in HBM (device_graph.hip): out-CSR + in-CSR, each with K weight
// columns of `stride` (>= m) entries, objective k at w + k * stride.
// Capacities let a spare graph be rebuilt in place by mosp_device_graph_apply.
struct mosp_device_graph {
  int n = 0;
  int m = 0;
  int K = 0;
  long long stride = 0;    // weight column length (edge capacity)
  int rowCap = 0;          // vertex capacity of the row arrays
  int *outRow = nullptr, *outCol = nullptr, *outW = nullptr;
  int *inRow = nullptr, *inCol = nullptr, *inW = nullptr;
  // Objective k's edges as {column, weight} records at k * stride, out and
  // in (built after every upload and apply).
  int2 *outEdge = nullptr, *inEdge = nullptr;
  // Objective k's out-edges as {column, weight, rowBegin(column),
  // rowEnd(column)} at k * stride, for the asynchronous near-far loop of
  // graphs without hubs; built on demand (graph_records4) and stale after
  // every upload and apply.  A cache of the arrays above, so it may be
  // built through a const graph (one launch on a graph at a time).
  mutable int4 *outEdge4 = nullptr;
  mutable bool rec4Valid = false;
  char *scratch = nullptr;  // build / apply temporaries (grow-only)
  size_t scratchBytes = 0;
  unsigned long long *stats = nullptr;  // weight statistics (2 * MOSP_MAX_OBJ words) + degree
  int maxDegree = 0;       // largest out- or in-degree
};

namespace mosp {

// Kernels and their launchers (sosp_kernel.hip).
// hubs: the variant for graphs with vertices of more than HUB_DEGREE edges.
// The persistent kernels are built for two workgroup shapes: 1024 threads,
// one workgroup per CU (16 waves per CU), and 512 threads, up to two per CU.
namespace shape1024 {
const void *sosp_kernel_symbol(bool hubs);
const void *sosp_batched_symbol(bool hubs);
}  // namespace shape1024
namespace shape512 {
const void *sosp_kernel_symbol(bool hubs);
const void *sosp_batched_symbol(bool hubs);
}  // namespace shape512

hipError_t launch_ctl_reset(Ctl *ctl, hipStream_t stream);
// abortWord (may be null): a launch-wide abort word cleared with the blocks.
hipError_t launch_ctl_reset_many(Ctl *ctls, int count, hipStream_t stream,
                                 unsigned *abortWord = nullptr);

namespace detail {
struct KernelShape {
  int threads;  // per workgroup
  const void *(*persistent)(bool hubs);
  const void *(*batched)(bool hubs);
};
// The shape of `threads` threads per workgroup (512; anything else: 1024).
const KernelShape &kernel_shape(int threads);
// Workgroups of a shape that fit on the device at once (0 on error).
int shape_max_blocks(const KernelShape &shape, bool batched);

// Objective k's {column, weight} record arrays of a device graph (out, in).
void graph_edges(const mosp_device_graph *g, int k, const int2 **out, const int2 **in);

// Largest out- or in-degree (inRow may be null) of a CSR with n rows, via
// one device word `scratch`; synchronous on `stream` (device_graph.hip).
int graph_max_degree(const int *outRow, const int *inRow, int n, unsigned *scratch,
                     hipStream_t stream, int *result);

// MOSP_HUBS=0|1 forces the kernel variant (tests, tuning); -1 when unset.
int hubs_override();
// Local relaxation rounds per near-far iteration: K updates, and the
// from-scratch / combined searches (measured best on MI355X for road-like
// graphs, single and batched launches).
constexpr int kHopsUpdate = 2;
constexpr int kHopsSearch = 6;
// MOSP_LOCAL_HOPS=u,s overrides the local relaxation rounds per near-far
// iteration of the K updates (u) and of the searches (s) (tuning).
void local_hops_override(int &update, int &search);
// Longest wait at a grid barrier before a launch gives up (ms): 2000, or
// MOSP_BARRIER_TIMEOUT_MS.  Far above any legitimate wait (a phase's
// imbalance is microseconds to milliseconds).
double barrier_timeout_ms_default();
// wall_clock64 ticks of `ms` on the current device (rate in kHz).
unsigned long long barrier_ticks(double ms, double clockKhz);
// Launches `kernel` cooperatively; if the runtime refuses the cooperative
// launch in this process's present state (another process holds the
// device's cooperative-launch resources), launches the same grid plainly.
// The grid never exceeds the occupancy limit, so on a device running no
// other persistent kernel every workgroup is resident; otherwise the
// bounded barrier waits end the launch with an error.  MOSP_LAUNCH=plain
// makes every such launch plain (the same grid, without the cooperative API).
hipError_t launch_persistent(const void *kernel, dim3 grid, dim3 block, void **params,
                             hipStream_t stream);
// MOSP_INIT_PUSH (LaunchArgs::initPush): 0 unset, 1 pull, 2 push.
int init_push_override();
// MOSP_LIST_CAP=c caps the near / far lists at c entries instead of n, so
// that small inputs take the sweep that follows a full list (tests); 0 unset.
int list_cap_override();
// MOSP_HUB_EDGES=e: the hub list takes push vertices of more than e edges
// (at least 64; default HUB_GRID).
int hub_edges_override();

// Records the message for mosp_last_error() and returns 0.
int fail(const std::string &message);
// 0 (with the error recorded, prefixed by `where`) when a launch's pack pass
// found an input parent or distance that cannot belong to the graph.
int input_tree_ok(unsigned overflow, const std::string &where);
void clear_error();
hipStream_t as_stream(void *s);
size_t align256(size_t bytes);

#define HIP_TRY(call)                                                      \
  do {                                                                     \
    hipError_t err_ = (call);                                              \
    if (err_ != hipSuccess) {                                              \
      return ::mosp::detail::fail(std::string("HIP error: ") +             \
                                  hipGetErrorString(err_) + " at " +       \
                                  __FILE__ + ":" + std::to_string(__LINE__)); \
    }                                                                      \
  } while (0)

// Packing of (distance << b | parent) words for n vertices
// (src/sospUpdateGpu.cu:127-143, 670-681).
struct PackChoice {
  int bits;
  u64 noParent;
  u64 bound;        // (n-1) * maxWeight: no shortest distance is larger
  u64 maxDistance;  // candidates beyond it are dropped
  bool optimistic;  // maxDistance < bound: a launch that drops one is redone
};

// Packed (distance << bits | parent) words when every distance up to the
// bound fits beside the parent; otherwise, with `optimistic`, packed words
// that hold distances up to their own limit (a launch that meets a larger
// one stops before writing and is redone without), else distance-only words.
int choose_packing(int n, long long maxWeight, PackChoice &pc, bool optimistic = false);
// Pointer-jumping rounds for n vertices: ceil(log2 n) + 1.
int jump_rounds(int n);
void fill_stats(const Ctl &c, int k, const ObjDesc &d, float ms,
                mosp_sosp_stats *stats);
ObjDesc make_desc(const mosp_csr &out, const mosp_csr *in,
                  const mosp_changes *ch, const PackChoice &pc, u64 delta,
                  const ObjState &st);
// Carves the scratch arrays of one objective for n vertices out of p
// (advanced past them); state_bytes(n) is their size.
// The out-edge records with row ranges of g (built when stale).
int graph_records4(const mosp_device_graph *g, hipStream_t stream);
// Entries of one pool segment of the asynchronous near-far loop for n
// vertices, and the bytes of the ASYNC_SEGS segments.
unsigned pool_seg_cap(int n);
size_t pool_bytes(unsigned segCap);
size_t state_bytes(int n, unsigned segCap = 0);
ObjState carve_state(char *&p, int n, unsigned segCap = 0);

}  // namespace detail
}  // namespace mosp
