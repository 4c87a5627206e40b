/*
 * mosp_hip.h -- C ABI of the MI355X (gfx950) DynaMOSP update engine.
 *
 * Plain C: POD structs, raw pointers and sizes, int status codes (1 = ok,
 * 0 = error with the message in mosp_last_error()).  Device pointers are HIP
 * device addresses (hipMalloc or any allocator); `stream` is a hipStream_t
 * passed as void* (NULL = the default stream).  Every compute call is
 * synchronous on its stream, like the reference's entry points, and returns
 * after its results are in device memory.
 *
 * Each declaration names the reference interface it replaces.  The
 * reference-named C++ API (csrGraph.h, sospUpdateGpu.h, combinedGraphGpu.h,
 * deviceGraph.h, mospUpdate.h in this directory) is a thin layer over these.
 */
#ifndef MOSP_HIP_H
#define MOSP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types -- */

/* One-weight CSR view in device memory.
 * Replaces DeviceCsr (headers/sospUpdateGpu.cuh:14-20). */
typedef struct mosp_csr {
  int32_t n;                 /* vertices */
  int32_t m;                 /* edges */
  const int32_t *row_ptr;    /* n + 1 */
  const int32_t *col_ind;    /* m */
  const int32_t *weights;    /* m (one objective) */
} mosp_csr;

/* The change batch as seen by one objective, device memory.
 * Replaces DeviceChanges (headers/sospUpdateGpu.cuh:22-32). */
typedef struct mosp_changes {
  const int32_t *changed_from;  /* deleted or weight-increased edges (u,v) */
  const int32_t *changed_to;
  int32_t n_changed;
  const int32_t *insert_heads;  /* heads of inserted edges */
  int32_t n_insert_heads;
} mosp_changes;

/* Counters of one update.  Replaces SospStats (sospUpdateGpu.cuh:35-45);
 * kernel_ms is new (HIP-event time of the persistent launch). */
typedef struct mosp_sosp_stats {
  int32_t invalidated;
  int32_t jump_rounds;
  int32_t iterations;
  int32_t epochs;
  int64_t pushes;
  int32_t packed_parents;    /* 0: distance-only words + parent recovery */
  float kernel_ms;
} mosp_sosp_stats;

/* Replaces CombineStats (headers/combinedGraphGpu.cuh:15-20). */
typedef struct mosp_combine_stats {
  int32_t n_edges;
  int64_t scale;             /* L: combined distances in units of 1/L */
  int64_t delta;
  mosp_sosp_stats search;
  /* Device-clock split of a fused launch (mosp_update_fused): time of the
   * K updates and of the combined step inside the one kernel. */
  float sosp_phase_ms;
  float combined_phase_ms;
} mosp_combine_stats;

/* Device scratch of the engine; replaces SospWorkspace
 * (sospUpdateGpu.cuh:53-78) and CombineWorkspace (combinedGraphGpu.cuh:25-46). */
typedef struct mosp_ws mosp_ws;

/* A multi-objective graph resident in HBM: out-CSR + in-CSR with one weight
 * column per objective.  Replaces DeviceGraph (headers/deviceGraph.cuh:16-27). */
typedef struct mosp_device_graph mosp_device_graph;

/* ------------------------------------------------------------- device ----- */
/* Plumbing for hosts without their own HIP allocator (no reference
 * counterpart): select a device, allocate / free / copy device memory,
 * wait for the device.  kind: 1 host->device, 2 device->host,
 * 3 device->device. */
int mosp_set_device(int32_t device);
int32_t mosp_device_count(void);
/* Compute units of the current device (0 on error). */
int32_t mosp_device_cus(void);
void *mosp_dev_alloc(uint64_t bytes);
int mosp_dev_free(void *ptr);
int mosp_memcpy(void *dst, const void *src, uint64_t bytes, int32_t kind);
int mosp_device_synchronize(void);

/* ------------------------------------------------------------ errors ----- */

/* Message of the last failed call on this thread ("" if none). */
const char *mosp_last_error(void);

/* --------------------------------------------------------- workspace ----- */

/* Replaces SospWorkspace::SospWorkspace / reserve / release
 * (src/sospUpdateGpu.cu:527-595).  slots = objectives updated per launch. */
mosp_ws *mosp_ws_create(void);
int mosp_ws_reserve(mosp_ws *ws, int32_t capacity, int32_t slots);
void mosp_ws_destroy(mosp_ws *ws);
/* Co-resident workgroups of the persistent launch (0 = occupancy maximum). */
int mosp_ws_set_grid(mosp_ws *ws, int32_t workgroups);
/* Workgroups of the last launch (before any: of a 1024-thread launch). */
int32_t mosp_ws_grid(const mosp_ws *ws);
/* Longest wait of a workgroup at a grid barrier before the persistent
 * launch gives up (ms, default 2000 or MOSP_BARRIER_TIMEOUT_MS).  No
 * reference counterpart: the reference relies on its cooperative launch.
 * Here the launch is cooperative when the runtime allows it and otherwise a
 * plain launch of the same occupancy-bounded grid; a launch whose workgroups
 * are not all resident then ends with an error (no result written; the
 * workspace's scratch is cleared before its next launch) instead of hanging. */
int mosp_ws_set_barrier_timeout(mosp_ws *ws, double ms);
/* Workgroup shape of the persistent launches: `threads` 1024 (one
 * workgroup per CU, 16 waves) or 512 (two per CU), 0 = automatic (512 for
 * graphs with a vertex of more than 32 edges, else 1024); `queue_cap` =
 * LDS queue entries per list the push loop uses, 256..4096 (0 = all).
 * Also MOSP_WG_THREADS / MOSP_QUEUE_CAP at workspace creation.  No
 * reference counterpart: tuning knobs of this implementation; results do
 * not depend on them. */
int mosp_ws_set_shape(mosp_ws *ws, int32_t threads, int32_t queue_cap);

/* Near-far width of the combined search in the fused launch when none is
 * given: the reference's rule (src/sospUpdateGpu.cu:607-616, evaluated on the
 * device over the combined graph) times `scale` (>= 1; default 1).  A tuning
 * knob with no reference counterpart: outputs do not depend on it;
 * mosp_combine_stats.delta reports the width used. */
int mosp_ws_set_combined_delta_scale(mosp_ws *ws, int32_t scale);

/* Device-clock profile of the last launch on this workspace (no reference
 * counterpart; the reference's --timing stages are host-side): ms[i] = time
 * of phase i (pack, jump, invalidate, pull, near-far, unpack, combined count,
 * scan, fill, search, unpack; 0 if it did not run), *barriers = grid
 * barriers executed.  n_phases <= 11.  Returns the number of phases filled. */
int mosp_ws_last_profile(const mosp_ws *ws, double *ms, int32_t n_phases,
                         int32_t *barriers);

/* Asynchronous near-far loop of the last launch on this workspace (no
 * reference counterpart: the loop is this engine's own schedule).  Graphs
 * without hub vertices run their near-far phases without per-iteration grid
 * barriers (MOSP_ASYNC=0 selects the iteration-synchronous loop instead);
 * events[0] = batches that dropped near entries for lack of local queue
 * room, events[1] = spills that found less pool room than asked, events[2] =
 * phases that listed every vertex under the threshold again (after a drop),
 * events[3] = 1 if the launch ran the asynchronous loop.  Returns 1. */
int mosp_ws_async_events(const mosp_ws *ws, int32_t events[4]);

/* Replaces defaultDelta (src/sospUpdateGpu.cu:607-616). */
int64_t mosp_default_delta(int64_t m, int32_t n, int64_t weight_sum);

/* --------------------------------------------------------- SOSP engine --- */

/* Incremental SOSP update of one objective; d_dist / d_parent hold the old
 * canonical tree on entry and the new one on return.
 * Replaces sospUpdateGpu (src/sospUpdateGpu.cu:685-717). */
int mosp_sosp_update(const mosp_csr *out, const mosp_csr *in,
                     const mosp_changes *changes, int32_t source, int64_t delta,
                     int64_t max_weight, mosp_ws *ws, int64_t *d_dist,
                     int32_t *d_parent, mosp_sosp_stats *stats, void *stream);

/* mosp_sosp_update that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, semantics and
 * results as mosp_sosp_update, without the workspace: the scratch is one
 * static device array (36 bytes per vertex, for at most
 * MOSP_NOALLOC_MAX_VERTICES = 2,097,152 vertices; about 76 MB that each GPU
 * using the library holds until the process ends).  Larger graphs are
 * refused with an error.  Every call on a GPU uses that GPU's array, so
 * calls that overlap on one GPU race: run them one at a time per device.
 * Synchronous on `stream`; stats->kernel_ms is taken from the device clock. */
int mosp_sosp_update_noalloc(const mosp_csr *out, const mosp_csr *in,
                             const mosp_changes *changes, int32_t source,
                             int64_t delta, int64_t max_weight, int64_t *d_dist,
                             int32_t *d_parent, mosp_sosp_stats *stats, void *stream);

/* K objectives updated in ONE persistent launch (no reference counterpart;
 * it replaces the K serial sospUpdateGpu calls of mospUpdate.cu:170-194).
 * Arrays of length k; d_dist_in may equal d_dist_out (in place). */
int mosp_sosp_update_multi(int32_t k, const mosp_csr *out, const mosp_csr *in,
                           const mosp_changes *changes, int32_t source,
                           const int64_t *delta, const int64_t *max_weight,
                           mosp_ws *ws, const int64_t *const *d_dist_in,
                           const int32_t *const *d_parent_in,
                           int64_t *const *d_dist_out,
                           int32_t *const *d_parent_out, mosp_sosp_stats *stats,
                           void *stream);

/* Single-source shortest paths from scratch (canonical tree).
 * Replaces sospFromScratchGpu (src/sospUpdateGpu.cu:719-745). */
int mosp_sosp_from_scratch(const mosp_csr *out, int32_t source, int64_t delta,
                           int64_t max_weight, mosp_ws *ws, int64_t *d_dist,
                           int32_t *d_parent, mosp_sosp_stats *stats,
                           void *stream);

/* Combined graph of K trees (d_parents objective-major, K*n) and its SOSP
 * tree; delta <= 0 selects the default.  pref: K values >= 1, or n_pref = 0
 * for all ones.  Replaces combinedGraphSospGpu (src/combinedGraphGpu.cu:251-317). */
int mosp_combined_sosp(const int32_t *d_parents, int32_t n, int32_t k,
                       const int32_t *pref, int32_t n_pref, int32_t source,
                       int64_t delta, mosp_ws *ws, int64_t *d_dist,
                       int32_t *d_parent, mosp_combine_stats *stats,
                       void *stream);

/* The whole device part of the MOSP update in ONE launch: K objective updates
 * + combined graph + its search.  Replaces the K sospUpdateGpu calls and the
 * combinedGraphSospGpu call of mospUpdate (src/mospUpdate.cu:167-209).
 * d_dist/d_parents: K*n objective-major, old trees in (unless d_dist_in /
 * d_parents_in are given: then those are read and d_dist/d_parents written),
 * new trees out. */
int mosp_update_fused(const mosp_device_graph *g, int32_t k,
                      const mosp_changes *changes, int32_t source,
                      const int64_t *delta, const int64_t *max_weight,
                      const int32_t *pref, int32_t n_pref, mosp_ws *ws,
                      const int64_t *d_dist_in, const int32_t *d_parents_in,
                      int64_t *d_dist, int32_t *d_parents, int64_t *d_comb_dist,
                      int32_t *d_comb_parent, mosp_sosp_stats *stats,
                      mosp_combine_stats *cstats, void *stream);

/* ------------------------------------------------------------ batched --- */
/* Many independent update instances (own graph, K, batch, source,
 * preferences) in ONE persistent launch.  No counterpart in the reference,
 * whose host loop runs one mospUpdate (src/mospUpdate.cu:57-225) at a time:
 * each instance computes exactly what mosp_update_fused computes for it.
 * The workgroups are split into groups, one per instance (sized to its
 * estimated frontier, placed on one XCD when it fits); each group runs its
 * instance with its own barrier, and a group given several instances runs
 * them in turn. */
typedef struct mosp_batch_instance {
  const mosp_device_graph *graph;  /* updated graph in HBM */
  int32_t k;                       /* objectives updated (<= graph's) */
  const mosp_changes *changes;     /* k entries (device pointers inside) */
  int32_t source;
  const int64_t *delta;            /* k near-far widths (> 0) */
  const int64_t *max_weight;       /* k */
  const int32_t *pref;             /* n_pref ints (0 or k) */
  int32_t n_pref;
  const int64_t *d_dist_in;        /* k*n old trees (objective-major) */
  const int32_t *d_parents_in;
  int64_t *d_dist;                 /* k*n new trees */
  int32_t *d_parents;
  int64_t *d_comb_dist;            /* n combined-graph search */
  int32_t *d_comb_parent;
} mosp_batch_instance;

typedef struct mosp_batch mosp_batch;

/* Copies the descriptions (the device buffers stay the caller's) and
 * reserves every instance's scratch.  NULL on error (mosp_last_error). */
mosp_batch *mosp_batch_create(const mosp_batch_instance *instances,
                              int32_t count);
void mosp_batch_destroy(mosp_batch *b);
/* A batch with room for max_instances instances of max_objectives
 * objectives in all and slab_bytes of scratch (mosp_batch_instance_bytes per
 * instance), to be loaded by mosp_batch_reset: a resident worker reuses it
 * for every set of instances without allocating. */
mosp_batch *mosp_batch_create_capacity(int32_t max_instances,
                                       int32_t max_objectives,
                                       int64_t slab_bytes);
/* Replace the batch's instances (same rules as mosp_batch_create; fails if
 * they exceed the capacity) and clear the scratch they use. */
int mosp_batch_reset(mosp_batch *b, const mosp_batch_instance *instances,
                     int32_t count);
/* Scratch bytes one instance of n vertices and k objectives needs. */
int64_t mosp_batch_instance_bytes(int32_t n, int32_t k);
/* The batch's own HIP stream (non-blocking; its launches run there). */
void *mosp_batch_stream(const mosp_batch *b);
/* Launch shape: workgroups (0 = every co-resident workgroup of the
 * batch's workgroup shape: one per CU at 1024 threads), lanes per frontier vertex as
 * 2^lane_bits (< 0 = keep), each instance's widths scaled by delta_scale
 * (> 0; widths stay >= 1), xcd_aware != 0 places small groups on one XCD. */
int mosp_batch_configure(mosp_batch *b, int32_t workgroups, int32_t lane_bits,
                         double delta_scale, int32_t xcd_aware);
/* Group sizing model: an instance of n vertices and k objectives is timed
 * as (0.2 sqrt(n) + 10) * latency_items + k n vertex relaxations, divided
 * over its workgroups for the second term; groups are sized so that every
 * instance finishes at the same time.  latency_items (default 400) is the
 * number of vertices a workgroup relaxes in one iteration's latency. */
int mosp_batch_set_cost_model(mosp_batch *b, double latency_items);
/* Workgroup shape of the batched launch: 1024 threads (one workgroup per
 * CU; the default) or 512 (two per CU), keeping the batch's share of the
 * device; `queue_cap` as in mosp_ws_set_shape.  No reference counterpart;
 * results do not depend on it. */
int mosp_batch_set_shape(mosp_batch *b, int32_t threads, int32_t queue_cap);
/* One pass over every instance, synchronous.  use_graph 1 replays a HIP
 * graph of the launch sequence (captured on first use and after
 * mosp_batch_configure / mosp_batch_reset); 0 launches cooperatively; 2
 * launches the same grid without the cooperative API (for callers that run
 * several batches at once on disjoint shares of the CUs).  *kernel_ms
 * (optional) = device time of the pass. */
int mosp_batch_run(mosp_batch *b, int32_t use_graph, float *kernel_ms);
/* Longest grid-barrier wait of the batch's launches (ms, as
 * mosp_ws_set_barrier_timeout); a run that gives up returns 0 and the next
 * run starts from cleared scratch.  No reference counterpart. */
int mosp_batch_set_barrier_timeout(mosp_batch *b, double ms);
/* After a run: 0 with a message if an instance's old tree did not fit its
 * graph (the check mosp_update_fused makes after every launch). */
int mosp_batch_check(mosp_batch *b);
/* Statistics of instance i from the last run: stats[k] for its k
 * objectives, cstats for its combined step. */
int mosp_batch_stats(mosp_batch *b, int32_t i, mosp_sosp_stats *stats,
                     mosp_combine_stats *cstats);
/* Schedule: group_size[i] = workgroups running instance i; graph_mode =
 * 0 no graph yet, 1 captured cooperative launch, 2 captured plain launch. */
int mosp_batch_schedule(const mosp_batch *b, int32_t *group_size,
                        int32_t *graph_mode);

/* ------------------------------------------------------------ served ---- */
/* Served updates: graphs stay resident in HBM and change batches arrive as
 * requests.  No counterpart in the reference, which runs one mospUpdate per
 * process (src/mosp.cu:236-412); each request computes exactly what
 * mospUpdate computes on its graph's current state (apply + K updates +
 * combined graph + its search; src/mospUpdate.cu:57-225), so a result never
 * depends on the other requests.  Requests on one graph run in order; slots
 * (worker threads, each with a reusable batched workspace) run launches of
 * up to max_coalesce requests concurrently under an HBM budget. */
typedef struct mosp_server_config {
  int32_t slots;           /* concurrent launches (worker threads), >= 1 */
  int32_t max_coalesce;    /* requests (of distinct graphs) per launch, >= 1 */
  int32_t max_batch;       /* most changes per request (sizes the change arenas) */
  int32_t cu_policy;       /* 0: each slot has CUs/slots workgroups; 1: a launch takes the idle CUs */
  int32_t min_workgroups;  /* policy 1: smallest grant (0 = CUs / (2 slots)) */
  int32_t keep_results;    /* != 0: snapshot each request's trees (tests) */
  double delta_scale;      /* objective near-far widths = default rule x this (> 0) */
  int64_t hbm_budget;      /* bytes for resident state + slot scratch (0 = room for
                              max_coalesce updates per slot) */
} mosp_server_config;

typedef struct mosp_request_info {
  int32_t graph;
  int32_t slot;            /* slot that ran it (-1: not yet admitted) */
  int32_t coalesced;       /* requests in its launch */
  int32_t status;          /* 0 waiting or running, 1 done, -1 failed */
  int64_t changes;
  double arrival, admitted, done;  /* seconds on the server clock (-1: not yet) */
} mosp_request_info;

typedef struct mosp_server_counters {
  int32_t slots, cus;
  int64_t launches;
  double kernel_ms;        /* summed over slots (HIP events on their streams) */
  double apply_s, launch_s, post_s;  /* slot wall time: preparation + apply,
                                        batched launch, statistics */
  int64_t hbm_budget, resident_bytes, slot_bytes;
} mosp_server_counters;

typedef struct mosp_server mosp_server;

mosp_server *mosp_server_create(const mosp_server_config *config);
/* Make a graph resident (host CSR, edge-major weights, kg objectives) with
 * its k current trees (k*n objective-major, host) and preferences (0 or k
 * values); returns its id, or -1.  Before mosp_server_start. */
int32_t mosp_server_add_graph(mosp_server *s, int32_t n, int32_t kg, const int32_t *row_ptr,
                              const int32_t *col_ind, const int32_t *weights, int32_t k,
                              const int64_t *dist, const int32_t *parents,
                              const int32_t *pref, int32_t n_pref, int32_t source);
/* Split the budget, create the slots, start the workers and the clock. */
int mosp_server_start(mosp_server *s);
/* Queue a change batch (host arrays, copied; ins_w n_ins*kg insertion-major)
 * for graph `gid`; returns the request id, or -1. */
int64_t mosp_server_submit(mosp_server *s, int32_t gid, int32_t n_ins, const int32_t *ins_from,
                           const int32_t *ins_to, const int32_t *ins_w, int32_t n_del,
                           const int32_t *del_from, const int32_t *del_to);
int mosp_server_wait(mosp_server *s, int64_t request, double timeout_s);
/* Wait until every submitted request is done (0 on timeout or failure). */
int mosp_server_drain(mosp_server *s, double timeout_s);
int mosp_server_request(mosp_server *s, int64_t request, mosp_request_info *info);
/* The graph's K trees (k*n), combined distances and MOSP tree (n) right
 * after the request (keep_results servers). */
int mosp_server_request_trees(mosp_server *s, int64_t request, int64_t *dist, int32_t *parents,
                              int64_t *comb_dist, int32_t *comb_parent);
/* Forget a finished request's record. */
int mosp_server_release(mosp_server *s, int64_t request);
/* Totals since the start or the last reset; slot_kernel_ms (n_slots
 * entries, may be NULL) receives each slot's device time. */
int mosp_server_stats(mosp_server *s, mosp_server_counters *stats, double *slot_kernel_ms,
                      int32_t n_slots);
int mosp_server_reset_stats(mosp_server *s);
/* Seconds since mosp_server_start. */
double mosp_server_clock(const mosp_server *s);
/* Stop the workers (after the running launches) and free everything. */
void mosp_server_destroy(mosp_server *s);

/* ------------------------------------------------------ device graph ----- */

/* Copy a host CSR (edge-major weights w[e*K + k], the Values.txt layout) to
 * HBM once and build the objective-major weight columns and the reverse
 * CSR there.  Replaces uploadDeviceGraph (src/deviceGraph.cu:87-135). */
mosp_device_graph *mosp_device_graph_upload(int32_t n, int32_t k,
                                            const int32_t *row_ptr,
                                            const int32_t *col_ind,
                                            const int32_t *weights);
void mosp_device_graph_free(mosp_device_graph *g);
/* An empty graph, to be filled by mosp_device_graph_apply (no reference
 * counterpart). */
mosp_device_graph *mosp_device_graph_create(void);
/* Apply a change batch to g in HBM, writing the updated graph to `out` (a
 * different graph: empty, or a spare whose allocations are reused when large
 * enough).  Same result, row order and weight-increase mask as the host
 * applyChangeBatch (src/csrGraph.cu:727-894), which the reference runs on
 * the host before every upload (src/mospUpdate.cu:94-99); the batch arrays are
 * host memory (ins_w: n_ins * K, insertion-major); inc_mask (host, n_ins,
 * may be NULL) receives the bits.  Synchronous on `stream`. */
int mosp_device_graph_apply(const mosp_device_graph *g, int32_t n_ins,
                            const int32_t *ins_from, const int32_t *ins_to,
                            const int32_t *ins_w, int32_t n_del,
                            const int32_t *del_from, const int32_t *del_to,
                            mosp_device_graph *out, uint32_t *inc_mask,
                            void *stream);
/* Edge count of g. */
int64_t mosp_device_graph_edges(const mosp_device_graph *g);
/* Per-objective maximum and sum of g's weights (k entries each, either may
 * be NULL), computed on the device: what mospUpdate takes from the original
 * graph for the packing bound and the near-far width (src/mospUpdate.cu:
 * 104-122, 177-181).  Synchronous on `stream`. */
int mosp_device_graph_weight_stats(mosp_device_graph *g, int64_t *max_weight,
                                   int64_t *weight_sum, void *stream);
/* Copy g's out-CSR to host arrays (row_ptr n+1, col_ind m, weights m*K
 * edge-major: the Values.txt layout). */
int mosp_device_graph_download(const mosp_device_graph *g, int32_t *row_ptr,
                               int32_t *col_ind, int32_t *weights);
/* Views of objective `obj`.  Replaces DeviceGraph::out/in
 * (src/deviceGraph.cu:67-85). */
int mosp_device_graph_view(const mosp_device_graph *g, int32_t obj,
                           mosp_csr *out, mosp_csr *in);
int32_t mosp_device_graph_objectives(const mosp_device_graph *g);

/* MOSP path costs on the GPU: costs[v*K + k] = objective-k length of the
 * MOSP tree path to v (first (p,v) edge of row p), DISTANCE_INF if
 * unreachable; K = all objectives of the graph.  Fails if a tree edge is
 * missing.  Replaces mospPathCosts (src/combinedGraphGpu.cu:161-208), which
 * runs on the host in the reference. */
int mosp_path_costs_gpu(const mosp_device_graph *g, const int32_t *d_parent,
                        int32_t source, int64_t *d_costs, void *stream);

/* --------------------------------------------------------- host helpers --- */

/* Replaces preferenceScale (src/combinedGraphGpu.cu:129-147). */
int64_t mosp_preference_scale(const int32_t *pref, int32_t n_pref, int32_t k);
/* Replaces combinedEdgeWeight (src/combinedGraphGpu.cu:149-159). */
int64_t mosp_combined_edge_weight(uint32_t tree_mask, const int32_t *pref,
                                  int32_t n_pref, int32_t k, int64_t scale);

/* Packing chosen for n vertices and max weight (src/sospUpdateGpu.cu:127-143,
 * 670-681): returns 0 if (n-1)*max_weight exceeds 62 bits, else 1 with
 * *parent_bits (0 = distance-only) and *bound = (n-1)*max_weight. */
int mosp_choose_packing(int32_t n, int64_t max_weight, int32_t *parent_bits,
                        uint64_t *bound);

/* Apply a change batch on the host (src/csrGraph.cu:727-894).  Output arrays
 * need capacity m + n_ins (col) and (m + n_ins)*K (w); inc_mask[n_ins].
 * Returns the updated edge count, or -1 on error. */
int64_t mosp_apply_batch(int32_t n, int32_t k, const int32_t *row_ptr,
                         const int32_t *col_ind, const int32_t *weights,
                         int32_t n_ins, const int32_t *ins_from,
                         const int32_t *ins_to, const int32_t *ins_w,
                         int32_t n_del, const int32_t *del_from,
                         const int32_t *del_to, int32_t *out_row_ptr,
                         int32_t *out_col_ind, int32_t *out_weights,
                         uint32_t *inc_mask);

/* Lowest-id tie rule on host arrays: parent[v] becomes the smallest u over
 * the tight edges u -> v of objective `objective` (d[u] + w = d[v]) and its
 * current parent, for every vertex with a parent (v != source).  Replaces
 * canonicalizeTree (src/mospUpdate.cu:40-55). */
void mosp_canonicalize_tree(int32_t n, int32_t k, const int32_t *row_ptr,
                            const int32_t *col_ind, const int32_t *weights,
                            int32_t objective, int32_t source, const int64_t *dist,
                            int32_t *parent);

/* MOSP path costs on the host (costs[v*K + k], as mosp_path_costs_gpu):
 * 1, or 0 if a tree edge is missing.  Replaces the reference's host
 * mospPathCosts (src/combinedGraphGpu.cu:161-208). */
int mosp_path_costs_host(int32_t n, int32_t k, const int32_t *row_ptr,
                         const int32_t *col_ind, const int32_t *weights,
                         const int32_t *parent, int32_t source, int64_t *costs);

/* The in-memory MOSP update on host arrays (src/mospUpdate.cu:57-225):
 * apply, upload, K updates + combined step on the GPU, download.
 * Inputs: original graph (edge-major KG weights), batch, K*n initial trees
 * (objective-major), options.  Outputs (caller-allocated): updated graph
 * (capacity m + n_ins), K*n trees, combined dist/parent (n).
 * fused != 0 runs the K updates and the combined step in one launch;
 * fused == 0 runs K separate launches + one combined launch like the
 * reference.  timings_ms[5] = apply, upload, sosp (all K), combined,
 * download; gpu_compute = timings_ms[2] + timings_ms[3].
 * stats[k] per objective; cstats for the combined step. */
int mosp_update_host(int32_t n, int32_t kg, const int32_t *row_ptr,
                     const int32_t *col_ind, const int32_t *weights,
                     int32_t n_ins, const int32_t *ins_from,
                     const int32_t *ins_to, const int32_t *ins_w, int32_t n_del,
                     const int32_t *del_from, const int32_t *del_to,
                     int32_t k_use, int32_t source, const int32_t *pref,
                     int32_t n_pref, int64_t delta, int32_t fused,
                     const int64_t *init_dist, const int32_t *init_parents,
                     int64_t *out_m, int32_t *upd_row_ptr, int32_t *upd_col_ind,
                     int32_t *upd_weights, int64_t *dist, int32_t *parents,
                     int64_t *comb_dist, int32_t *comb_parent,
                     mosp_sosp_stats *stats, mosp_combine_stats *cstats,
                     double *timings_ms);

/* ------------------------------------------------- file-based wrappers --- */
/* parallelSOSPUpdate (headers/parallelSOSPUpdate.cuh:32-39): reads the CSR
 * files <csr_prefix>{RowPtr,ColInd,Values}.txt, objective `objective`'s old
 * tree (distances + parent files), insert.txt / delete.txt; applies the
 * batch, updates the tree on the GPU and writes it (tree file format).
 * 1 on success; 0 after printing the reference's error message. */
int mosp_parallel_sosp_update(const char *csr_prefix, const char *distances_in,
                              const char *tree_in, const char *insert_path,
                              const char *delete_path, int32_t objective, int32_t source,
                              const char *distances_out, const char *tree_out);

/* parallelCombinedGraph (headers/parallelCombinedGraph.cuh:46-53): the
 * vertex count from <csr_prefix>RowPtr.txt, the K parent files tree_paths[k];
 * the combined graph and its SOSP tree on the GPU, written in the tree file
 * format (distances in units of 1/L).  pref: n_pref values (0: all ones).
 * 1 on success; 0 after printing the reference's error message. */
int mosp_parallel_combined_graph(const char *csr_prefix, const char *const *tree_paths,
                                 int32_t k, int32_t source, const char *distances_out,
                                 const char *tree_out, const int32_t *pref, int32_t n_pref);

/* ----------------------------------------------------- synthetic inputs --- */
/* Seeded generators (xoshiro256** + unbiased bounded integers): identical
 * instances on every host.  No reference counterpart (the reference's own
 * generators live in src/changeGenerator.cu and src/mospTest.cu:64-106). */

/* Road-like W x H grid: random spanning tree kept, `drop` of the grid edges
 * removed, `diag` * n diagonal shortcuts added; both directions stored.
 * Call with row_ptr == NULL to get *n / *m, then with buffers of capacity
 * >= *m edges (weights edge-major, K per edge). */
int mosp_gen_road_grid(int32_t width, int32_t height, int32_t k, int32_t wmin,
                       int32_t wmax, double drop, double diag, uint64_t seed,
                       int32_t *n, int64_t *m, int32_t *row_ptr,
                       int32_t *col_ind, int32_t *weights, int64_t capacity);

/* R-MAT (a, b, c; d = 1-a-b-c), 2^scale vertices, edge_factor * 2^scale
 * draws, self-loops and duplicates removed.  Two-call protocol as above. */
int mosp_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                  double c, int32_t k, int32_t wmin, int32_t wmax, uint64_t seed,
                  int32_t *n, int64_t *m, int32_t *row_ptr, int32_t *col_ind,
                  int32_t *weights, int64_t capacity);

/* Change batch: round(changes * ins_pct / 100) insertions of random pairs
 * u != v (weights in [wmin, wmax]) and the rest deletions of distinct
 * existing arcs; safe != 0 keeps every vertex reachable from `source`
 * (src/changeGenerator.cu:78-116 rule).  Buffers: ins_from/ins_to n_ins,
 * ins_w n_ins * K, del_* n_del (capacity = changes). */
int mosp_gen_batch(int32_t n, int32_t k, const int32_t *row_ptr,
                   const int32_t *col_ind, const int32_t *weights,
                   int32_t changes, double ins_pct, int32_t safe, int32_t wmin,
                   int32_t wmax, uint64_t seed, int32_t source, int32_t *n_ins,
                   int32_t *ins_from, int32_t *ins_to, int32_t *ins_w,
                   int32_t *n_del, int32_t *del_from, int32_t *del_to);

#ifdef __cplusplus
}
#endif

#endif /* MOSP_HIP_H */
