/*
 * mosp_oracle.h -- CPU restatement of the reference's DynaMOSP update path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  Nothing in the product (mosp-cuda_amd/) links or calls it.
 *
 * Graph layout is the reference's CsrGraph (headers/csrGraph.cuh:18-29):
 * rowPtr[n+1], colInd[m], edge-major weights w[e*K + k].  Distances are
 * int64 with OR_INF = INT64_MAX/4 for "unreachable"; parents are int32 with
 * -1 for "none".
 *
 * Parity pins (each function is checked against at least one of these in
 * tests/test_oracle.py):
 *   or_dijkstra            brute-force Bellman-Ford with the lowest-id rule
 *                          (tiny graphs), thesis known-answer trees, _ref
 *   or_apply_batch         pure-Python list restatement, _ref applyChangeBatch
 *   or_sequential_update   == or_dijkstra on the updated graph (canonical
 *                          fixed point), _ref sequentialSOSPUpdate files
 *   or_combined_graph      thesis known-answer MOSP paths, closed form
 *   or_path_costs          thesis known-answer costs (15,3,20) / (15,24,7)
 *   or_choose_packing      closed-form cases of the packing-boundary test
 *   or_default_delta       closed form
 *   or_alg_bytes           hand count on a 4-vertex graph
 */
#ifndef MOSP_ORACLE_H
#define MOSP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_INF (INT64_MAX / 4)

/* Dijkstra.cu:239-272 (dijkstraCsrGraph). */
void or_dijkstra(int n, const int *rowPtr, const int *colInd, const int *w,
                 int K, int objective, int source, int64_t *dist, int *parent);

/* csrGraph.cu:727-894 (applyChangeBatch).  Output arrays need capacity
 * m + nIns (colInd) and (m + nIns) * K (w).  incMask[nIns] receives the
 * weight-increase mask.  Returns the updated m, or -1 on error (K > 32). */
int or_apply_batch(int n, int K, const int *rowPtr, const int *colInd,
                   const int *w, int nIns, const int *insFrom,
                   const int *insTo, const int *insW, int nDel,
                   const int *delFrom, const int *delTo, int *outRowPtr,
                   int *outColInd, int *outW, unsigned *incMask);

/* csrGraph.cu:896-922 (transposeCsrGraph). */
void or_transpose(int n, int K, const int *rowPtr, const int *colInd,
                  const int *w, int *tRowPtr, int *tColInd, int *tW);

/* sequentialSOSPUpdate.cu:457-612 (Phases 1-2) on an already-updated graph
 * given as out-CSR and in-CSR (edge-major weights, objective index).
 * changed = deletions + weight increases of this objective, heads = insert
 * heads in batch order.  dist/parent hold the old tree on entry and the new
 * tree on return.  Returns the Phase-2 iteration count, or -1 if it fails to
 * converge (cannot happen for positive weights). */
int or_sequential_update(int n, const int *outRow, const int *outCol,
                         const int *outW, const int *inRow, const int *inCol,
                         const int *inW, int K, int objective, int source,
                         int nChanged, const int *chFrom, const int *chTo,
                         int nHeads, const int *heads, int64_t *dist,
                         int *parent);

/* combinedGraphGpu.cu:129-147 (preferenceScale): L = lcm(pref) or 0. */
int64_t or_preference_scale(const int *pref, int nPref, int K);

/* combinedGraphGpu.cu:149-159 (combinedEdgeWeight). */
int64_t or_combined_edge_weight(unsigned mask, const int *pref, int nPref,
                                int K, int64_t scale);

/* validation.cu:88-125 (combinedGraphReference).  colInd/weights need
 * capacity K*n.  Returns the edge count, or -1 for an invalid Pref. */
int or_combined_graph(const int *parents, int n, int K, int source,
                      const int *pref, int nPref, int *rowPtr, int *colInd,
                      int *weights);

/* combinedGraphGpu.cu:161-208 (mospPathCosts).  costs[n*K].  Returns 1, or
 * 0 if a tree edge is missing from the graph. */
int or_path_costs(int n, int K, const int *rowPtr, const int *colInd,
                  const int *w, const int *parent, int source,
                  int64_t *costs);

/* mospUpdate.cu:40-55 (canonicalizeTree). */
void or_canonicalize_tree(int n, const int *rowPtr, const int *colInd,
                          const int *w, int K, int objective, int source,
                          const int64_t *dist, int *parent);

/* validation.cu:28-86 (checkSospTree): out[4] = distance mismatches,
 * inconsistent parents, non-canonical parents, parent mismatches. */
void or_check_tree(int n, const int *revRow, const int *revCol,
                   const int *revW, int K, int objective, int source,
                   const int64_t *dist, const int *parent,
                   const int64_t *refDist, const int *refParent,
                   int64_t *out);

/* sospUpdateGpu.cu:607-616 (defaultDelta). */
int64_t or_default_delta(int64_t m, int n, int64_t weightSum);

/* sospUpdateGpu.cu:127-143, 670-681 (makePacking / choosePacking).
 * Returns 0 if (n-1)*maxW does not fit 62 bits; else 1 with parentBits
 * (0 = distance-only fallback) and the distance bound. */
int or_choose_packing(int n, int64_t maxWeight, int *parentBits,
                      uint64_t *bound);

/* Whole in-memory pipeline, mospUpdate.cu:57-225 with the sequential update
 * as engine (the CPU path of the reference) and host combined graph +
 * Dijkstra for Steps 2-3 (validation.cu:88-125, Dijkstra.cu:239-272).
 * Kuse <= KG objectives of the graph are updated.  outColInd/outW need
 * capacity m + nIns.  dist/parents: Kuse*n in (old trees) and out.
 * times_ms[3] = apply, K sequential updates, combined graph + Dijkstra.
 * info[4] = updated m, combined edges, L, status (1 ok, 0 error). */
void or_mosp_update(int n, int KG, const int *rowPtr, const int *colInd,
                    const int *w, int nIns, const int *insFrom,
                    const int *insTo, const int *insW, int nDel,
                    const int *delFrom, const int *delTo, int Kuse,
                    int source, const int *pref, int nPref, int64_t *dist,
                    int *parents, int *outRowPtr, int *outColInd, int *outW,
                    int64_t *combDist, int *combParent, double *times_ms,
                    int64_t *info);

/* Algorithmic HBM bytes of one update (the measurement formula): per
 * objective k, 28n + sum over C_k of (8 + 16 indeg) + sum over D_k of
 * (8 + 16 outdeg); combined 4Kn + 4(n+1) + 8E + 8n + 16E + 12n.
 * C_k = I_k u (H \ {s}); D_k = {u != s : d'(u) < INF and (u in I_k or
 * d'(u) != d(u))}.  oldParents/oldDist/newDist are Kuse*n objective-major;
 * changed lists per objective are passed flattened with offsets. */
int64_t or_alg_bytes(int n, const int *outRow, const int *inRow, int Kuse,
                     int source, const int *oldParents, const int64_t *oldDist,
                     const int64_t *newDist, const int *chOffsets,
                     const int *chFrom, const int *chTo, int nHeads,
                     const int *heads, int64_t combinedEdges);

#ifdef __cplusplus
}
#endif

#endif /* MOSP_ORACLE_H */
