/*
 * mosp_oracle.c -- CPU restatement of the reference's DynaMOSP update path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; see the header).
 * Every function cites the reference file:line whose behaviour it restates.
 * Written in plain C from the reference's documented semantics; the data
 * structures are this file's own (flat arrays, generation stamps instead of
 * per-iteration flag vectors).
 */
#define _POSIX_C_SOURCE 199309L
#include "mosp_oracle.h"

#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------------ */
/* small helpers                                                            */
/* ------------------------------------------------------------------------ */

static void *xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) abort();
  return p;
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* Binary min-heap of (distance, vertex), lexicographic like std::pair. */
typedef struct {
  int64_t d;
  int v;
} HeapItem;

typedef struct {
  HeapItem *a;
  size_t size, cap;
} Heap;

static int item_less(HeapItem x, HeapItem y) {
  return x.d < y.d || (x.d == y.d && x.v < y.v);
}

static void heap_push(Heap *h, int64_t d, int v) {
  if (h->size == h->cap) {
    h->cap = h->cap ? 2 * h->cap : 1024;
    h->a = (HeapItem *)realloc(h->a, h->cap * sizeof(HeapItem));
    if (!h->a) abort();
  }
  size_t i = h->size++;
  HeapItem it = {d, v};
  while (i > 0) {
    size_t p = (i - 1) / 2;
    if (!item_less(it, h->a[p])) break;
    h->a[i] = h->a[p];
    i = p;
  }
  h->a[i] = it;
}

static HeapItem heap_pop(Heap *h) {
  HeapItem top = h->a[0];
  HeapItem last = h->a[--h->size];
  size_t i = 0;
  for (;;) {
    size_t l = 2 * i + 1, r = l + 1, s = i;
    HeapItem best = last;
    if (l < h->size && item_less(h->a[l], best)) {
      s = l;
      best = h->a[l];
    }
    if (r < h->size && item_less(h->a[r], best)) s = r;
    if (s == i) break;
    h->a[i] = h->a[s];
    i = s;
  }
  if (h->size > 0) h->a[i] = last;
  return top;
}

/* ------------------------------------------------------------------------ */
/* Dijkstra.cu:239-272                                                      */
/* ------------------------------------------------------------------------ */

void or_dijkstra(int n, const int *rowPtr, const int *colInd, const int *w,
                 int K, int objective, int source, int64_t *dist,
                 int *parent) {
  for (int v = 0; v < n; ++v) {
    dist[v] = OR_INF;
    parent[v] = -1;
  }
  if (source < 0 || source >= n) return;
  Heap h = {0, 0, 0};
  dist[source] = 0;
  heap_push(&h, 0, source);
  while (h.size > 0) {
    HeapItem top = heap_pop(&h);
    int u = top.v;
    if (top.d != dist[u]) continue; /* stale entry */
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      int v = colInd[e];
      int64_t cand = top.d + (int64_t)w[(size_t)e * K + objective];
      if (cand < dist[v]) {
        dist[v] = cand;
        parent[v] = u;
        heap_push(&h, cand, v);
      } else if (cand == dist[v] && v != source && u < parent[v]) {
        parent[v] = u; /* Dijkstra.cu:266-268: lowest parent id wins ties */
      }
    }
  }
  free(h.a);
}

/* ------------------------------------------------------------------------ */
/* csrGraph.cu:727-894                                                      */
/* ------------------------------------------------------------------------ */

typedef struct {
  int vertex;
  long long slot; /* >= 0: original edge; < 0: insertion -slot-1 */
  int alive;
} RowEntry;

int or_apply_batch(int n, int K, const int *rowPtr, const int *colInd,
                   const int *w, int nIns, const int *insFrom,
                   const int *insTo, const int *insW, int nDel,
                   const int *delFrom, const int *delTo, int *outRowPtr,
                   int *outColInd, int *outW, unsigned *incMask) {
  if (K > 32) return -1; /* csrGraph.cu:736-739 */
  /* Group changes by source row, keeping batch order within a row
   * (csrGraph.cu:743-765). */
  int *delStart = (int *)calloc((size_t)n + 1, sizeof(int));
  int *insStart = (int *)calloc((size_t)n + 1, sizeof(int));
  if (!delStart || !insStart) abort();
  for (int i = 0; i < nDel; ++i) ++delStart[delFrom[i] + 1];
  for (int i = 0; i < nIns; ++i) ++insStart[insFrom[i] + 1];
  for (int u = 0; u < n; ++u) {
    delStart[u + 1] += delStart[u];
    insStart[u + 1] += insStart[u];
  }
  int *delOrder = (int *)xmalloc(sizeof(int) * (size_t)(nDel ? nDel : 1));
  int *insOrder = (int *)xmalloc(sizeof(int) * (size_t)(nIns ? nIns : 1));
  int *cursor = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  memcpy(cursor, delStart, sizeof(int) * (size_t)n);
  for (int i = 0; i < nDel; ++i) delOrder[cursor[delFrom[i]]++] = i;
  memcpy(cursor, insStart, sizeof(int) * (size_t)n);
  for (int i = 0; i < nIns; ++i) insOrder[cursor[insFrom[i]]++] = i;

  /* Rebuild each changed row (csrGraph.cu:780-815) and assemble the updated
   * CSR row by row (csrGraph.cu:817-859). */
  int maxRow = 0;
  for (int u = 0; u < n; ++u) {
    int len = rowPtr[u + 1] - rowPtr[u] + insStart[u + 1] - insStart[u];
    if (len > maxRow) maxRow = len;
  }
  RowEntry *row = (RowEntry *)xmalloc(sizeof(RowEntry) * (size_t)(maxRow + 1));
  int out = 0;
  for (int u = 0; u < n; ++u) {
    outRowPtr[u] = out;
    if (delStart[u] == delStart[u + 1] && insStart[u] == insStart[u + 1]) {
      for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e, ++out) {
        outColInd[out] = colInd[e];
        memcpy(&outW[(size_t)out * K], &w[(size_t)e * K], sizeof(int) * K);
      }
      continue;
    }
    int len = 0;
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      row[len].vertex = colInd[e];
      row[len].slot = e;
      row[len].alive = 1;
      ++len;
    }
    /* Deletions: the first alive (u,v) entry dies (csrGraph.cu:788-796). */
    for (int j = delStart[u]; j < delStart[u + 1]; ++j) {
      int v = delTo[delOrder[j]];
      for (int x = 0; x < len; ++x) {
        if (row[x].alive && row[x].vertex == v) {
          row[x].alive = 0;
          break;
        }
      }
    }
    /* Insertions: overwrite the first alive (u,v) or append
     * (csrGraph.cu:797-811). */
    for (int j = insStart[u]; j < insStart[u + 1]; ++j) {
      int i = insOrder[j];
      int v = insTo[i];
      int replaced = 0;
      for (int x = 0; x < len; ++x) {
        if (row[x].alive && row[x].vertex == v) {
          row[x].slot = -(long long)i - 1;
          replaced = 1;
          break;
        }
      }
      if (!replaced) {
        row[len].vertex = v;
        row[len].slot = -(long long)i - 1;
        row[len].alive = 1;
        ++len;
      }
    }
    for (int x = 0; x < len; ++x) {
      if (!row[x].alive) continue;
      outColInd[out] = row[x].vertex;
      const int *src = row[x].slot >= 0
                           ? &w[(size_t)row[x].slot * K]
                           : &insW[(size_t)(-row[x].slot - 1) * K];
      memcpy(&outW[(size_t)out * K], src, sizeof(int) * K);
      ++out;
    }
  }
  outRowPtr[n] = out;

  /* Weight increases: final first-occurrence weight vs original
   * first-occurrence weight of (u,v) (csrGraph.cu:861-892). */
  for (int i = 0; i < nIns; ++i) {
    int u = insFrom[i], v = insTo[i];
    int origEdge = -1, finalEdge = -1;
    incMask[i] = 0;
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      if (colInd[e] == v) {
        origEdge = e;
        break;
      }
    }
    if (origEdge < 0) continue;
    for (int e = outRowPtr[u]; e < outRowPtr[u + 1]; ++e) {
      if (outColInd[e] == v) {
        finalEdge = e;
        break;
      }
    }
    if (finalEdge < 0) continue;
    unsigned mask = 0;
    for (int k = 0; k < K; ++k) {
      if (outW[(size_t)finalEdge * K + k] > w[(size_t)origEdge * K + k])
        mask |= 1u << k;
    }
    incMask[i] = mask;
  }
  free(row);
  free(cursor);
  free(insOrder);
  free(delOrder);
  free(insStart);
  free(delStart);
  return out;
}

/* ------------------------------------------------------------------------ */
/* csrGraph.cu:896-922                                                      */
/* ------------------------------------------------------------------------ */

void or_transpose(int n, int K, const int *rowPtr, const int *colInd,
                  const int *w, int *tRowPtr, int *tColInd, int *tW) {
  int m = rowPtr[n];
  memset(tRowPtr, 0, sizeof(int) * ((size_t)n + 1));
  for (int e = 0; e < m; ++e) ++tRowPtr[colInd[e] + 1];
  for (int v = 0; v < n; ++v) tRowPtr[v + 1] += tRowPtr[v];
  int *cursor = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  memcpy(cursor, tRowPtr, sizeof(int) * (size_t)n);
  for (int u = 0; u < n; ++u) {
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      int pos = cursor[colInd[e]]++;
      tColInd[pos] = u;
      memcpy(&tW[(size_t)pos * K], &w[(size_t)e * K], sizeof(int) * K);
    }
  }
  free(cursor);
}

/* ------------------------------------------------------------------------ */
/* sequentialSOSPUpdate.cu:256-284 (findBestParent), 457-612 (Phases 1-2)   */
/* ------------------------------------------------------------------------ */

typedef struct {
  int n, K, objective;
  const int *inRow, *inCol, *inW;
  int64_t *dist;
  int *parent;
} SeqState;

/* findBestParent + relax (sequentialSOSPUpdate.cu:256-284, 524-539).
 * Returns 1 if the distance of v decreased. */
static int seq_relax(SeqState *s, int v) {
  int bestP = -1;
  int64_t bestD = OR_INF;
  for (int e = s->inRow[v]; e < s->inRow[v + 1]; ++e) {
    int u = s->inCol[e];
    if (s->dist[u] >= OR_INF / 2) continue;
    int64_t cand = s->dist[u] + (int64_t)s->inW[(size_t)e * s->K + s->objective];
    if (cand < bestD || (cand == bestD && u < bestP)) {
      bestD = cand;
      bestP = u;
    }
  }
  int better = bestD < s->dist[v] ||
               (bestD == s->dist[v] && bestP >= 0 && bestP < s->parent[v]);
  if (!better) return 0;
  int decreased = bestD < s->dist[v];
  s->dist[v] = bestD;
  s->parent[v] = bestP;
  return decreased;
}

int or_sequential_update(int n, const int *outRow, const int *outCol,
                         const int *outW, const int *inRow, const int *inCol,
                         const int *inW, int K, int objective, int source,
                         int nChanged, const int *chFrom, const int *chTo,
                         int nHeads, const int *heads, int64_t *dist,
                         int *parent) {
  (void)outW;
  SeqState s = {n, K, objective, inRow, inCol, inW, dist, parent};
  char *isInvalid = (char *)calloc((size_t)n + 1, 1);
  int *invalidated = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int nInv = 0;
  /* 1a: roots = heads of changed edges that are tree edges
   * (sequentialSOSPUpdate.cu:470-483). */
  for (int i = 0; i < nChanged; ++i) {
    int u = chFrom[i], v = chTo[i];
    if (parent[v] == u && !isInvalid[v]) {
      isInvalid[v] = 1;
      invalidated[nInv++] = v;
    }
  }
  /* 1b: invalidate the subtrees of the roots through the children lists of
   * the old tree (sequentialSOSPUpdate.cu:485-518). */
  {
    int *childStart = (int *)calloc((size_t)n + 1, sizeof(int));
    int *children = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
    if (!childStart) abort();
    for (int v = 0; v < n; ++v)
      if (v != source && parent[v] >= 0) ++childStart[parent[v] + 1];
    for (int v = 0; v < n; ++v) childStart[v + 1] += childStart[v];
    int *cursor = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
    memcpy(cursor, childStart, sizeof(int) * (size_t)n);
    for (int v = 0; v < n; ++v)
      if (v != source && parent[v] >= 0) children[cursor[parent[v]]++] = v;
    for (int i = 0; i < nInv; ++i) {
      int x = invalidated[i];
      for (int c = childStart[x]; c < childStart[x + 1]; ++c) {
        int y = children[c];
        if (!isInvalid[y]) {
          isInvalid[y] = 1;
          invalidated[nInv++] = y;
        }
      }
    }
    for (int i = 0; i < nInv; ++i) {
      dist[invalidated[i]] = OR_INF;
      parent[invalidated[i]] = -1;
    }
    free(cursor);
    free(children);
    free(childStart);
  }
  /* 1c: re-evaluate invalidated vertices, then insert heads
   * (sequentialSOSPUpdate.cu:520-557). */
  char *isAffected = (char *)calloc((size_t)n + 1, 1);
  int *affected = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int nAff = 0;
  for (int i = 0; i < nInv; ++i) {
    int v = invalidated[i];
    if (seq_relax(&s, v) && !isAffected[v]) {
      isAffected[v] = 1;
      affected[nAff++] = v;
    }
  }
  for (int i = 0; i < nHeads; ++i) {
    int v = heads[i];
    if (v != source && seq_relax(&s, v) && !isAffected[v]) {
      isAffected[v] = 1;
      affected[nAff++] = v;
    }
  }
  /* Phase 2: sweep the out-neighbours of the affected vertices until no
   * distance decreases (sequentialSOSPUpdate.cu:571-612).  Candidate
   * de-duplication uses a generation stamp per vertex. */
  int *candStamp = (int *)calloc((size_t)n + 1, sizeof(int));
  int *cands = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int iterations = 0;
  while (nAff > 0) {
    ++iterations;
    if (iterations > n) {
      iterations = -1;
      break;
    }
    int nC = 0;
    for (int i = 0; i < nAff; ++i) {
      int a = affected[i];
      isAffected[a] = 0;
      for (int e = outRow[a]; e < outRow[a + 1]; ++e) {
        int v = outCol[e];
        if (v == source) continue;
        if (candStamp[v] != iterations) {
          candStamp[v] = iterations;
          cands[nC++] = v;
        }
      }
    }
    nAff = 0;
    for (int i = 0; i < nC; ++i) {
      int v = cands[i];
      if (seq_relax(&s, v) && !isAffected[v]) {
        isAffected[v] = 1;
        affected[nAff++] = v;
      }
    }
  }
  free(cands);
  free(candStamp);
  free(affected);
  free(isAffected);
  free(invalidated);
  free(isInvalid);
  return iterations;
}

/* ------------------------------------------------------------------------ */
/* combinedGraphGpu.cu:129-159, validation.cu:88-125                         */
/* ------------------------------------------------------------------------ */

static int64_t gcd64(int64_t a, int64_t b) {
  while (b != 0) {
    int64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int64_t or_preference_scale(const int *pref, int nPref, int K) {
  if (nPref == 0) return 1;
  if (nPref != K) return 0;
  int64_t scale = 1;
  for (int i = 0; i < nPref; ++i) {
    if (pref[i] < 1) return 0;
    scale = scale / gcd64(scale, pref[i]) * pref[i];
    if (scale > ((int64_t)1 << 20)) return 0;
  }
  return scale;
}

int64_t or_combined_edge_weight(unsigned mask, const int *pref, int nPref,
                                int K, int64_t scale) {
  int64_t weight = scale * (K + 1);
  for (int i = 0; i < K; ++i)
    if (mask & (1u << i)) weight -= nPref == 0 ? scale : scale / pref[i];
  return weight;
}

int or_combined_graph(const int *parents, int n, int K, int source,
                      const int *pref, int nPref, int *rowPtr, int *colInd,
                      int *weights) {
  int64_t scale = or_preference_scale(pref, nPref, K);
  if (scale == 0 || K < 1 || K > 32) return -1;
  /* Edges are grouped by parent row, in increasing v, then k order (the
   * row order of validation.cu:94-123). */
  memset(rowPtr, 0, sizeof(int) * ((size_t)n + 1));
  for (int pass = 0; pass < 2; ++pass) {
    int *cursor = NULL;
    if (pass == 1) {
      for (int u = 0; u < n; ++u) rowPtr[u + 1] += rowPtr[u];
      cursor = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
      memcpy(cursor, rowPtr, sizeof(int) * (size_t)n);
    }
    for (int v = 0; v < n; ++v) {
      if (v == source) continue;
      for (int k = 0; k < K; ++k) {
        int p = parents[(size_t)k * n + v];
        int seen = p < 0;
        for (int j = 0; j < k && !seen; ++j)
          seen = parents[(size_t)j * n + v] == p;
        if (seen) continue;
        if (pass == 0) {
          ++rowPtr[p + 1];
          continue;
        }
        unsigned mask = 0;
        for (int j = 0; j < K; ++j)
          if (parents[(size_t)j * n + v] == p) mask |= 1u << j;
        int pos = cursor[p]++;
        colInd[pos] = v;
        weights[pos] = (int)or_combined_edge_weight(mask, pref, nPref, K, scale);
      }
    }
    free(cursor);
  }
  return rowPtr[n];
}

/* ------------------------------------------------------------------------ */
/* combinedGraphGpu.cu:161-208                                              */
/* ------------------------------------------------------------------------ */

int or_path_costs(int n, int K, const int *rowPtr, const int *colInd,
                  const int *w, const int *parent, int source,
                  int64_t *costs) {
  for (size_t i = 0; i < (size_t)n * K; ++i) costs[i] = OR_INF;
  int *childStart = (int *)calloc((size_t)n + 1, sizeof(int));
  int *children = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int *queue = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  if (!childStart) abort();
  for (int v = 0; v < n; ++v)
    if (v != source && parent[v] >= 0) ++childStart[parent[v] + 1];
  for (int v = 0; v < n; ++v) childStart[v + 1] += childStart[v];
  int *cursor = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  memcpy(cursor, childStart, sizeof(int) * (size_t)n);
  for (int v = 0; v < n; ++v)
    if (v != source && parent[v] >= 0) children[cursor[parent[v]]++] = v;
  for (int k = 0; k < K; ++k) costs[(size_t)source * K + k] = 0;
  int head = 0, tail = 0, ok = 1;
  queue[tail++] = source;
  while (head < tail && ok) {
    int p = queue[head++];
    for (int c = childStart[p]; c < childStart[p + 1]; ++c) {
      int v = children[c];
      int edge = -1;
      for (int e = rowPtr[p]; e < rowPtr[p + 1]; ++e) {
        if (colInd[e] == v) {
          edge = e; /* first occurrence of (p, v) */
          break;
        }
      }
      if (edge < 0) {
        ok = 0;
        break;
      }
      for (int k = 0; k < K; ++k)
        costs[(size_t)v * K + k] =
            costs[(size_t)p * K + k] + (int64_t)w[(size_t)edge * K + k];
      queue[tail++] = v;
    }
  }
  free(cursor);
  free(queue);
  free(children);
  free(childStart);
  return ok;
}

/* ------------------------------------------------------------------------ */
/* mospUpdate.cu:40-55, validation.cu:28-86                                 */
/* ------------------------------------------------------------------------ */

void or_canonicalize_tree(int n, const int *rowPtr, const int *colInd,
                          const int *w, int K, int objective, int source,
                          const int64_t *dist, int *parent) {
  for (int u = 0; u < n; ++u) {
    int64_t du = dist[u];
    if (du >= OR_INF / 2) continue;
    for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
      int v = colInd[e];
      if (v != source && u < parent[v] &&
          du + (int64_t)w[(size_t)e * K + objective] == dist[v])
        parent[v] = u;
    }
  }
}

void or_check_tree(int n, const int *revRow, const int *revCol,
                   const int *revW, int K, int objective, int source,
                   const int64_t *dist, const int *parent,
                   const int64_t *refDist, const int *refParent,
                   int64_t *out) {
  out[0] = out[1] = out[2] = out[3] = 0;
  for (int v = 0; v < n; ++v) {
    int64_t dv = dist[v];
    int reach = dv < OR_INF / 2;
    if (reach != (refDist[v] < OR_INF / 2) || (reach && dv != refDist[v]))
      ++out[0];
    if (refParent && parent[v] != refParent[v]) ++out[3];
    if (v == source) {
      if (parent[v] != -1 || dv != 0) ++out[1];
      continue;
    }
    if (!reach) {
      if (parent[v] != -1) ++out[1];
      continue;
    }
    int p = parent[v], consistent = 0, lowest = -1;
    for (int e = revRow[v]; e < revRow[v + 1]; ++e) {
      int u = revCol[e];
      if (dist[u] >= OR_INF / 2) continue;
      if (dist[u] + (int64_t)revW[(size_t)e * K + objective] != dv) continue;
      if (u == p) consistent = 1;
      if (lowest < 0 || u < lowest) lowest = u;
    }
    if (!consistent)
      ++out[1];
    else if (p != lowest)
      ++out[2];
  }
}

/* ------------------------------------------------------------------------ */
/* sospUpdateGpu.cu:127-143, 607-616, 670-681                               */
/* ------------------------------------------------------------------------ */

int64_t or_default_delta(int64_t m, int n, int64_t weightSum) {
  if (m <= 0 || n <= 0) return 1;
  double avgW = (double)weightSum / (double)m;
  double avgDeg = (double)m / (double)n;
  int64_t d = (int64_t)(32.0 * avgW / avgDeg);
  return d > 1 ? d : 1;
}

int or_choose_packing(int n, int64_t maxWeight, int *parentBits,
                      uint64_t *bound) {
  const uint64_t outputMax = (uint64_t)(OR_INF / 2 - 1);
  uint64_t weight = (uint64_t)(maxWeight > 1 ? maxWeight : 1);
  uint64_t hops = (uint64_t)(n - 1 > 1 ? n - 1 : 1);
  if (weight > outputMax / hops) return 0;
  *bound = weight * hops;
  int bits = 1;
  while ((((uint64_t)1) << bits) - 1 < (uint64_t)n) ++bits;
  uint64_t maxDistance = (~(uint64_t)0 >> bits) - 1;
  *parentBits = *bound <= maxDistance ? bits : 0;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* mospUpdate.cu:57-225 with the CPU engines                                */
/* ------------------------------------------------------------------------ */

void or_mosp_update(int n, int KG, const int *rowPtr, const int *colInd,
                    const int *w, int nIns, const int *insFrom,
                    const int *insTo, const int *insW, int nDel,
                    const int *delFrom, const int *delTo, int Kuse,
                    int source, const int *pref, int nPref, int64_t *dist,
                    int *parents, int *outRowPtr, int *outColInd, int *outW,
                    int64_t *combDist, int *combParent, double *times_ms,
                    int64_t *info) {
  info[0] = info[1] = info[2] = info[3] = 0;
  times_ms[0] = times_ms[1] = times_ms[2] = 0;
  if (n <= 0 || Kuse <= 0 || Kuse > KG || source < 0 || source >= n) return;
  /* Initial trees must be rooted at the source (mospUpdate.cu:77-92). */
  for (int k = 0; k < Kuse; ++k) {
    size_t at = (size_t)k * n + source;
    if (dist[at] != 0 || parents[at] != -1) return;
  }
  double t0 = now_ms();
  unsigned *mask = (unsigned *)xmalloc(sizeof(unsigned) * (size_t)(nIns ? nIns : 1));
  int m2 = or_apply_batch(n, KG, rowPtr, colInd, w, nIns, insFrom, insTo, insW,
                          nDel, delFrom, delTo, outRowPtr, outColInd, outW,
                          mask);
  times_ms[0] = now_ms() - t0;
  if (m2 < 0) {
    free(mask);
    return;
  }
  info[0] = m2;
  /* In-CSR of the updated graph (the sequential path's inAdjacency). */
  int *inRow = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int *inCol = (int *)xmalloc(sizeof(int) * (size_t)(m2 ? m2 : 1));
  int *inW = (int *)xmalloc(sizeof(int) * (size_t)(m2 ? m2 : 1) * KG);
  or_transpose(n, KG, outRowPtr, outColInd, outW, inRow, inCol, inW);
  /* changed_k = deletions + inserts that raised objective k
   * (mospUpdate.cu:124-136). */
  int *chFrom = (int *)xmalloc(sizeof(int) * (size_t)(nDel + nIns + 1));
  int *chTo = (int *)xmalloc(sizeof(int) * (size_t)(nDel + nIns + 1));
  double t1 = now_ms();
  for (int k = 0; k < Kuse; ++k) {
    int nCh = 0;
    for (int i = 0; i < nDel; ++i) {
      chFrom[nCh] = delFrom[i];
      chTo[nCh++] = delTo[i];
    }
    for (int i = 0; i < nIns; ++i) {
      if (mask[i] & (1u << k)) {
        chFrom[nCh] = insFrom[i];
        chTo[nCh++] = insTo[i];
      }
    }
    int it = or_sequential_update(n, outRowPtr, outColInd, outW, inRow, inCol,
                                  inW, KG, k, source, nCh, chFrom, chTo, nIns,
                                  insTo, dist + (size_t)k * n,
                                  parents + (size_t)k * n);
    if (it < 0) goto done;
  }
  times_ms[1] = now_ms() - t1;
  {
    double t2 = now_ms();
    int *cRow = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
    int *cCol = (int *)xmalloc(sizeof(int) * (size_t)Kuse * n);
    int *cW = (int *)xmalloc(sizeof(int) * (size_t)Kuse * n);
    int ec = or_combined_graph(parents, n, Kuse, source, pref, nPref, cRow, cCol,
                               cW);
    if (ec >= 0) {
      or_dijkstra(n, cRow, cCol, cW, 1, 0, source, combDist, combParent);
      info[1] = ec;
      info[2] = or_preference_scale(pref, nPref, Kuse);
      info[3] = 1;
    }
    times_ms[2] = now_ms() - t2;
    free(cW);
    free(cCol);
    free(cRow);
  }
done:
  free(chTo);
  free(chFrom);
  free(inW);
  free(inCol);
  free(inRow);
  free(mask);
}

/* ------------------------------------------------------------------------ */
/* Algorithmic bytes of one update                                          */
/* ------------------------------------------------------------------------ */

int64_t or_alg_bytes(int n, const int *outRow, const int *inRow, int Kuse,
                     int source, const int *oldParents, const int64_t *oldDist,
                     const int64_t *newDist, const int *chOffsets,
                     const int *chFrom, const int *chTo, int nHeads,
                     const int *heads, int64_t combinedEdges) {
  int64_t total = 0;
  char *inI = (char *)xmalloc((size_t)n + 1);
  char *inC = (char *)xmalloc((size_t)n + 1);
  int *stack = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int *childStart = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int *children = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  int *cursor = (int *)xmalloc(sizeof(int) * ((size_t)n + 1));
  for (int k = 0; k < Kuse; ++k) {
    const int *par = oldParents + (size_t)k * n;
    const int64_t *od = oldDist + (size_t)k * n;
    const int64_t *nd = newDist + (size_t)k * n;
    memset(inI, 0, (size_t)n);
    memset(inC, 0, (size_t)n);
    memset(childStart, 0, sizeof(int) * ((size_t)n + 1));
    for (int v = 0; v < n; ++v)
      if (v != source && par[v] >= 0) ++childStart[par[v] + 1];
    for (int v = 0; v < n; ++v) childStart[v + 1] += childStart[v];
    memcpy(cursor, childStart, sizeof(int) * (size_t)n);
    for (int v = 0; v < n; ++v)
      if (v != source && par[v] >= 0) children[cursor[par[v]]++] = v;
    int top = 0;
    for (int i = chOffsets[k]; i < chOffsets[k + 1]; ++i) {
      int v = chTo[i];
      if (par[v] == chFrom[i] && !inI[v]) {
        inI[v] = 1;
        stack[top++] = v;
      }
    }
    while (top > 0) {
      int x = stack[--top];
      for (int c = childStart[x]; c < childStart[x + 1]; ++c) {
        int y = children[c];
        if (!inI[y]) {
          inI[y] = 1;
          stack[top++] = y;
        }
      }
    }
    int64_t b = 28LL * n;
    for (int v = 0; v < n; ++v)
      if (inI[v]) inC[v] = 1;
    for (int i = 0; i < nHeads; ++i)
      if (heads[i] != source) inC[heads[i]] = 1;
    for (int v = 0; v < n; ++v) {
      if (inC[v]) b += 8 + 16LL * (inRow[v + 1] - inRow[v]);
      if (v != source && nd[v] < OR_INF / 2 && (inI[v] || nd[v] != od[v]))
        b += 8 + 16LL * (outRow[v + 1] - outRow[v]);
    }
    total += b;
  }
  total += 4LL * Kuse * n + (4LL * (n + 1) + 8 * combinedEdges) +
           (8LL * n + 16 * combinedEdges) + 12LL * n;
  free(cursor);
  free(children);
  free(childStart);
  free(stack);
  free(inC);
  free(inI);
  return total;
}
