// device_graph.hip -- the graph in HBM (C ABI mosp_device_graph_*).
//
// Upload: the out-CSR is copied once, the edge-major weights (Values.txt
// layout) are split into K objective-major columns and the reverse CSR is
// built on the device (src/deviceGraph.cu:28-135 in the reference).
//
// Apply: a change batch is applied to a resident graph on the device, writing
// the updated graph into a spare graph whose allocations are reused.  The
// result is the host apply's exactly (src/csrGraph.cu:727-894): per row,
// deletions in file order remove the first alive (u,v); insertions in file
// order overwrite the first alive (u,v) -- original entries first, then the
// row's appended ones -- or append; survivors keep their order and appended
// edges follow; the weight-increase mask compares the final (u,v) weight with
// the original first occurrence.  Changes to different targets of a row never
// interact, so the replay is split by (u,v) pair: of a pair with D deletions
// and I insertions, the first D occurrences of v in the original row die, the
// next one takes the last insertion's weights, and with none left the pair
// appends one entry; a row's appended entries follow the order of each pair's
// first insertion.  The host sorts the (small) batch into pairs; one wave per
// pair finds its occurrences, one wave per changed row places the appended
// entries, a scan gives the new row pointers, and rows are copied (edge-
// parallel) or compacted.  No step is serial in a row's degree beyond one
// wave's scan of it, so power-law hub rows cost microseconds, not a thread's
// walk over tens of thousands of entries.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "engine.hpp"
#include "engine_internal.hpp"
#include "mosp_hip.h"

using namespace mosp;
using namespace mosp::detail;

namespace {

constexpr int TILE = 2048;  // scan tile
constexpr int DEAD = -2;    // apply: deleted entry
constexpr int KEEP = -1;    // apply: original weights

__global__ void split_weights_kernel(const int *edgeMajor, long long total, long long stride,
                                     int K, int *columns) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= total) {
    return;
  }
  const long long e = i / K;
  const int k = static_cast<int>(i - e * K);
  columns[k * stride + e] = edgeMajor[i];
}


// {column, weight} records of every objective: edge[k * stride + e] =
// {col[e], w[k * stride + e]} for e < m = rowPtr[n] (read on the device, so
// the host need not know m yet).
__global__ void pair_edges_kernel(const int *rowPtr, int n, const int *col, const int *w,
                                  long long stride, int K, int2 *edge) {
  const long long m = rowPtr[n];
  const long long total = m * K;
  for (long long i = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; i < total;
       i += static_cast<long long>(gridDim.x) * blockDim.x) {
    const long long k = i / m, e = i - k * m;
    edge[k * stride + e] = make_int2(col[e], w[k * stride + e]);
  }
}

// {column, weight, rowPtr[column], rowPtr[column + 1]} records of every
// objective, laid out as pair_edges_kernel's.
__global__ void records4_kernel(const int *rowPtr, int n, const int *col, const int *w,
                                long long stride, int K, int4 *edge) {
  const long long m = rowPtr[n];
  for (long long e = blockIdx.x * static_cast<long long>(blockDim.x) + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const int c = col[e];
    const int b = rowPtr[c], end = rowPtr[c + 1];
    for (int k = 0; k < K; ++k) {
      edge[k * stride + e] = make_int4(c, w[k * stride + e], b, end);
    }
  }
}

__global__ void copy_ints_kernel(int *dst, const int *src, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    dst[i] = src[i];
  }
}

__global__ void tile_sums_kernel(const int *in, int count, int *sums) {
  __shared__ int red[BLOCK];
  const int base = blockIdx.x * TILE;
  int s = 0;
  for (int i = threadIdx.x; i < TILE; i += BLOCK) {
    const int idx = base + i;
    s += idx < count ? in[idx] : 0;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      red[threadIdx.x] += red[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sums[blockIdx.x] = red[0];
  }
}

// Exclusive scan of `sums` in place by one workgroup (any length).
__global__ void scan_sums_kernel(int *sums, int count) {
  __shared__ int buf[BLOCK];
  __shared__ int carry;
  if (threadIdx.x == 0) {
    carry = 0;
  }
  __syncthreads();
  for (int base = 0; base < count; base += BLOCK) {
    const int idx = base + threadIdx.x;
    const int v = idx < count ? sums[idx] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < BLOCK; off <<= 1) {
      const int add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (idx < count) {
      sums[idx] = carry + buf[threadIdx.x] - v;
    }
    __syncthreads();
    if (threadIdx.x == BLOCK - 1) {
      carry += buf[BLOCK - 1];
    }
    __syncthreads();
  }
}

// out[i] = exclusive prefix of in[] (tile offsets from scan_sums_kernel).
__global__ void tile_scan_kernel(const int *in, int count, const int *offsets, int *out) {
  __shared__ int buf[BLOCK];
  constexpr int PER = TILE / BLOCK;
  const int base = blockIdx.x * TILE + threadIdx.x * PER;
  int vals[PER];
  int mine = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int idx = base + j;
    vals[j] = idx < count ? in[idx] : 0;
    mine += vals[j];
  }
  buf[threadIdx.x] = mine;
  __syncthreads();
  for (int off = 1; off < BLOCK; off <<= 1) {
    const int add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += add;
    __syncthreads();
  }
  int running = offsets[blockIdx.x] + buf[threadIdx.x] - mine;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int idx = base + j;
    if (idx < count) {
      out[idx] = running;
    }
    running += vals[j];
  }
}

// Largest u in [lo, hi] with rowPtr[u] <= e: the row holding entry e when
// rowPtr[lo] <= e < rowPtr[hi + 1].
__device__ __forceinline__ int row_of(const int *rowPtr, int lo, int hi, long long e) {
  while (lo < hi) {
    const int mid = lo + (hi - lo + 1) / 2;
    if (rowPtr[mid] <= e) {
      lo = mid;
    } else {
      hi = mid - 1;
    }
  }
  return lo;
}

// Edge-parallel kernels: a workgroup of SWEEP_THREADS lanes takes a
// contiguous chunk of CSR entries, SWEEP_PASS at a time (SWEEP_UNROLL per
// lane, their loads issued together).
constexpr int SWEEP_THREADS = 256;
constexpr int SWEEP_UNROLL = 8;
constexpr int SWEEP_PASS = SWEEP_THREADS * SWEEP_UNROLL;
constexpr int ROW_WINDOW = 2048;  // row pointers staged in LDS at a time

// Rows of a workgroup's entries [c0, c1): row pointers are staged in LDS,
// ROW_WINDOW at a time, and reloaded only when the passes have used them up
// (so a pass costs no global search); each entry's row is a binary search
// there.  pass() and its barriers are workgroup-uniform.
struct RowWindow {
  int *rows;  // LDS, ROW_WINDOW ints
  const int *rowPtr;
  int n;
  int want;   // rowPtr[want] <= the next pass's first entry
  int first = -1, count = 0;
  long long end = 0;  // the window's entries are [rows[0], end)

  __device__ RowWindow(const int *rp, int nRows, long long c0, int *lds)
      : rows(lds), rowPtr(rp), n(nRows), want(row_of(rp, 0, nRows - 1, c0)) {}

  // Stages the window holding entry `base` if needed; returns the end of
  // the pass starting there: min(base + span, c1, window end).
  __device__ long long pass(long long base, long long span, long long c1) {
    if (want != first) {
      __syncthreads();  // the last window's searches are done
      first = want;
      count = n - first < ROW_WINDOW ? n - first : ROW_WINDOW;
      for (int i = static_cast<int>(threadIdx.x); i < count; i += static_cast<int>(blockDim.x)) {
        rows[i] = rowPtr[first + i];
      }
      end = rowPtr[first + count];
      __syncthreads();
    }
    long long stop = std::min(base + span, c1);
    stop = std::min(stop, end);
    if (stop == end) {
      want = first + count;  // rowPtr[first + count] == end == the next base
    }
    return stop;
  }

  // Index in the window of the row holding entry e (staged entries only).
  __device__ int slot(long long e) const {
    int a = 0, b = count - 1;
    while (a < b) {
      const int mid = a + (b - a + 1) / 2;
      if (rows[mid] <= e) {
        a = mid;
      } else {
        b = mid - 1;
      }
    }
    return a;
  }
};

// Entries per chunk of an edge-parallel kernel over at most mBound entries:
// ~2048 chunks (enough workgroups, few per-chunk counts), 4096-32768
// entries each, at most 2^20 chunks.
long long sweep_chunk(long long mBound) {
  long long chunk = std::max(4096LL, std::min(32768LL, (mBound + 2047) / 2048));
  chunk = std::max(chunk, (mBound + (1LL << 20) - 1) >> 20);
  return (chunk + SWEEP_PASS - 1) / SWEEP_PASS * SWEEP_PASS;
}

int sweep_grid(long long mBound) {
  const long long chunk = sweep_chunk(mBound);
  return static_cast<int>(std::max(1LL, (mBound + chunk - 1) / chunk));
}

// Reverse CSR, direct: one atomic per out-entry on its target's in-degree,
// then on its cursor.  Cheapest when no target is a hub (the atomics do not
// contend); graphs with hubs use the bucketed transpose below.
// In-degrees of the m = *edges arcs (edge count read on the device, so a
// graph rebuilt on the device needs no host round trip).
__global__ void in_degree_kernel(const int *colInd, const int *edges, int *degree) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < *edges) {
    atomicAdd(&degree[colInd[e]], 1);
  }
}

// Places every out-entry in its target's in-row (cursor: the in-rows'
// starts, advanced by one atomic per entry; the order inside an in-row
// follows the atomics, and every consumer of the in-CSR is order-free).
__global__ void __launch_bounds__(SWEEP_THREADS)
fill_reverse_kernel(int n, long long stride, int K, const int *rowPtr, const int *colInd,
                    const int *outW, long long chunk, int *cursor, int *inCol, int *inW) {
  __shared__ int rows[ROW_WINDOW];
  const long long m = rowPtr[n];
  const long long c0 = static_cast<long long>(blockIdx.x) * chunk;
  if (c0 >= m) {
    return;
  }
  const long long c1 = std::min(c0 + chunk, m);
  RowWindow win(rowPtr, n, c0, rows);
  for (long long base = c0; base < c1;) {
    const long long stop = win.pass(base, SWEEP_PASS, c1);
    int t[SWEEP_UNROLL], pos[SWEEP_UNROLL];
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      const long long e = base + j * SWEEP_THREADS + threadIdx.x;
      t[j] = e < stop ? colInd[e] : -1;
    }
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      if (t[j] >= 0) {
        const long long e = base + j * SWEEP_THREADS + threadIdx.x;
        pos[j] = atomicAdd(&cursor[t[j]], 1);
        inCol[pos[j]] = win.first + win.slot(e);
      }
    }
    for (int k = 0; k < K; ++k) {
      int w[SWEEP_UNROLL];
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        const long long e = base + j * SWEEP_THREADS + threadIdx.x;
        w[j] = t[j] >= 0 ? outW[k * stride + e] : 0;
      }
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        if (t[j] >= 0) {
          inW[k * stride + pos[j]] = w[j];
        }
      }
    }
    base = stop;
  }
}

// Reverse CSR by a bucketed transpose.  Targets fall in buckets of 2^bits
// consecutive ids.  (1) Every chunk of out-entries counts its entries per
// bucket in LDS; (2) a scan of the bucket-major (bucket, chunk) counts gives
// every chunk its places in a record array sorted by bucket; (3) the chunk
// scatters (target, source, weights) records there, a bucket's entries of a
// chunk in one run; (4) in-degrees and (5) in-row places are then taken per
// tile of bucket-sorted records with one LDS counter per target of the
// tile's (at most TR_WINDOW) targets, so a target costs one global atomic per
// tile instead of one per entry -- power-law hubs no longer serialise the
// memory-side atomics, and a target's entries of a tile land in one run of
// its in-row.  A tile whose records span more targets (sparse buckets) uses
// one global atomic per record.  The order inside an in-row is unspecified;
// every consumer of the in-CSR is order-free.
constexpr int TR_MAX_BUCKETS = 8192;
constexpr int TR_MIN_BITS = 11;
constexpr int TR_WINDOW = 4096;         // targets one tile can count in LDS
constexpr int TR_TILE = 16384;          // records per tile of passes (4) and (5)

// (1) counts[b * chunks + c] = entries of chunk c whose target is in bucket b.
__global__ void __launch_bounds__(SWEEP_THREADS)
tr_count_kernel(const int *rowPtr, int n, const int *colInd, int bits, int buckets,
                long long chunk, int chunks, int *counts) {
  extern __shared__ int hist[];  // buckets counters
  const long long m = rowPtr[n];
  const int c = static_cast<int>(blockIdx.x);
  for (int b = static_cast<int>(threadIdx.x); b < buckets; b += SWEEP_THREADS) {
    hist[b] = 0;
  }
  __syncthreads();
  const long long c0 = c * chunk;
  const long long c1 = std::min(c0 + chunk, m);
  for (long long e = c0 + threadIdx.x; e < c1; e += SWEEP_THREADS) {
    atomicAdd(&hist[colInd[e] >> bits], 1);
  }
  __syncthreads();
  for (int b = static_cast<int>(threadIdx.x); b < buckets; b += SWEEP_THREADS) {
    counts[static_cast<long long>(b) * chunks + c] = hist[b];
  }
}

// (3) Records of chunk c at places[b * chunks + c] onward, per bucket b
// (rows from a RowWindow).
__global__ void __launch_bounds__(SWEEP_THREADS)
tr_scatter_kernel(const int *rowPtr, int n, const int *colInd, const int *outW,
                  long long stride, int K, int bits, int buckets, long long chunk, int chunks,
                  const int *places, long long recCap, int *recT, int *recU, int *recW) {
  extern __shared__ int lds[];  // buckets next-places, then ROW_WINDOW row pointers
  int *const next = lds;
  int *const rows = lds + buckets;
  const long long m = rowPtr[n];
  const int c = static_cast<int>(blockIdx.x);
  const long long c0 = c * chunk;
  if (c0 >= m) {
    return;
  }
  const long long c1 = std::min(c0 + chunk, m);
  for (int b = static_cast<int>(threadIdx.x); b < buckets; b += SWEEP_THREADS) {
    next[b] = places[static_cast<long long>(b) * chunks + c];
  }
  RowWindow win(rowPtr, n, c0, rows);  // its first pass's barrier also orders next[]
  for (long long base = c0; base < c1;) {
    const long long stop = win.pass(base, SWEEP_PASS, c1);
    int t[SWEEP_UNROLL], at[SWEEP_UNROLL];
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      const long long e = base + j * SWEEP_THREADS + threadIdx.x;
      t[j] = e < stop ? colInd[e] : -1;
    }
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      if (t[j] < 0) {
        continue;
      }
      const long long e = base + j * SWEEP_THREADS + threadIdx.x;
      at[j] = atomicAdd(&next[t[j] >> bits], 1);
      recT[at[j]] = t[j];
      recU[at[j]] = win.first + win.slot(e);
    }
    for (int k = 0; k < K; ++k) {
      int w[SWEEP_UNROLL];
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        const long long e = base + j * SWEEP_THREADS + threadIdx.x;
        w[j] = t[j] >= 0 ? outW[k * stride + e] : 0;
      }
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        if (t[j] >= 0) {
          recW[k * recCap + at[j]] = w[j];
        }
      }
    }
    base = stop;
  }
}

// The buckets of a tile of bucket-sorted records: those of its first and
// last records; its targets lie in [lo, lo + width) (tr_tile_targets).
__device__ __forceinline__ void tr_tile_span(const int *recT, long long i0, long long i1,
                                             int bits, int *span) {
  if (threadIdx.x == 0) {
    span[0] = recT[i0] >> bits;
    span[1] = recT[i1 - 1] >> bits;
  }
}

__device__ __forceinline__ void tr_tile_targets(const int *span, int bits, int &lo,
                                                long long &width) {
  lo = span[0] << bits;
  width = static_cast<long long>(span[1] - span[0] + 1) << bits;
}

// (4) In-degrees (degree[] zeroed; m read from rowPtr[n]).
__global__ void __launch_bounds__(SWEEP_THREADS)
tr_degree_kernel(const int *rowPtr, int n, const int *recT, int bits, int *degree) {
  __shared__ int counts[TR_WINDOW];
  __shared__ int span[2];
  const long long m = rowPtr[n];
  const long long i0 = static_cast<long long>(blockIdx.x) * TR_TILE;
  if (i0 >= m) {
    return;
  }
  const long long i1 = std::min(i0 + TR_TILE, m);
  tr_tile_span(recT, i0, i1, bits, span);
  for (int j = static_cast<int>(threadIdx.x); j < TR_WINDOW; j += SWEEP_THREADS) {
    counts[j] = 0;
  }
  __syncthreads();
  int lo;
  long long width;
  tr_tile_targets(span, bits, lo, width);
  const bool local = width <= TR_WINDOW;
  for (long long i = i0 + threadIdx.x; i < i1; i += SWEEP_THREADS * SWEEP_UNROLL) {
    int t[SWEEP_UNROLL];
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      const long long r = i + j * SWEEP_THREADS;
      t[j] = r < i1 ? recT[r] : -1;
    }
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      if (t[j] < 0) {
        continue;
      }
      if (local) {
        atomicAdd(&counts[t[j] - lo], 1);
      } else {
        atomicAdd(&degree[t[j]], 1);
      }
    }
  }
  __syncthreads();
  if (local) {
    for (int j = static_cast<int>(threadIdx.x); j < width; j += SWEEP_THREADS) {
      if (counts[j] > 0) {
        atomicAdd(&degree[lo + j], counts[j]);
      }
    }
  }
}

// (5) In-row places: a tile reserves each target's run with one atomic on
// the target's cursor (the in-row starts, advanced by reservations), then
// hands out the run's places with LDS atomics.
__global__ void __launch_bounds__(SWEEP_THREADS)
tr_place_kernel(const int *rowPtr, int n, int K, const int *recT, const int *recU,
                const int *recW, long long recCap, int bits, int *cursor, int *inCol, int *inW,
                long long stride) {
  __shared__ int slots[TR_WINDOW];  // counts, then each target's next place
  __shared__ int span[2];
  const long long m = rowPtr[n];
  const long long i0 = static_cast<long long>(blockIdx.x) * TR_TILE;
  if (i0 >= m) {
    return;
  }
  const long long i1 = std::min(i0 + TR_TILE, m);
  tr_tile_span(recT, i0, i1, bits, span);
  for (int j = static_cast<int>(threadIdx.x); j < TR_WINDOW; j += SWEEP_THREADS) {
    slots[j] = 0;
  }
  __syncthreads();
  int lo;
  long long width;
  tr_tile_targets(span, bits, lo, width);
  const bool local = width <= TR_WINDOW;
  if (local) {
    for (long long i = i0 + threadIdx.x; i < i1; i += SWEEP_THREADS) {
      atomicAdd(&slots[recT[i] - lo], 1);
    }
    __syncthreads();
    for (int j = static_cast<int>(threadIdx.x); j < width; j += SWEEP_THREADS) {
      if (slots[j] > 0) {
        slots[j] = atomicAdd(&cursor[lo + j], slots[j]);
      }
    }
    __syncthreads();
  }
  for (long long i = i0 + threadIdx.x; i < i1; i += SWEEP_THREADS * SWEEP_UNROLL) {
    int t[SWEEP_UNROLL], u[SWEEP_UNROLL], pos[SWEEP_UNROLL];
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      const long long r = i + j * SWEEP_THREADS;
      t[j] = r < i1 ? recT[r] : -1;
      u[j] = r < i1 ? recU[r] : 0;
    }
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      if (t[j] >= 0) {
        pos[j] = local ? atomicAdd(&slots[t[j] - lo], 1) : atomicAdd(&cursor[t[j]], 1);
        inCol[pos[j]] = u[j];
      }
    }
    for (int k = 0; k < K; ++k) {
      int w[SWEEP_UNROLL];
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        const long long r = i + j * SWEEP_THREADS;
        w[j] = t[j] >= 0 ? recW[k * recCap + r] : 0;
      }
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        if (t[j] >= 0) {
          inW[k * stride + pos[j]] = w[j];
        }
      }
    }
  }
}

__global__ void zero_kernel(int *a, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    a[i] = 0;
  }
}

// Largest row length of rowPtr[0..n] folded into *result: a workgroup
// reduction, then ONE atomicMax per workgroup (same-address atomics
// serialize, so one per wave cost ~0.2 ms at n = 2M).
__global__ void max_degree_kernel(const int *rowPtr, int n, unsigned *result) {
  __shared__ unsigned red[BLOCK / 64];
  unsigned best = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    best = max(best, static_cast<unsigned>(rowPtr[i + 1] - rowPtr[i]));
  }
  for (int offset = 32; offset > 0; offset >>= 1) {
    best = max(best, static_cast<unsigned>(__shfl_xor(static_cast<int>(best), offset)));
  }
  if (__lane_id() == 0) {
    red[threadIdx.x / 64] = best;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned top = 0;
    for (unsigned i = 0; i < blockDim.x / 64; ++i) {
      top = max(top, red[i]);
    }
    if (top > 0) {
      atomicMax(result, top);
    }
  }
}

// Per-objective maximum and sum of the weights of column `k` (one grid-stride
// pass per objective; workgroup reductions, then one atomic each).
__global__ void weight_stats_kernel(const int *w, int m, unsigned long long *sum,
                                    unsigned long long *maxv) {
  __shared__ unsigned long long redSum[BLOCK / 64];
  __shared__ unsigned long long redMax[BLOCK / 64];
  unsigned long long s = 0, mx = 0;
  for (long long e = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<long long>(gridDim.x) * blockDim.x) {
    const unsigned long long v = static_cast<unsigned long long>(w[e]);
    s += v;
    mx = v > mx ? v : mx;
  }
  for (int off = 32; off > 0; off >>= 1) {
    s += __shfl_xor(s, off);
    const unsigned long long o = __shfl_xor(mx, off);
    mx = o > mx ? o : mx;
  }
  if (__lane_id() == 0) {
    redSum[threadIdx.x / 64] = s;
    redMax[threadIdx.x / 64] = mx;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long ts = 0, tm = 0;
    for (unsigned i = 0; i < blockDim.x / 64; ++i) {
      ts += redSum[i];
      tm = redMax[i] > tm ? redMax[i] : tm;
    }
    atomicAdd(sum, ts);
    atomicMax(maxv, tm);
  }
}

// ------------------------------------------------------------------ apply --

// Degrees, row tags and every original entry's slot (KEEP) before the pair
// pass, which then only writes the entries its pair kills or overwrites.
__global__ void apply_init_kernel(int n, long long m, const int *oldRow, int *newDeg,
                                  int *rowTag, int *slot) {
  const long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    newDeg[i] = oldRow[i + 1] - oldRow[i];
    rowTag[i] = -1;
  } else if (i == n) {
    newDeg[n] = 0;
  }
  if (i < m) {
    slot[i] = KEEP;
  }
}

__device__ __forceinline__ int wave_sum(int x) {
  for (int off = 32; off > 0; off >>= 1) {
    x += __shfl_xor(x, off);
  }
  return x;
}

// One wave per (u, v) pair: D = pairDel deletions and the insertions
// insSorted[pairInsBeg .. +pairInsCnt) (file order; the last one is
// pairLast).  Marks the first D occurrences of v in the original row DEAD and
// the next one with pairLast; pairApp = 1 when no occurrence is left for an
// insertion (the pair appends one entry).  The pair's insertions all get the
// same weight-increase bits: the last insertion's weights against the first
// original occurrence's (none when v was not in the row).  Pairs are distinct
// (u, v), so no two waves write one entry.
__global__ void apply_pairs_kernel(int P, int K, const int *pairU, const int *pairV,
                                   const int *pairDel, const int *pairInsBeg,
                                   const int *pairInsCnt, const int *pairLast,
                                   const int *insSorted, const int *oldRow, const int *oldCol,
                                   const int *oldW, long long oldStride, const int *insW,
                                   int *slot, int *pairDead, int *pairApp, unsigned *mask) {
  const int p = static_cast<int>(blockIdx.x);  // one wave per workgroup
  if (p >= P) {
    return;
  }
  const int lane = static_cast<int>(__lane_id());
  const unsigned long long below = (1ull << lane) - 1ull;
  const int u = pairU[p], v = pairV[p], D = pairDel[p], I = pairInsCnt[p], last = pairLast[p];
  const int need = D + (I > 0 ? 1 : 0);
  const int ob = oldRow[u], oe = oldRow[u + 1];
  // PAIR_UNROLL groups of 64 entries a step, their loads issued together
  // (entries read past the last needed occurrence are only read).
  constexpr int PAIR_UNROLL = 8;
  int found = 0, first = -1;
  for (int base = ob; base < oe && found < need; base += 64 * PAIR_UNROLL) {
    int col[PAIR_UNROLL];
    for (int g = 0; g < PAIR_UNROLL; ++g) {
      const int x = base + 64 * g + lane;
      col[g] = x < oe ? oldCol[x] : -1;
    }
    for (int g = 0; g < PAIR_UNROLL; ++g) {
      const int x = base + 64 * g + lane;
      const bool hit = col[g] == v;
      const unsigned long long bal = __ballot(hit);
      if (hit) {
        const int j = found + __popcll(bal & below);
        if (j < D) {
          slot[x] = DEAD;
        } else if (j == D && I > 0) {
          slot[x] = last;
        }
      }
      if (first < 0 && bal != 0) {
        first = base + 64 * g + __ffsll(static_cast<long long>(bal)) - 1;
      }
      found += __popcll(bal);
    }
  }
  if (lane == 0) {
    pairDead[p] = std::min(D, found);
    pairApp[p] = (I > 0 && found <= D) ? 1 : 0;
  }
  if (I > 0) {
    const bool up = lane < K && first >= 0 &&
                    insW[static_cast<size_t>(last) * K + lane] > oldW[lane * oldStride + first];
    const unsigned bits = static_cast<unsigned>(__ballot(up));
    const int beg = pairInsBeg[p];
    for (int t = lane; t < I; t += 64) {
      mask[insSorted[beg + t]] = bits;
    }
  }
}

// One wave per changed row: its new degree, and each appending pair's index
// among the row's appended entries (a row's pairs are ordered by their first
// insertion in file order, pairs without insertions last).
__global__ void apply_rows_kernel(int R, const int *rows, const int *rowPairs, const int *oldRow,
                                  const int *pairDead, const int *pairApp, int *pairAppPos,
                                  int *newDeg, int *rowTag) {
  const int r = static_cast<int>(blockIdx.x);  // one wave per workgroup
  if (r >= R) {
    return;
  }
  const int lane = static_cast<int>(__lane_id());
  const unsigned long long below = (1ull << lane) - 1ull;
  const int u = rows[r], pe = rowPairs[r + 1];
  int dead = 0, app = 0;
  for (int base = rowPairs[r]; base < pe; base += 64) {
    const int p = base + lane;
    const bool in = p < pe;
    const bool a = in && pairApp[p] != 0;
    const unsigned long long bal = __ballot(a);
    if (a) {
      pairAppPos[p] = app + __popcll(bal & below);
    }
    app += __popcll(bal);
    dead += in ? pairDead[p] : 0;
  }
  dead = wave_sum(dead);
  if (lane == 0) {
    newDeg[u] = oldRow[u + 1] - oldRow[u] - dead + app;
    rowTag[u] = r;
  }
}

// Unchanged rows: every original entry's column and K weights move to the
// same offset in its row's new place (rows from a RowWindow).
__global__ void __launch_bounds__(SWEEP_THREADS)
apply_copy_kernel(int n, int K, const int *oldRow, const int *oldCol, const int *oldW,
                  long long oldStride, const int *rowTag, const int *newRow, int *newCol,
                  int *newW, long long newStride, long long chunk) {
  __shared__ int rows[ROW_WINDOW];
  const long long m = oldRow[n];
  const long long c0 = static_cast<long long>(blockIdx.x) * chunk;
  if (c0 >= m) {
    return;
  }
  const long long c1 = std::min(c0 + chunk, m);
  RowWindow win(oldRow, n, c0, rows);
  for (long long base = c0; base < c1;) {
    const long long stop = win.pass(base, SWEEP_PASS, c1);
    long long pos[SWEEP_UNROLL];
    int col[SWEEP_UNROLL];
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      const long long e = base + j * SWEEP_THREADS + threadIdx.x;
      pos[j] = -1;
      if (e < stop) {
        const int a = win.slot(e);
        const int u = win.first + a;
        if (rowTag[u] < 0) {
          pos[j] = newRow[u] + (e - rows[a]);
          col[j] = oldCol[e];
        }
      }
    }
    for (int j = 0; j < SWEEP_UNROLL; ++j) {
      if (pos[j] >= 0) {
        newCol[pos[j]] = col[j];
      }
    }
    for (int k = 0; k < K; ++k) {
      int w[SWEEP_UNROLL];
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        const long long e = base + j * SWEEP_THREADS + threadIdx.x;
        w[j] = pos[j] >= 0 ? oldW[k * oldStride + e] : 0;
      }
      for (int j = 0; j < SWEEP_UNROLL; ++j) {
        if (pos[j] >= 0) {
          newW[k * newStride + pos[j]] = w[j];
        }
      }
    }
    base = stop;
  }
}

// Changed rows: one workgroup per row compacts the alive original entries in
// order, COMPACT_PASS entries a pass (each lane holds COMPACT_PER_THREAD,
// loaded together; per-wave alive counts in LDS give every entry its place),
// then writes the appended ones at their pairs' places.  A power-law hub row
// of tens of thousands of entries is a few dozen passes, not hundreds of one
// wave's dependent ones.
constexpr int COMPACT_THREADS = 256;
constexpr int COMPACT_PER_THREAD = 4;
constexpr int COMPACT_WAVES = COMPACT_THREADS / 64;
constexpr int COMPACT_PASS = COMPACT_THREADS * COMPACT_PER_THREAD;

__global__ void __launch_bounds__(COMPACT_THREADS)
apply_compact_kernel(int R, int K, const int *__restrict__ rows,
                     const int *__restrict__ rowPairs, const int *__restrict__ oldRow,
                     const int *__restrict__ oldCol, const int *__restrict__ oldW,
                     long long oldStride, const int *__restrict__ slot,
                     const int *__restrict__ pairV, const int *__restrict__ pairLast,
                     const int *__restrict__ pairApp, const int *__restrict__ pairAppPos,
                     const int *__restrict__ insW, const int *__restrict__ newRow,
                     int *__restrict__ newCol, int *__restrict__ newW, long long newStride) {
  // Alive counts of a pass, (entry group j, wave) in entry order; two
  // buffers, so a pass's counts are not overwritten while a slower wave
  // still reads the previous pass's.
  __shared__ int counts[2][COMPACT_PER_THREAD * COMPACT_WAVES];
  const int r = static_cast<int>(blockIdx.x);
  if (r >= R) {
    return;
  }
  const int lane = static_cast<int>(__lane_id());
  const int wave = static_cast<int>(threadIdx.x) / 64;
  const int u = rows[r];
  const int ob = oldRow[u], oe = oldRow[u + 1];
  int out = newRow[u];
  const unsigned long long below = (1ull << lane) - 1ull;
  int buf = 0;
  for (int base = ob; base < oe; base += COMPACT_PASS, buf ^= 1) {
    int s[COMPACT_PER_THREAD], col[COMPACT_PER_THREAD], rank[COMPACT_PER_THREAD];
    for (int j = 0; j < COMPACT_PER_THREAD; ++j) {
      const int e = base + j * COMPACT_THREADS + static_cast<int>(threadIdx.x);
      s[j] = e < oe ? slot[e] : DEAD;
      col[j] = e < oe ? oldCol[e] : 0;
    }
    for (int j = 0; j < COMPACT_PER_THREAD; ++j) {
      const unsigned long long alive = __ballot(s[j] != DEAD);
      rank[j] = __popcll(alive & below);
      if (lane == 0) {
        counts[buf][j * COMPACT_WAVES + wave] = __popcll(alive);
      }
    }
    __syncthreads();
    int before[COMPACT_PER_THREAD];
    int total = 0;
    for (int i = 0; i < COMPACT_PER_THREAD * COMPACT_WAVES; ++i) {
      for (int j = 0; j < COMPACT_PER_THREAD; ++j) {
        if (i == j * COMPACT_WAVES + wave) {
          before[j] = total;
        }
      }
      total += counts[buf][i];
    }
    for (int j = 0; j < COMPACT_PER_THREAD; ++j) {
      if (s[j] == DEAD) {
        continue;
      }
      const int e = base + j * COMPACT_THREADS + static_cast<int>(threadIdx.x);
      const int pos = out + before[j] + rank[j];
      newCol[pos] = col[j];
      for (int k = 0; k < K; ++k) {
        newW[k * newStride + pos] =
            s[j] == KEEP ? oldW[k * oldStride + e] : insW[static_cast<size_t>(s[j]) * K + k];
      }
    }
    out += total;
  }
  for (int p = rowPairs[r] + static_cast<int>(threadIdx.x); p < rowPairs[r + 1];
       p += COMPACT_THREADS) {
    if (pairApp[p] != 0) {
      const int pos = out + pairAppPos[p];
      newCol[pos] = pairV[p];
      for (int k = 0; k < K; ++k) {
        newW[k * newStride + pos] = insW[static_cast<size_t>(pairLast[p]) * K + k];
      }
    }
  }
}

int grid_for(long long work) {
  return static_cast<int>((std::max(work, 1LL) + BLOCK - 1) / BLOCK);
}


// Exclusive scan of count ints (count >= 1); `sums` holds scan_tiles(count).
int scan_tiles(int count) { return (count + TILE - 1) / TILE; }

int exclusive_scan(const int *in, int *out, int count, int *sums, hipStream_t stream) {
  const int tiles = scan_tiles(count);
  hipLaunchKernelGGL(tile_sums_kernel, dim3(tiles), dim3(BLOCK), 0, stream, in, count, sums);
  hipLaunchKernelGGL(scan_sums_kernel, dim3(1), dim3(BLOCK), 0, stream, sums, tiles);
  hipLaunchKernelGGL(tile_scan_kernel, dim3(tiles), dim3(BLOCK), 0, stream, in, count, sums,
                     out);
  HIP_TRY(hipGetLastError());
  return 1;
}

// Carves typed arrays out of a graph's scratch buffer.
struct Carve {
  char *p;
  template <typename T>
  T *take(size_t count) {
    T *r = reinterpret_cast<T *>(p);
    p += align256(sizeof(T) * std::max<size_t>(count, 1));
    return r;
  }
  template <typename T>
  static size_t bytes(size_t count) {
    return align256(sizeof(T) * std::max<size_t>(count, 1));
  }
};

// The graph's scratch holds at least `bytes` (grow-only: growing frees the
// old buffer, a device-wide synchronization).
int scratch_reserve(mosp_device_graph *g, size_t bytes) {
  if (g->scratchBytes >= bytes) {
    return 1;
  }
  (void)hipFree(g->scratch);
  g->scratch = nullptr;
  g->scratchBytes = 0;
  const size_t want = bytes + bytes / 4 + (1u << 20);
  HIP_TRY(hipMalloc(&g->scratch, want));
  g->scratchBytes = want;
  return 1;
}

void release_arrays(mosp_device_graph *g) {
  (void)hipFree(g->stats);
  g->stats = nullptr;
  (void)hipFree(g->scratch);
  g->scratch = nullptr;
  g->scratchBytes = 0;
  (void)hipFree(g->outRow);
  (void)hipFree(g->outCol);
  (void)hipFree(g->outW);
  (void)hipFree(g->inRow);
  (void)hipFree(g->inCol);
  (void)hipFree(g->inW);
  (void)hipFree(g->outEdge);
  (void)hipFree(g->inEdge);
  (void)hipFree(g->outEdge4);
  g->outRow = g->outCol = g->outW = g->inRow = g->inCol = g->inW = nullptr;
  g->outEdge = g->inEdge = nullptr;
  g->outEdge4 = nullptr;
  g->rec4Valid = false;
  g->stride = 0;
  g->rowCap = 0;
}

// Room for n vertices and m edges of K objectives; grows with headroom and
// keeps the arrays when they are large enough (a spare graph is rebuilt in
// place).  Growing frees the old arrays (a device-wide synchronization).
int reserve(mosp_device_graph *g, int n, long long m, int K) {
  const long long need = std::max(m, 1LL);
  if (g->rowCap >= n + 1 && g->stride >= need && g->K >= K && g->outRow != nullptr) {
    return 1;
  }
  release_arrays(g);
  const long long stride = need + std::max(need / 8, 4096LL);
  const size_t rows = sizeof(int) * static_cast<size_t>(n + 1);
  const size_t cols = sizeof(int) * static_cast<size_t>(stride);
  HIP_TRY(hipMalloc(&g->outRow, rows));
  HIP_TRY(hipMalloc(&g->outCol, cols));
  HIP_TRY(hipMalloc(&g->outW, cols * K));
  HIP_TRY(hipMalloc(&g->inRow, rows));
  HIP_TRY(hipMalloc(&g->inCol, cols));
  HIP_TRY(hipMalloc(&g->inW, cols * K));
  HIP_TRY(hipMalloc(&g->outEdge, sizeof(int2) * static_cast<size_t>(stride) * K));
  HIP_TRY(hipMalloc(&g->inEdge, sizeof(int2) * static_cast<size_t>(stride) * K));
  g->stride = stride;
  g->rowCap = n + 1;
  g->K = K;
  return 1;
}

// Bucket size, chunking and record capacity of the transpose of a graph with
// n vertices and at most mBound entries.
struct TransposePlan {
  int bits;          // bucket of target t = t >> bits
  int buckets;
  long long chunk;   // out-entries per chunk of passes (1) and (3)
  int chunks;
  long long recCap;  // records (>= mBound)
};

TransposePlan transpose_plan(int n, long long mBound) {
  TransposePlan p{};
  p.bits = TR_MIN_BITS;
  while (((static_cast<long long>(n) + (1LL << p.bits) - 1) >> p.bits) > TR_MAX_BUCKETS) {
    ++p.bits;
  }
  p.buckets = static_cast<int>((static_cast<long long>(n) + (1LL << p.bits) - 1) >> p.bits);
  p.chunk = sweep_chunk(mBound);
  p.chunks = sweep_grid(mBound);
  p.recCap = std::max(mBound, 1LL);
  return p;
}

// Scratch bytes of build_reverse (hubs: the bucketed transpose, else the
// direct build).
size_t reverse_bytes(int n, long long mBound, int K, bool hubs) {
  const size_t direct = Carve::bytes<int>(n + 1) + Carve::bytes<int>(scan_tiles(n + 1));
  if (!hubs) {
    return direct;
  }
  const TransposePlan p = transpose_plan(n, mBound);
  const long long cells = static_cast<long long>(p.buckets) * p.chunks;
  return Carve::bytes<int>(n + 1) * 2 + Carve::bytes<int>(scan_tiles(n + 1)) +
         Carve::bytes<int>(cells) * 2 +
         Carve::bytes<int>(scan_tiles(static_cast<int>(cells))) +
         Carve::bytes<int>(p.recCap) * (2 + static_cast<size_t>(K));
}

// In-CSR by one atomic per entry (no hubs).
int build_reverse_direct(mosp_device_graph *g, char *at, long long mBound, hipStream_t stream) {
  const int n = g->n, K = g->K;
  Carve c{at};
  int *cursor = c.take<int>(n + 1);
  int *sums = c.take<int>(scan_tiles(n + 1));
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n + 1)), dim3(BLOCK), 0, stream, cursor, n + 1);
  hipLaunchKernelGGL(in_degree_kernel, dim3(grid_for(mBound)), dim3(BLOCK), 0, stream, g->outCol,
                     g->outRow + n, cursor);
  if (!exclusive_scan(cursor, g->inRow, n + 1, sums, stream)) {
    return 0;
  }
  hipLaunchKernelGGL(copy_ints_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, stream, cursor,
                     g->inRow, n);
  hipLaunchKernelGGL(fill_reverse_kernel, dim3(sweep_grid(mBound)), dim3(SWEEP_THREADS), 0, stream,
                     n, g->stride, K, g->outRow, g->outCol, g->outW, sweep_chunk(mBound), cursor,
                     g->inCol, g->inW);
  HIP_TRY(hipGetLastError());
  return 1;
}

// In-CSR (rows, columns, K weight columns) from the out-CSR, on `stream`;
// temporaries at `at` (reverse_bytes(n, mBound, K, hubs)); mBound >= the
// edge count, which is read from outRow[n] on the device.  Only kernels on
// one stream: each sees its predecessors' writes.  `hubs`: some vertex has
// more than HUB_DEGREE edges, so one atomic per entry would contend.
int build_reverse(mosp_device_graph *g, char *at, long long mBound, bool hubs,
                  hipStream_t stream) {
  if (!hubs) {
    return build_reverse_direct(g, at, mBound, stream);
  }
  const int n = g->n, K = g->K;
  const TransposePlan p = transpose_plan(n, mBound);
  const long long cells = static_cast<long long>(p.buckets) * p.chunks;
  Carve c{at};
  int *degree = c.take<int>(n + 1);
  int *cursor = c.take<int>(n + 1);
  int *sums = c.take<int>(scan_tiles(n + 1));
  int *counts = c.take<int>(cells);
  int *places = c.take<int>(cells);
  int *cellSums = c.take<int>(scan_tiles(static_cast<int>(cells)));
  int *recT = c.take<int>(p.recCap);
  int *recU = c.take<int>(p.recCap);
  int *recW = c.take<int>(p.recCap * std::max(K, 1));
  const int tiles = static_cast<int>((p.recCap + TR_TILE - 1) / TR_TILE);
  const size_t countLds = sizeof(int) * p.buckets;
  const size_t scatterLds = sizeof(int) * (p.buckets + ROW_WINDOW);
  hipLaunchKernelGGL(tr_count_kernel, dim3(p.chunks), dim3(SWEEP_THREADS), countLds, stream,
                     g->outRow, n, g->outCol, p.bits, p.buckets, p.chunk, p.chunks, counts);
  if (!exclusive_scan(counts, places, static_cast<int>(cells), cellSums, stream)) {
    return 0;
  }
  hipLaunchKernelGGL(tr_scatter_kernel, dim3(p.chunks), dim3(SWEEP_THREADS), scatterLds, stream,
                     g->outRow, n, g->outCol, g->outW, g->stride, K, p.bits, p.buckets, p.chunk,
                     p.chunks, places, p.recCap, recT, recU, recW);
  hipLaunchKernelGGL(zero_kernel, dim3(grid_for(n + 1)), dim3(BLOCK), 0, stream, degree, n + 1);
  hipLaunchKernelGGL(tr_degree_kernel, dim3(tiles), dim3(SWEEP_THREADS), 0, stream, g->outRow, n,
                     recT, p.bits, degree);
  if (!exclusive_scan(degree, g->inRow, n + 1, sums, stream)) {
    return 0;
  }
  hipLaunchKernelGGL(copy_ints_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, stream, cursor,
                     g->inRow, n);
  hipLaunchKernelGGL(tr_place_kernel, dim3(tiles), dim3(SWEEP_THREADS), 0, stream, g->outRow, n, K,
                     recT, recU, recW, p.recCap, p.bits, cursor, g->inCol, g->inW, g->stride);
  HIP_TRY(hipGetLastError());
  return 1;
}

// Queues the largest out- or in-degree into *scratch (zeroed first).
int queue_max_degree(const int *outRow, const int *inRow, int n, unsigned *scratch,
                     hipStream_t stream) {
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<int *>(scratch),
                     1);
  const int blocks = std::max(1, std::min(256, grid_for(n)));
  hipLaunchKernelGGL(max_degree_kernel, dim3(blocks), dim3(BLOCK), 0, stream, outRow, n, scratch);
  if (inRow != nullptr) {
    hipLaunchKernelGGL(max_degree_kernel, dim3(blocks), dim3(BLOCK), 0, stream, inRow, n,
                       scratch);
  }
  HIP_TRY(hipGetLastError());
  return 1;
}

// The out- and in-edge records of g from its arrays (after an upload or an
// apply, once the reverse CSR is built); mBound >= m sizes the grid.
int pair_edges(mosp_device_graph *g, long long mBound, hipStream_t stream) {
  const long long total = std::max(mBound, 1LL) * g->K;
  const int blocks = static_cast<int>(std::min<long long>((total + BLOCK - 1) / BLOCK, 8192));
  hipLaunchKernelGGL(pair_edges_kernel, dim3(blocks), dim3(BLOCK), 0, stream, g->outRow, g->n,
                     g->outCol, g->outW, g->stride, g->K, g->outEdge);
  hipLaunchKernelGGL(pair_edges_kernel, dim3(blocks), dim3(BLOCK), 0, stream, g->inRow, g->n,
                     g->inCol, g->inW, g->stride, g->K, g->inEdge);
  HIP_TRY(hipGetLastError());
  g->rec4Valid = false;
  return 1;
}

// Statistics words of a graph: 2 * MOSP_MAX_OBJ weight words, then the degree.
int ensure_stats(mosp_device_graph *g) {
  if (g->stats == nullptr) {
    HIP_TRY(hipMalloc(&g->stats, sizeof(unsigned long long) * (2 * MOSP_MAX_OBJ + 1)));
  }
  return 1;
}

unsigned *degree_word(mosp_device_graph *g) {
  return reinterpret_cast<unsigned *>(g->stats + 2 * MOSP_MAX_OBJ);
}

// The rules readCsrGraph enforces on a text graph, checked on the caller's
// arrays before anything is copied (the reverse build indexes by them, so a
// bad column would be an out-of-bounds device access): rowPtr[0] == 0,
// non-decreasing row pointers, columns in [0, n), weights >= 1.  The arrays
// are split over threads.  0 with the first violation recorded.
int host_csr_valid(int n, int K, const int32_t *rowPtr, const int32_t *colInd,
                   const int32_t *weights) {
  if (rowPtr == nullptr || rowPtr[0] != 0 || rowPtr[n] < 0) {
    return fail("Error: invalid graph: the row pointers must start at 0 and end at m >= 0.");
  }
  const long long m = rowPtr[n];
  if (m > 0 && (colInd == nullptr || weights == nullptr)) {
    return fail("Error: invalid graph: missing column or weight array.");
  }
  const unsigned workers =
      m < (1 << 20) ? 1u : std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<int> bad(workers, 0);  // 1: row pointers, 2: columns, 3: weights
  auto check = [&](unsigned t) {
    const long long r0 = static_cast<long long>(n) * t / workers;
    const long long r1 = static_cast<long long>(n) * (t + 1) / workers;
    for (long long v = r0; v < r1; ++v) {
      if (rowPtr[v + 1] < rowPtr[v]) { bad[t] = 1; return; }
    }
    const long long e0 = m * t / workers, e1 = m * (t + 1) / workers;
    for (long long e = e0; e < e1; ++e) {
      if (static_cast<uint32_t>(colInd[e]) >= static_cast<uint32_t>(n)) { bad[t] = 2; return; }
    }
    const long long w0 = e0 * K, w1 = e1 * K;
    for (long long i = w0; i < w1; ++i) {
      if (weights[i] < 1) { bad[t] = 3; return; }
    }
  };
  std::vector<std::thread> threads;
  for (unsigned t = 1; t < workers; ++t) threads.emplace_back(check, t);
  check(0);
  for (std::thread &t : threads) t.join();
  for (int b : bad) {
    if (b == 1) return fail("Error: invalid graph: the row pointers decrease.");
    if (b == 2) return fail("Error: invalid graph: a column index is outside [0, n).");
    if (b == 3) return fail("Error: invalid graph: a weight is below 1.");
  }
  return 1;
}

}  // namespace

extern "C" mosp_device_graph *mosp_device_graph_create(void) { return new mosp_device_graph(); }

extern "C" mosp_device_graph *mosp_device_graph_upload(int32_t n, int32_t K, const int32_t *rowPtr,
                                                       const int32_t *colInd,
                                                       const int32_t *weights) {
  clear_error();
  if (n <= 0 || K <= 0 || K > MOSP_MAX_OBJ) {
    fail("Error: invalid graph dimensions.");
    return nullptr;
  }
  if (!host_csr_valid(n, K, rowPtr, colInd, weights)) {
    return nullptr;
  }
  auto *g = new mosp_device_graph();
  auto build = [&]() -> int {
    const int m = rowPtr[n];
    if (!reserve(g, n, m, K)) return 0;
    g->n = n;
    g->m = m;
    g->K = K;
    const size_t wBytes = Carve::bytes<int>(static_cast<size_t>(m) * K);
    int maxOut = 0;
    for (int v = 0; v < n; ++v) {
      maxOut = std::max(maxOut, rowPtr[v + 1] - rowPtr[v]);
    }
    const bool hubs = maxOut > HUB_DEGREE;
    if (!scratch_reserve(g, wBytes + reverse_bytes(n, m, K, hubs))) return 0;
    int *edgeMajor = reinterpret_cast<int *>(g->scratch);
    HIP_TRY(hipMemcpy(g->outRow, rowPtr, sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    if (m > 0) {
      HIP_TRY(hipMemcpy(g->outCol, colInd, sizeof(int) * m, hipMemcpyHostToDevice));
      HIP_TRY(hipMemcpy(edgeMajor, weights, sizeof(int) * static_cast<size_t>(m) * K,
                        hipMemcpyHostToDevice));
    }
    const long long total = static_cast<long long>(m) * K;
    hipLaunchKernelGGL(split_weights_kernel, dim3(grid_for(total)), dim3(BLOCK), 0, 0, edgeMajor,
                       total, g->stride, K, g->outW);
    if (!build_reverse(g, g->scratch + wBytes, m, hubs, 0) || !pair_edges(g, m, 0)) return 0;
    if (!ensure_stats(g) ||
        !graph_max_degree(g->outRow, g->inRow, n, degree_word(g), 0, &g->maxDegree)) {
      return 0;
    }
    HIP_TRY(hipDeviceSynchronize());
    return 1;
  };
  if (!build()) {
    release_arrays(g);
    delete g;
    return nullptr;
  }
  return g;
}

extern "C" void mosp_device_graph_free(mosp_device_graph *g) {
  if (g != nullptr) {
    release_arrays(g);
    delete g;
  }
}

extern "C" int mosp_device_graph_view(const mosp_device_graph *g, int32_t obj, mosp_csr *out,
                                      mosp_csr *in) {
  if (g == nullptr || obj < 0 || obj >= g->K || g->outRow == nullptr) {
    return fail("Error: objective out of range.");
  }
  const size_t off = static_cast<size_t>(obj) * g->stride;
  if (out != nullptr) {
    *out = mosp_csr{g->n, g->m, g->outRow, g->outCol, g->outW + off};
  }
  if (in != nullptr) {
    *in = mosp_csr{g->n, g->m, g->inRow, g->inCol, g->inW + off};
  }
  return 1;
}

extern "C" int32_t mosp_device_graph_objectives(const mosp_device_graph *g) { return g->K; }

extern "C" int mosp_device_graph_weight_stats(mosp_device_graph *g, int64_t *maxWeight,
                                              int64_t *weightSum, void *streamArg) {
  clear_error();
  if (g == nullptr || g->outRow == nullptr) return fail("Error: empty graph.");
  hipStream_t stream = as_stream(streamArg);
  const int K = g->K;
  if (!ensure_stats(g)) return 0;
  hipLaunchKernelGGL(zero_kernel, dim3(1), dim3(BLOCK), 0, stream, reinterpret_cast<int *>(g->stats),
                     4 * MOSP_MAX_OBJ);
  const int blocks = std::max(1, std::min(1024, grid_for(g->m)));
  for (int k = 0; k < K; ++k) {
    hipLaunchKernelGGL(weight_stats_kernel, dim3(blocks), dim3(BLOCK), 0, stream,
                       g->outW + static_cast<size_t>(k) * g->stride, g->m, g->stats + k,
                       g->stats + MOSP_MAX_OBJ + k);
  }
  HIP_TRY(hipGetLastError());
  unsigned long long host[2 * MOSP_MAX_OBJ];
  HIP_TRY(hipMemcpyAsync(host, g->stats, sizeof(host), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipStreamSynchronize(stream));
  for (int k = 0; k < K; ++k) {
    if (weightSum != nullptr) weightSum[k] = static_cast<int64_t>(host[k]);
    if (maxWeight != nullptr) maxWeight[k] = static_cast<int64_t>(host[MOSP_MAX_OBJ + k]);
  }
  return 1;
}

extern "C" int64_t mosp_device_graph_edges(const mosp_device_graph *g) {
  return g == nullptr ? -1 : g->m;
}

extern "C" int mosp_device_graph_download(const mosp_device_graph *g, int32_t *rowPtr,
                                          int32_t *colInd, int32_t *weights) {
  clear_error();
  if (g == nullptr || g->outRow == nullptr) return fail("Error: empty graph.");
  const int n = g->n, m = g->m, K = g->K;
  HIP_TRY(hipMemcpy(rowPtr, g->outRow, sizeof(int) * (n + 1), hipMemcpyDeviceToHost));
  if (m == 0) return 1;
  HIP_TRY(hipMemcpy(colInd, g->outCol, sizeof(int) * m, hipMemcpyDeviceToHost));
  std::vector<int> column(m);
  for (int k = 0; k < K; ++k) {
    HIP_TRY(hipMemcpy(column.data(), g->outW + static_cast<size_t>(k) * g->stride,
                      sizeof(int) * m, hipMemcpyDeviceToHost));
    for (int e = 0; e < m; ++e) {
      weights[static_cast<size_t>(e) * K + k] = column[e];
    }
  }
  return 1;
}

extern "C" int mosp_device_graph_apply(const mosp_device_graph *g, int32_t nIns,
                                       const int32_t *insFrom, const int32_t *insTo,
                                       const int32_t *insW, int32_t nDel, const int32_t *delFrom,
                                       const int32_t *delTo, mosp_device_graph *out,
                                       uint32_t *incMask, void *streamArg) {
  clear_error();
  if (g == nullptr || out == nullptr || g == out || g->outRow == nullptr || nIns < 0 ||
      nDel < 0) {
    return fail("Error: invalid graph apply arguments (need two distinct graphs).");
  }
  const int n = g->n, K = g->K;
  for (int i = 0; i < nIns; ++i) {
    if (insFrom[i] < 0 || insFrom[i] >= n || insTo[i] < 0 || insTo[i] >= n) {
      return fail("Error: insertion " + std::to_string(i) + " has a vertex out of range.");
    }
    for (int k = 0; k < K; ++k) {
      if (insW[static_cast<size_t>(i) * K + k] < 1) {
        return fail("Error: insertion " + std::to_string(i) + " has a weight below 1.");
      }
    }
  }
  for (int i = 0; i < nDel; ++i) {
    if (delFrom[i] < 0 || delFrom[i] >= n || delTo[i] < 0 || delTo[i] >= n) {
      return fail("Error: deletion " + std::to_string(i) + " has a vertex out of range.");
    }
  }
  hipStream_t stream = as_stream(streamArg);
  // Changes sorted by (row, target, file position) and merged into (u, v)
  // pairs; a row's pairs then ordered by their first insertion.
  struct Key {
    long long uv;
    int idx;
  };
  auto sorted_keys = [](const int32_t *from, const int32_t *to, int count) {
    std::vector<Key> keys(count);
    for (int i = 0; i < count; ++i) {
      keys[i] = {(static_cast<long long>(from[i]) << 32) | static_cast<unsigned>(to[i]), i};
    }
    std::sort(keys.begin(), keys.end(), [](const Key &x, const Key &y) {
      return x.uv != y.uv ? x.uv < y.uv : x.idx < y.idx;
    });
    return keys;
  };
  const std::vector<Key> dk = sorted_keys(delFrom, delTo, nDel);
  const std::vector<Key> ik = sorted_keys(insFrom, insTo, nIns);
  struct Pair {
    int u, v, del, insBeg, insCnt, first, last;
  };
  std::vector<Pair> pairs;
  pairs.reserve(static_cast<size_t>(nDel) + nIns);
  constexpr long long NO_KEY = std::numeric_limits<long long>::max();
  for (size_t a = 0, b = 0; a < dk.size() || b < ik.size();) {
    const long long key = std::min(a < dk.size() ? dk[a].uv : NO_KEY,
                                   b < ik.size() ? ik[b].uv : NO_KEY);
    Pair pr{static_cast<int>(key >> 32), static_cast<int>(key & 0xffffffffLL), 0,
            static_cast<int>(b), 0, std::numeric_limits<int>::max(), -1};
    for (; a < dk.size() && dk[a].uv == key; ++a) ++pr.del;
    for (; b < ik.size() && ik[b].uv == key; ++b) ++pr.insCnt;
    if (pr.insCnt > 0) {
      pr.first = ik[pr.insBeg].idx;
      pr.last = ik[b - 1].idx;
    }
    pairs.push_back(pr);
  }
  std::sort(pairs.begin(), pairs.end(), [](const Pair &x, const Pair &y) {
    return x.u != y.u ? x.u < y.u : x.first != y.first ? x.first < y.first : x.v < y.v;
  });
  const int P = static_cast<int>(pairs.size());
  // One host staging array -> one upload: rows, rowPairs, the six pair
  // columns, the sorted insertion indices.
  std::vector<int> rowList;
  std::vector<int> rowStart;
  for (int i = 0; i < P; ++i) {
    if (i == 0 || pairs[i].u != pairs[i - 1].u) {
      rowList.push_back(pairs[i].u);
      rowStart.push_back(i);
    }
  }
  const int R = static_cast<int>(rowList.size());
  rowStart.push_back(P);
  const size_t hostInts = static_cast<size_t>(R) + (R + 1) + 6 * static_cast<size_t>(P) + nIns;
  std::vector<int> host(std::max<size_t>(hostInts, 1));
  int *h = host.data();
  std::copy(rowList.begin(), rowList.end(), h);
  std::copy(rowStart.begin(), rowStart.end(), h + R);
  int *hp = h + R + (R + 1);
  for (int i = 0; i < P; ++i) {
    hp[i] = pairs[i].u;
    hp[P + i] = pairs[i].v;
    hp[2 * P + i] = pairs[i].del;
    hp[3 * P + i] = pairs[i].insBeg;
    hp[4 * P + i] = pairs[i].insCnt;
    hp[5 * P + i] = pairs[i].last;
  }
  for (int i = 0; i < nIns; ++i) {
    hp[6 * P + i] = ik[i].idx;
  }
  const long long mOld = g->m;
  // The reverse build's method follows the graph's degrees before the batch
  // (a batch changes them little).
  const bool hubs = g->maxDegree > HUB_DEGREE;
  if (!reserve(out, n, mOld + nIns, K)) return 0;
  // Temporaries in the destination graph's scratch (no allocation once it
  // is large enough).
  const size_t W = static_cast<size_t>(nIns) * K;
  const size_t bytes = Carve::bytes<int>(hostInts) + Carve::bytes<int>(W) +
                       Carve::bytes<int>(P) * 3 +
                       Carve::bytes<int>(static_cast<size_t>(std::max(mOld, 1LL))) +
                       Carve::bytes<int>(n) + Carve::bytes<int>(n + 1) +
                       Carve::bytes<int>(scan_tiles(n + 1)) + Carve::bytes<unsigned>(nIns) +
                       reverse_bytes(n, mOld + nIns, K, hubs);
  if (!scratch_reserve(out, bytes)) return 0;
  Carve c{out->scratch};
  int *dHost = c.take<int>(hostInts), *dInsW = c.take<int>(W);
  int *pairDead = c.take<int>(P), *pairApp = c.take<int>(P), *pairAppPos = c.take<int>(P);
  int *slot = c.take<int>(static_cast<size_t>(std::max(mOld, 1LL)));
  int *rowTag = c.take<int>(n), *newDeg = c.take<int>(n + 1);
  int *sums = c.take<int>(scan_tiles(n + 1));
  unsigned *dMask = c.take<unsigned>(nIns);
  char *reverseAt = c.p;
  const int *dRows = dHost, *dRowPairs = dHost + R;
  const int *pU = dHost + R + (R + 1), *pV = pU + P, *pDel = pU + 2 * P, *pInsBeg = pU + 3 * P;
  const int *pInsCnt = pU + 4 * P, *pLast = pU + 5 * P, *insSorted = pU + 6 * P;
  if (hostInts > 0) {
    HIP_TRY(hipMemcpyAsync(dHost, host.data(), sizeof(int) * hostInts, hipMemcpyHostToDevice,
                           stream));
  }
  if (W > 0) {
    HIP_TRY(hipMemcpyAsync(dInsW, insW, sizeof(int) * W, hipMemcpyHostToDevice, stream));
  }
  hipLaunchKernelGGL(apply_init_kernel, dim3(grid_for(std::max<long long>(n + 1, mOld))),
                     dim3(BLOCK), 0, stream, n, mOld, g->outRow, newDeg, rowTag, slot);
  if (P > 0) {
    hipLaunchKernelGGL(apply_pairs_kernel, dim3(P), dim3(64), 0, stream, P, K, pU, pV, pDel,
                       pInsBeg, pInsCnt, pLast, insSorted, g->outRow, g->outCol, g->outW,
                       g->stride, dInsW, slot, pairDead, pairApp, dMask);
    hipLaunchKernelGGL(apply_rows_kernel, dim3(R), dim3(64), 0, stream, R, dRows, dRowPairs,
                       g->outRow, pairDead, pairApp, pairAppPos, newDeg, rowTag);
  }
  if (!exclusive_scan(newDeg, out->outRow, n + 1, sums, stream)) return 0;
  out->n = n;
  out->K = K;
  hipLaunchKernelGGL(apply_copy_kernel, dim3(sweep_grid(mOld)), dim3(SWEEP_THREADS), 0, stream,
                     n, K, g->outRow, g->outCol, g->outW, g->stride, rowTag, out->outRow,
                     out->outCol, out->outW, out->stride, sweep_chunk(mOld));
  if (R > 0) {
    hipLaunchKernelGGL(apply_compact_kernel, dim3(R), dim3(COMPACT_THREADS), 0, stream, R, K,
                       dRows, dRowPairs, g->outRow, g->outCol, g->outW, g->stride, slot, pV,
                       pLast, pairApp, pairAppPos, dInsW, out->outRow, out->outCol, out->outW,
                       out->stride);
  }
  HIP_TRY(hipGetLastError());
  if (!build_reverse(out, reverseAt, mOld + nIns, hubs, stream) ||
      !pair_edges(out, mOld + nIns, stream)) {
    return 0;
  }
  if (!ensure_stats(out) ||
      !queue_max_degree(out->outRow, out->inRow, n, degree_word(out), stream)) {
    return 0;
  }
  // Results to the host only after every kernel of the rebuild.
  int mNew = 0;
  unsigned maxDegree = 0;
  HIP_TRY(hipMemcpyAsync(&mNew, out->outRow + n, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIP_TRY(hipMemcpyAsync(&maxDegree, degree_word(out), sizeof(unsigned), hipMemcpyDeviceToHost,
                         stream));
  if (nIns > 0 && incMask != nullptr) {
    HIP_TRY(hipMemcpyAsync(incMask, dMask, sizeof(unsigned) * nIns, hipMemcpyDeviceToHost,
                           stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  out->m = mNew;
  out->maxDegree = static_cast<int>(maxDegree);
  HIP_TRY(hipStreamSynchronize(stream));
  return 1;
}

namespace mosp {
namespace detail {

int graph_records4(const mosp_device_graph *g, hipStream_t stream) {
  if (g->rec4Valid) {
    return 1;
  }
  if (g->outEdge4 == nullptr) {
    HIP_TRY(hipMalloc(&g->outEdge4, sizeof(int4) * static_cast<size_t>(g->stride) * g->K));
  }
  const long long m = std::max(g->m, 1);
  const int blocks = static_cast<int>(std::min<long long>((m + BLOCK - 1) / BLOCK, 8192));
  hipLaunchKernelGGL(records4_kernel, dim3(blocks), dim3(BLOCK), 0, stream, g->outRow, g->n,
                     g->outCol, g->outW, g->stride, g->K, g->outEdge4);
  HIP_TRY(hipGetLastError());
  g->rec4Valid = true;
  return 1;
}

void graph_edges(const mosp_device_graph *g, int k, const int2 **out, const int2 **in) {
  const size_t off = static_cast<size_t>(k) * g->stride;
  *out = g->outEdge + off;
  *in = g->inEdge + off;
}

int graph_max_degree(const int *outRow, const int *inRow, int n, unsigned *scratch,
                     hipStream_t stream, int *result) {
  unsigned host = 0;
  if (n > 0) {
    if (!queue_max_degree(outRow, inRow, n, scratch, stream)) return 0;
    HIP_TRY(hipMemcpyAsync(&host, scratch, sizeof(unsigned), hipMemcpyDeviceToHost, stream));
    HIP_TRY(hipStreamSynchronize(stream));
  }
  *result = static_cast<int>(host);
  return 1;
}

}  // namespace detail
}  // namespace mosp
