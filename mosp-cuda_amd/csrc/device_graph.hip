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
// the original first occurrence.  The host only groups the (small) batch by
// row; one wave per changed row replays its changes with ballot searches, a
// scan gives the new row pointers, and the rows are copied or compacted.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
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

// In-degrees of the m = *edges arcs (edge count read on the device, so a
// graph rebuilt on the device needs no host round trip).
__global__ void in_degree_kernel(const int *colInd, const int *edges, int *degree) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < *edges) {
    atomicAdd(&degree[colInd[e]], 1);
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

__global__ void fill_reverse_kernel(int n, long long stride, int K, const int *rowPtr,
                                    const int *colInd, const int *outW, int *cursor, int *inCol,
                                    int *inW) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) {
    return;
  }
  for (int e = rowPtr[u]; e < rowPtr[u + 1]; ++e) {
    const int pos = atomicAdd(&cursor[colInd[e]], 1);
    inCol[pos] = u;
    for (int k = 0; k < K; ++k) {
      inW[k * stride + pos] = outW[k * stride + e];
    }
  }
}

__global__ void zero_kernel(int *a, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) {
    a[i] = 0;
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

// Degrees, row tags and every original entry's slot (KEEP), before the
// row replay: the replay's only writer per row is then lane 0.
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

// First i in [begin, end) with pred(i), searched 64 at a time by one wave;
// wave-uniform result, -1 if none.
template <typename Pred>
__device__ __forceinline__ int wave_find_first(int begin, int end, Pred pred) {
  const int lane = static_cast<int>(__lane_id());
  for (int base = begin; base < end; base += 64) {
    const int i = base + lane;
    const unsigned long long hits = __ballot(i < end && pred(i));
    if (hits != 0) {
      return base + __ffsll(static_cast<long long>(hits)) - 1;
    }
  }
  return -1;
}

// Lane 0's store, complete and visible before any lane's next load: an
// agent-scope release + acquire (waits for the store, writes back and
// invalidates the caches the next loads could hit).
__device__ __forceinline__ void publish_lane0_store() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "agent");
}

// One wave per changed row: replay the row's deletions, then its insertions,
// in file order.  slot[e] of an original entry: KEEP (set by the init
// kernel), DEAD or the insertion whose weights it takes; appended entries
// live at appV/appSlot[insOff[r]..].  Only lane 0 writes.
__global__ void apply_rows_kernel(int R, const int *rows, const int *oldRow, const int *oldCol,
                                  const int *delOff, const int *delV, const int *insOff,
                                  const int *insV, const int *insIdx, int *slot, int *appV,
                                  int *appSlot, int *appCnt, int *newDeg, int *rowTag) {
  const int r = static_cast<int>(blockIdx.x);  // one wave per workgroup
  if (r >= R) {
    return;
  }
  const int lane = static_cast<int>(__lane_id());
  const int u = rows[r];
  const int ob = oldRow[u], oe = oldRow[u + 1];
  for (int j = delOff[r]; j < delOff[r + 1]; ++j) {
    const int v = delV[j];
    const int e = wave_find_first(ob, oe, [&](int x) { return oldCol[x] == v && slot[x] != DEAD; });
    if (e >= 0 && lane == 0) {
      slot[e] = DEAD;
    }
    publish_lane0_store();
  }
  const int ab = insOff[r];
  int count = 0;
  for (int j = insOff[r]; j < insOff[r + 1]; ++j) {
    const int v = insV[j], i = insIdx[j];
    const int e = wave_find_first(ob, oe, [&](int x) { return oldCol[x] == v && slot[x] != DEAD; });
    if (e >= 0) {
      if (lane == 0) slot[e] = i;
    } else {
      const int a = wave_find_first(ab, ab + count, [&](int x) { return appV[x] == v; });
      if (a >= 0) {
        if (lane == 0) appSlot[a] = i;
      } else {
        if (lane == 0) {
          appV[ab + count] = v;
          appSlot[ab + count] = i;
        }
        ++count;
      }
    }
    publish_lane0_store();
  }
  int alive = 0;
  for (int base = ob; base < oe; base += 64) {
    const int e = base + lane;
    alive += __popcll(__ballot(e < oe && slot[e] != DEAD));
  }
  if (lane == 0) {
    newDeg[u] = alive + count;
    appCnt[r] = count;
    rowTag[u] = r;
  }
}

// Unchanged rows: 16 lanes per row copy columns and the K weight columns.
__global__ void apply_copy_kernel(int n, int K, const int *oldRow, const int *oldCol,
                                  const int *oldW, long long oldStride, const int *rowTag,
                                  const int *newRow, int *newCol, int *newW, long long newStride) {
  const long long gid = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int u = static_cast<int>(gid / 16);
  const int sub = static_cast<int>(gid % 16);
  if (u >= n || rowTag[u] >= 0) {
    return;
  }
  const int ob = oldRow[u], deg = oldRow[u + 1] - ob, nb = newRow[u];
  for (int x = sub; x < deg; x += 16) {
    newCol[nb + x] = oldCol[ob + x];
    for (int k = 0; k < K; ++k) {
      newW[k * newStride + nb + x] = oldW[k * oldStride + ob + x];
    }
  }
}

// Changed rows: one wave compacts the alive original entries in order, then
// writes the appended ones.
__global__ void apply_compact_kernel(int R, int K, const int *rows, const int *oldRow,
                                     const int *oldCol, const int *oldW, long long oldStride,
                                     const int *slot, const int *insOff, const int *appV,
                                     const int *appSlot, const int *appCnt, const int *insW,
                                     const int *newRow, int *newCol, int *newW,
                                     long long newStride) {
  const int r = static_cast<int>(blockIdx.x);  // one wave per workgroup
  if (r >= R) {
    return;
  }
  const int lane = static_cast<int>(__lane_id());
  const int u = rows[r];
  const int ob = oldRow[u], oe = oldRow[u + 1];
  int out = newRow[u];
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int base = ob; base < oe; base += 64) {
    const int e = base + lane;
    const int s = e < oe ? slot[e] : DEAD;
    const unsigned long long alive = __ballot(s != DEAD);
    if (s != DEAD) {
      const int pos = out + __popcll(alive & below);
      newCol[pos] = oldCol[e];
      for (int k = 0; k < K; ++k) {
        newW[k * newStride + pos] =
            s == KEEP ? oldW[k * oldStride + e] : insW[static_cast<size_t>(s) * K + k];
      }
    }
    out += __popcll(alive);
  }
  const int ab = insOff[r];
  for (int a = lane; a < appCnt[r]; a += 64) {
    const int pos = out + a;
    const int s = appSlot[ab + a];
    newCol[pos] = appV[ab + a];
    for (int k = 0; k < K; ++k) {
      newW[k * newStride + pos] = insW[static_cast<size_t>(s) * K + k];
    }
  }
}

// One wave per insertion: bit k set if the final (u,v) weight exceeds the
// original first occurrence's.
__global__ void apply_mask_kernel(int nIns, int K, const int *insFrom, const int *insTo,
                                  const int *oldRow, const int *oldCol, const int *oldW,
                                  long long oldStride, const int *newRow, const int *newCol,
                                  const int *newW, long long newStride, unsigned *mask) {
  const int i = static_cast<int>(blockIdx.x);  // one wave per workgroup
  if (i >= nIns) {
    return;
  }
  const int u = insFrom[i], v = insTo[i];
  const int eo = wave_find_first(oldRow[u], oldRow[u + 1], [&](int x) { return oldCol[x] == v; });
  const int en = wave_find_first(newRow[u], newRow[u + 1], [&](int x) { return newCol[x] == v; });
  if (__lane_id() != 0) {
    return;
  }
  unsigned bits = 0;
  if (eo >= 0 && en >= 0) {
    for (int k = 0; k < K; ++k) {
      if (newW[k * newStride + en] > oldW[k * oldStride + eo]) {
        bits |= 1u << k;
      }
    }
  }
  mask[i] = bits;
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
  g->outRow = g->outCol = g->outW = g->inRow = g->inCol = g->inW = nullptr;
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
  g->stride = stride;
  g->rowCap = n + 1;
  g->K = K;
  return 1;
}

size_t reverse_bytes(int n) {
  return Carve::bytes<int>(n + 1) + Carve::bytes<int>(scan_tiles(n + 1));
}

// In-CSR (rows, columns, K weight columns) from the out-CSR, on `stream`;
// temporaries at `at` (reverse_bytes(n)); mBound >= the edge count, which is
// read from outRow[n] on the device.  Only kernels on one stream: each sees
// its predecessors' writes.
int build_reverse(mosp_device_graph *g, char *at, long long mBound, hipStream_t stream) {
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
  hipLaunchKernelGGL(fill_reverse_kernel, dim3(grid_for(n)), dim3(BLOCK), 0, stream, n,
                     g->stride, K, g->outRow, g->outCol, g->outW, cursor, g->inCol, g->inW);
  HIP_TRY(hipGetLastError());
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
  auto *g = new mosp_device_graph();
  auto build = [&]() -> int {
    const int m = rowPtr[n];
    if (!reserve(g, n, m, K)) return 0;
    g->n = n;
    g->m = m;
    g->K = K;
    const size_t wBytes = Carve::bytes<int>(static_cast<size_t>(m) * K);
    if (!scratch_reserve(g, wBytes + reverse_bytes(n))) return 0;
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
    if (!build_reverse(g, g->scratch + wBytes, m, 0)) return 0;
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
  if (g->stats == nullptr) {
    HIP_TRY(hipMalloc(&g->stats, sizeof(unsigned long long) * 2 * MOSP_MAX_OBJ));
  }
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
  // Group the changes by row, file order inside a row.
  std::vector<int> dOrder(nDel), iOrder(nIns);
  std::iota(dOrder.begin(), dOrder.end(), 0);
  std::iota(iOrder.begin(), iOrder.end(), 0);
  std::stable_sort(dOrder.begin(), dOrder.end(),
                   [&](int a, int b) { return delFrom[a] < delFrom[b]; });
  std::stable_sort(iOrder.begin(), iOrder.end(),
                   [&](int a, int b) { return insFrom[a] < insFrom[b]; });
  std::vector<int> rows, delOff{0}, insOff{0}, delV, insV, insIdx;
  rows.reserve(nDel + nIns);
  size_t a = 0, b = 0;
  while (a < dOrder.size() || b < iOrder.size()) {
    const int ua = a < dOrder.size() ? delFrom[dOrder[a]] : n;
    const int ub = b < iOrder.size() ? insFrom[iOrder[b]] : n;
    const int u = std::min(ua, ub);
    while (a < dOrder.size() && delFrom[dOrder[a]] == u) delV.push_back(delTo[dOrder[a++]]);
    while (b < iOrder.size() && insFrom[iOrder[b]] == u) {
      insV.push_back(insTo[iOrder[b]]);
      insIdx.push_back(iOrder[b++]);
    }
    rows.push_back(u);
    delOff.push_back(static_cast<int>(delV.size()));
    insOff.push_back(static_cast<int>(insV.size()));
  }
  const int R = static_cast<int>(rows.size());
  const long long mOld = g->m;
  if (!reserve(out, n, mOld + nIns, K)) return 0;
  // Temporaries in the destination graph's scratch (no allocation once it
  // is large enough).
  const size_t W = static_cast<size_t>(nIns) * K;
  const size_t bytes = Carve::bytes<int>(R) * 2 + Carve::bytes<int>(R + 1) * 2 +
                       Carve::bytes<int>(delV.size()) + Carve::bytes<int>(insV.size()) * 2 +
                       Carve::bytes<int>(W) + Carve::bytes<int>(nIns) * 4 +
                       Carve::bytes<int>(static_cast<size_t>(std::max(mOld, 1LL))) +
                       Carve::bytes<int>(n) + Carve::bytes<int>(n + 1) +
                       Carve::bytes<int>(scan_tiles(n + 1)) + Carve::bytes<unsigned>(nIns) +
                       reverse_bytes(n);
  if (!scratch_reserve(out, bytes)) return 0;
  Carve c{out->scratch};
  int *dRows = c.take<int>(R), *dAppCnt = c.take<int>(R);
  int *dDelOff = c.take<int>(R + 1), *dInsOff = c.take<int>(R + 1);
  int *dDelV = c.take<int>(delV.size()), *dInsV = c.take<int>(insV.size());
  int *dInsIdx = c.take<int>(insIdx.size()), *dInsW = c.take<int>(W);
  int *dInsFrom = c.take<int>(nIns), *dInsTo = c.take<int>(nIns);
  int *appV = c.take<int>(nIns), *appSlot = c.take<int>(nIns);
  int *slot = c.take<int>(static_cast<size_t>(std::max(mOld, 1LL)));
  int *rowTag = c.take<int>(n), *newDeg = c.take<int>(n + 1);
  int *sums = c.take<int>(scan_tiles(n + 1));
  unsigned *dMask = c.take<unsigned>(nIns);
  char *reverseAt = c.p;
  auto up = [&](int *dst, const int *src, size_t count) -> int {
    if (count > 0) {
      HIP_TRY(hipMemcpyAsync(dst, src, sizeof(int) * count, hipMemcpyHostToDevice, stream));
    }
    return 1;
  };
  if (!up(dRows, rows.data(), R) || !up(dDelOff, delOff.data(), R + 1) ||
      !up(dInsOff, insOff.data(), R + 1) || !up(dDelV, delV.data(), delV.size()) ||
      !up(dInsV, insV.data(), insV.size()) || !up(dInsIdx, insIdx.data(), insIdx.size()) ||
      !up(dInsW, insW, W) || !up(dInsFrom, insFrom, nIns) || !up(dInsTo, insTo, nIns)) {
    return 0;
  }
  hipLaunchKernelGGL(apply_init_kernel, dim3(grid_for(std::max<long long>(n + 1, mOld))),
                     dim3(BLOCK), 0, stream, n, mOld, g->outRow, newDeg, rowTag, slot);
  if (R > 0) {
    hipLaunchKernelGGL(apply_rows_kernel, dim3(R), dim3(64), 0, stream, R,
                       dRows, g->outRow, g->outCol, dDelOff, dDelV, dInsOff, dInsV, dInsIdx,
                       slot, appV, appSlot, dAppCnt, newDeg, rowTag);
  }
  if (!exclusive_scan(newDeg, out->outRow, n + 1, sums, stream)) return 0;
  out->n = n;
  out->K = K;
  hipLaunchKernelGGL(apply_copy_kernel, dim3(grid_for(static_cast<long long>(n) * 16)),
                     dim3(BLOCK), 0, stream, n, K, g->outRow, g->outCol, g->outW, g->stride,
                     rowTag, out->outRow, out->outCol, out->outW, out->stride);
  if (R > 0) {
    hipLaunchKernelGGL(apply_compact_kernel, dim3(R), dim3(64), 0, stream, R,
                       K, dRows, g->outRow, g->outCol, g->outW, g->stride, slot, dInsOff, appV,
                       appSlot, dAppCnt, dInsW, out->outRow, out->outCol, out->outW,
                       out->stride);
  }
  if (nIns > 0) {
    hipLaunchKernelGGL(apply_mask_kernel, dim3(nIns), dim3(64), 0, stream, nIns, K, dInsFrom, dInsTo, g->outRow, g->outCol,
                       g->outW, g->stride, out->outRow, out->outCol, out->outW, out->stride,
                       dMask);
  }
  HIP_TRY(hipGetLastError());
  if (!build_reverse(out, reverseAt, mOld + nIns, stream)) return 0;
  // Results to the host only after every kernel of the rebuild.
  int mNew = 0;
  HIP_TRY(hipMemcpyAsync(&mNew, out->outRow + n, sizeof(int), hipMemcpyDeviceToHost, stream));
  if (nIns > 0 && incMask != nullptr) {
    HIP_TRY(hipMemcpyAsync(incMask, dMask, sizeof(unsigned) * nIns, hipMemcpyDeviceToHost,
                           stream));
  }
  HIP_TRY(hipStreamSynchronize(stream));
  out->m = mNew;
  HIP_TRY(hipStreamSynchronize(stream));
  return 1;
}
