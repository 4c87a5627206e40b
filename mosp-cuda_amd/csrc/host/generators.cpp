// generators.cpp -- seeded synthetic inputs of the benchmark workloads.
//
// An explicit PRNG (xoshiro256** seeded by SplitMix64, unbiased bounded
// integers by Lemire's multiply-and-reject method) makes every instance
// bit-identical across hosts and compilers, unlike std::uniform_int_
// distribution.  Workloads:
//   * road-like grid: W x H 4-neighbour grid, a uniformly random spanning
//     tree kept (Kruskal on a random edge order), a fraction of the other
//     grid edges dropped, random diagonal shortcuts added; both directions
//     stored, K iid weights per arc;
//   * R-MAT (a, b, c, d) with self-loops and duplicate arcs removed, drawn
//     in parallel over fixed chunks with their own streams;
//   * change batches: insertions join random pairs u != v, deletions pick
//     distinct existing arcs; optionally connectivity-safe (deletions that
//     cut a reachable vertex off the source are dropped and replaced, the
//     rule of the reference's src/changeGenerator.cu:78-116).
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <unordered_set>
#include <vector>

#include "csrGraph.h"
#include "mosp_hip.h"

namespace {

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto &w : s) {  // SplitMix64
      x += 0x9e3779b97f4a7c15ULL;
      uint64_t z = x;
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      w = z ^ (z >> 31);
    }
  }
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {  // xoshiro256**
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  // Uniform in [0, range) without bias (Lemire 2019).
  uint64_t below(uint64_t range) {
    if (range <= 1) return 0;
    unsigned __int128 m = static_cast<unsigned __int128>(next()) * range;
    uint64_t low = static_cast<uint64_t>(m);
    if (low < range) {
      const uint64_t threshold = (0 - range) % range;
      while (low < threshold) {
        m = static_cast<unsigned __int128>(next()) * range;
        low = static_cast<uint64_t>(m);
      }
    }
    return static_cast<uint64_t>(m >> 64);
  }
  int64_t between(int64_t lo, int64_t hi) {  // inclusive
    return lo + static_cast<int64_t>(below(static_cast<uint64_t>(hi - lo) + 1));
  }
  double unit() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
};

struct Dsu {
  std::vector<int> p;
  explicit Dsu(int n) : p(n) { std::iota(p.begin(), p.end(), 0); }
  int find(int x) {
    while (p[x] != x) {
      p[x] = p[p[x]];
      x = p[x];
    }
    return x;
  }
  bool unite(int a, int b) {
    a = find(a);
    b = find(b);
    if (a == b) return false;
    p[a] = b;
    return true;
  }
};

// Arc list -> CSR (rows in arc order of generation, stable).
void to_csr(int n, int K, const std::vector<int> &us, const std::vector<int> &vs,
            const std::vector<int> &ws, CsrGraph &g) {
  g = CsrGraph();
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  g.rowPtr.assign(n + 1, 0);
  for (int u : us) ++g.rowPtr[u + 1];
  for (int i = 0; i < n; ++i) g.rowPtr[i + 1] += g.rowPtr[i];
  const size_t m = us.size();
  g.colInd.resize(m);
  g.weights.resize(m * K);
  std::vector<int> fill(g.rowPtr.begin(), g.rowPtr.end() - 1);
  for (size_t a = 0; a < m; ++a) {
    const size_t pos = static_cast<size_t>(fill[us[a]]++);
    g.colInd[pos] = vs[a];
    std::memcpy(&g.weights[pos * K], &ws[a * K], sizeof(int) * K);
  }
}

void road_grid(int width, int height, int K, int wmin, int wmax, double drop,
               double diag, uint64_t seed, CsrGraph &g) {
  Rng rng(seed);
  const int n = width * height;
  // Undirected grid edges: (a, b) with b the right or lower neighbour.
  std::vector<std::pair<int, int>> edges;
  edges.reserve(2 * static_cast<size_t>(n));
  for (int y = 0; y < height; ++y) {
    for (int x = 0; x < width; ++x) {
      const int a = y * width + x;
      if (x + 1 < width) edges.push_back({a, a + 1});
      if (y + 1 < height) edges.push_back({a, a + width});
    }
  }
  // Random order (Fisher-Yates); the first edges joining two components form
  // a uniformly ordered Kruskal spanning tree, kept to guarantee connectivity.
  for (size_t i = edges.size(); i > 1; --i) {
    std::swap(edges[i - 1], edges[rng.below(i)]);
  }
  Dsu dsu(n);
  std::vector<char> inTree(edges.size(), 0);
  std::vector<size_t> others;
  for (size_t i = 0; i < edges.size(); ++i) {
    if (dsu.unite(edges[i].first, edges[i].second)) {
      inTree[i] = 1;
    } else {
      others.push_back(i);
    }
  }
  size_t dropCount = static_cast<size_t>(drop * static_cast<double>(edges.size()) + 0.5);
  dropCount = std::min(dropCount, others.size());
  // `others` is already in random order: drop its first dropCount edges.
  std::vector<char> keep(edges.size(), 1);
  for (size_t i = 0; i < dropCount; ++i) keep[others[i]] = 0;
  std::vector<int> us, vs, ws;
  us.reserve(2 * edges.size());
  vs.reserve(2 * edges.size());
  auto add = [&](int a, int b) {
    us.push_back(a);
    vs.push_back(b);
    for (int k = 0; k < K; ++k) ws.push_back(static_cast<int>(rng.between(wmin, wmax)));
  };
  for (size_t i = 0; i < edges.size(); ++i) {
    if (!keep[i]) continue;
    add(edges[i].first, edges[i].second);
    add(edges[i].second, edges[i].first);
  }
  // Diagonal shortcuts: distinct (cell, orientation) pairs.
  if (width > 1 && height > 1 && diag > 0) {
    const uint64_t cells = static_cast<uint64_t>(width - 1) * (height - 1);
    size_t want = static_cast<size_t>(diag * n + 0.5);
    want = std::min<size_t>(want, cells * 2);
    std::unordered_set<uint64_t> used;
    used.reserve(want * 2);
    while (used.size() < want) {
      const uint64_t pick = rng.below(cells * 2);
      if (!used.insert(pick).second) continue;
      const uint64_t cell = pick >> 1;
      const int x = static_cast<int>(cell % (width - 1)), y = static_cast<int>(cell / (width - 1));
      int a, b;
      if (pick & 1) {
        a = y * width + x;
        b = (y + 1) * width + x + 1;
      } else {
        a = y * width + x + 1;
        b = (y + 1) * width + x;
      }
      add(a, b);
      add(b, a);
    }
  }
  to_csr(n, K, us, vs, ws, g);
}

// R-MAT draws in fixed chunks, each with its own stream (seeded from the
// seed and the chunk index), so the instance does not depend on how many
// threads draw it; arcs are then bucketed by the top bits of their source,
// each bucket sorted and de-duplicated on its own (duplicates share a
// bucket), and each bucket's weights drawn from its own stream.
void rmat(int scale, int edgeFactor, double a, double b, double c, int K, int wmin, int wmax,
          uint64_t seed, CsrGraph &g) {
  const int n = 1 << scale;
  const uint64_t m = static_cast<uint64_t>(edgeFactor) << scale;
  constexpr uint64_t NONE = ~0ULL;
  constexpr int CHUNKS = 256;
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  auto parallel = [&](int count, auto &&f) {
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; ++t) {
      pool.emplace_back([&, t] {
        for (int i = static_cast<int>(t); i < count; i += static_cast<int>(threads)) f(i);
      });
    }
    for (int i = 0; i < count; i += static_cast<int>(threads)) f(i);
    for (std::thread &th : pool) th.join();
  };
  std::vector<uint64_t> draws(m);
  parallel(CHUNKS, [&](int ch) {
    Rng rng(seed ^ (0x5851f42d4c957f2dULL * static_cast<uint64_t>(ch + 1)));
    const uint64_t lo = m * ch / CHUNKS, hi = m * (ch + 1) / CHUNKS;
    for (uint64_t i = lo; i < hi; ++i) {
      uint64_t u = 0, v = 0;
      for (int bit = 0; bit < scale; ++bit) {
        const double r = rng.unit();
        const uint64_t du = r >= a + b ? 1 : 0;
        const uint64_t dv = (r >= a && r < a + b) || r >= a + b + c ? 1 : 0;
        u = (u << 1) | du;
        v = (v << 1) | dv;
      }
      draws[i] = u != v ? (u << 32) | v : NONE;
    }
  });
  // Buckets by the source's top bits: count per (chunk, bucket), scan, scatter.
  const int bucketBits = std::min(scale, 8);
  const int buckets = 1 << bucketBits;
  const int shift = 32 + scale - bucketBits;
  std::vector<uint64_t> counts(static_cast<size_t>(CHUNKS) * buckets, 0);
  parallel(CHUNKS, [&](int ch) {
    const uint64_t lo = m * ch / CHUNKS, hi = m * (ch + 1) / CHUNKS;
    uint64_t *cnt = &counts[static_cast<size_t>(ch) * buckets];
    for (uint64_t i = lo; i < hi; ++i) {
      if (draws[i] != NONE) ++cnt[draws[i] >> shift];
    }
  });
  std::vector<uint64_t> bucketStart(buckets + 1, 0), at(counts.size());
  {
    uint64_t total = 0;
    for (int bk = 0; bk < buckets; ++bk) {
      bucketStart[bk] = total;
      for (int ch = 0; ch < CHUNKS; ++ch) {
        at[static_cast<size_t>(ch) * buckets + bk] = total;
        total += counts[static_cast<size_t>(ch) * buckets + bk];
      }
    }
    bucketStart[buckets] = total;
  }
  std::vector<uint64_t> arcs(bucketStart[buckets]);
  parallel(CHUNKS, [&](int ch) {
    const uint64_t lo = m * ch / CHUNKS, hi = m * (ch + 1) / CHUNKS;
    uint64_t *pos = &at[static_cast<size_t>(ch) * buckets];
    for (uint64_t i = lo; i < hi; ++i) {
      if (draws[i] != NONE) arcs[pos[draws[i] >> shift]++] = draws[i];
    }
  });
  std::vector<uint64_t>().swap(draws);
  std::vector<uint64_t> kept(buckets, 0);
  parallel(buckets, [&](int bk) {
    auto first = arcs.begin() + static_cast<std::ptrdiff_t>(bucketStart[bk]);
    auto last = arcs.begin() + static_cast<std::ptrdiff_t>(bucketStart[bk + 1]);
    std::sort(first, last);
    kept[bk] = static_cast<uint64_t>(std::unique(first, last) - first);
  });
  std::vector<uint64_t> out(buckets + 1, 0);
  for (int bk = 0; bk < buckets; ++bk) out[bk + 1] = out[bk] + kept[bk];
  const size_t arcsKept = out[buckets];
  g = CsrGraph();
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  g.rowPtr.assign(static_cast<size_t>(n) + 1, 0);
  g.colInd.resize(arcsKept);
  g.weights.resize(arcsKept * K);
  parallel(buckets, [&](int bk) {
    Rng rng(seed ^ (0x9e3779b97f4a7c15ULL * static_cast<uint64_t>(bk + 1)) ^ 0xd1b54a32d192ed03ULL);
    const uint64_t *src = &arcs[bucketStart[bk]];
    for (uint64_t j = 0; j < kept[bk]; ++j) {
      const size_t e = out[bk] + j;
      g.colInd[e] = static_cast<int>(src[j] & 0xffffffffu);
      for (int k = 0; k < K; ++k) {
        g.weights[e * K + k] = static_cast<int>(rng.between(wmin, wmax));
      }
    }
  });
  // Row pointers: arcs are sorted by source across buckets.
  for (int bk = 0; bk < buckets; ++bk) {
    const uint64_t *src = &arcs[bucketStart[bk]];
    for (uint64_t j = 0; j < kept[bk]; ++j) ++g.rowPtr[static_cast<size_t>(src[j] >> 32) + 1];
  }
  for (int v = 0; v < n; ++v) g.rowPtr[v + 1] += g.rowPtr[v];
}

std::vector<char> reachable(const CsrGraph &g, int source) {
  std::vector<char> seen(g.numberOfNodes, 0);
  std::vector<int> queue{source};
  seen[source] = 1;
  for (size_t i = 0; i < queue.size(); ++i) {
    const int u = queue[i];
    for (int e = g.rowPtr[u]; e < g.rowPtr[u + 1]; ++e) {
      const int w = g.colInd[e];
      if (!seen[w]) {
        seen[w] = 1;
        queue.push_back(w);
      }
    }
  }
  return seen;
}

void make_batch(const CsrGraph &g, int changes, double insPct, int safe, int wmin, int wmax,
                uint64_t seed, int source, ChangeBatch &b) {
  Rng rng(seed);
  const int n = g.numberOfNodes, K = g.numberOfObjectives;
  const int m = g.numberOfEdges();
  const int nIns = static_cast<int>(changes * insPct / 100.0 + 0.5);
  const int nDel = std::min(changes - nIns, m);
  b = ChangeBatch();
  b.numberOfObjectives = K;
  while (b.numberOfInserts() < nIns && n > 1) {
    const int u = static_cast<int>(rng.below(n)), v = static_cast<int>(rng.below(n));
    if (u == v) continue;
    b.insertFrom.push_back(u);
    b.insertTo.push_back(v);
    for (int k = 0; k < K; ++k) b.insertWeights.push_back(static_cast<int>(rng.between(wmin, wmax)));
  }
  // Source vertex of every arc (for drawing arcs uniformly).
  std::vector<int> src(m);
  for (int u = 0; u < n; ++u) {
    for (int e = g.rowPtr[u]; e < g.rowPtr[u + 1]; ++e) src[e] = u;
  }
  std::vector<char> taken(m, 0);
  std::vector<int> chosen;
  auto draw = [&](int count) {
    while (count > 0 && static_cast<int>(chosen.size()) < m) {
      const int e = static_cast<int>(rng.below(m));
      if (taken[e]) continue;
      taken[e] = 1;
      chosen.push_back(e);
      --count;
    }
  };
  draw(nDel);
  if (safe) {
    const std::vector<char> before = reachable(g, source);
    for (int round = 0; round < 64; ++round) {
      ChangeBatch probe = b;
      for (int e : chosen) {
        probe.deleteFrom.push_back(src[e]);
        probe.deleteTo.push_back(g.colInd[e]);
      }
      CsrGraph upd;
      applyChangeBatch(g, probe, upd);
      const std::vector<char> after = reachable(upd, source);
      std::vector<char> lost(n, 0);
      bool any = false;
      for (int v = 0; v < n; ++v) {
        if (before[v] && !after[v]) {
          lost[v] = 1;
          any = true;
        }
      }
      if (any) {
        std::vector<int> kept;
        for (int e : chosen) {
          if (!lost[src[e]] && !lost[g.colInd[e]]) kept.push_back(e);
        }
        if (kept.size() == chosen.size()) break;  // cannot happen
        chosen.swap(kept);
      }
      if (static_cast<int>(chosen.size()) >= nDel) {
        if (!any) break;
        continue;
      }
      draw(nDel - static_cast<int>(chosen.size()));  // top up, then re-check
    }
  }
  for (int e : chosen) {
    b.deleteFrom.push_back(src[e]);
    b.deleteTo.push_back(g.colInd[e]);
  }
}

int export_graph(const CsrGraph &g, int32_t *n, int64_t *m, int32_t *rowPtr, int32_t *colInd,
                 int32_t *weights, int64_t capacity) {
  *n = g.numberOfNodes;
  *m = g.numberOfEdges();
  if (rowPtr == nullptr) return 1;  // size query
  if (*m > capacity) return 0;
  std::copy(g.rowPtr.begin(), g.rowPtr.end(), rowPtr);
  std::copy(g.colInd.begin(), g.colInd.end(), colInd);
  std::copy(g.weights.begin(), g.weights.end(), weights);
  return 1;
}

}  // namespace

// Road-like grid; call with row_ptr == NULL to get n and m, then again with
// buffers (capacity = edges they hold).
extern "C" int mosp_gen_road_grid(int32_t width, int32_t height, int32_t K, int32_t wmin,
                                  int32_t wmax, double drop, double diag, uint64_t seed,
                                  int32_t *n, int64_t *m, int32_t *rowPtr, int32_t *colInd,
                                  int32_t *weights, int64_t capacity) {
  CsrGraph g;
  road_grid(width, height, K, wmin, wmax, drop, diag, seed, g);
  return export_graph(g, n, m, rowPtr, colInd, weights, capacity);
}

extern "C" int mosp_gen_rmat(int32_t scale, int32_t edgeFactor, double a, double b, double c,
                             int32_t K, int32_t wmin, int32_t wmax, uint64_t seed, int32_t *n,
                             int64_t *m, int32_t *rowPtr, int32_t *colInd, int32_t *weights,
                             int64_t capacity) {
  CsrGraph g;
  rmat(scale, edgeFactor, a, b, c, K, wmin, wmax, seed, g);
  return export_graph(g, n, m, rowPtr, colInd, weights, capacity);
}

// Change batch for a graph; output capacity: n_ins * (1 + 1 + K) and n_del.
extern "C" int mosp_gen_batch(int32_t n, int32_t K, const int32_t *rowPtr, const int32_t *colInd,
                              const int32_t *weights, int32_t changes, double insPct,
                              int32_t safe, int32_t wmin, int32_t wmax, uint64_t seed,
                              int32_t source, int32_t *nIns, int32_t *insFrom, int32_t *insTo,
                              int32_t *insW, int32_t *nDel, int32_t *delFrom, int32_t *delTo) {
  CsrGraph g;
  g.numberOfNodes = n;
  g.numberOfObjectives = K;
  const int m = rowPtr[n];
  g.rowPtr.assign(rowPtr, rowPtr + n + 1);
  g.colInd.assign(colInd, colInd + m);
  g.weights.assign(weights, weights + static_cast<size_t>(m) * K);
  ChangeBatch b;
  make_batch(g, changes, insPct, safe, wmin, wmax, seed, source, b);
  *nIns = b.numberOfInserts();
  *nDel = b.numberOfDeletes();
  std::copy(b.insertFrom.begin(), b.insertFrom.end(), insFrom);
  std::copy(b.insertTo.begin(), b.insertTo.end(), insTo);
  std::copy(b.insertWeights.begin(), b.insertWeights.end(), insW);
  std::copy(b.deleteFrom.begin(), b.deleteFrom.end(), delFrom);
  std::copy(b.deleteTo.begin(), b.deleteTo.end(), delTo);
  return 1;
}
