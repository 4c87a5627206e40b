// mosp -- command-line driver of the MOSP update on MI355X.
//
// Drop-in for the reference's bin/mosp (src/mosp.cu): the same options,
// input and output files, "VALIDATE ..." lines and the summary line
//   RESULT gpu_compute_ms=<a> end_to_end_ms=<b>
// where (a) is the K SOSP updates + Steps 2-3 (host wall clock around the
// synchronous GPU calls) and (b) runs from reading the inputs to writing the
// outputs.  Extra options: --serial (K launches + 1, the reference's
// schedule, instead of one fused launch), --delta-scale and
// --comb-delta-scale (near-far widths; outputs do not depend on them).
//
// The driver is a table of options and a sequence of stages (inputs,
// canonical ties, update, outputs, report, validation), each a function
// over one Run record.  The MOSP path costs are computed on the GPU; the
// validation compares against independent from-scratch GPU searches of the
// updated graph and of a combined graph built on the host.
#include <algorithm>
#include <charconv>
#include <chrono>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <filesystem>
#include <functional>
#include <iostream>
#include <string>
#include <vector>

#include "csrGraph.h"
#include "deviceArray.h"
#include "deviceGraph.h"
#include "mospUpdate.h"
#include "mosp_hip.h"
#include "parallelCombinedGraph.h"
#include "sospUpdateGpu.h"
#include "stageTimer.h"

namespace {

using Clock = std::chrono::steady_clock;

double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// ------------------------------------------------------------- options --

struct Settings {
  std::string graph, changes, init, out = "mosp-output", timing, cache;
  std::vector<int> pref;
  long long K = 0, source = 0, delta = 0, combScale = 1;
  double deltaScale = 1.0;
  bool canonicalize = false, noOutput = false, validate = false, quiet = false, serial = false;
};

// One command-line option: a string, an integer in [lo, hi], a switch, a
// comma list of integers >= 1, or a positive real.
struct Option {
  const char *name;
  enum Kind { TEXT, INTEGER, SWITCH, LIST, REAL } kind;
  long long lo, hi;
  std::function<void(Settings &, const std::string &, long long)> set;
};

const std::vector<Option> &options() {
  using S = Settings;
  static const std::vector<Option> table = {
      {"--graph", Option::TEXT, 0, 0, [](S &s, const std::string &v, long long) { s.graph = v; }},
      {"--changes", Option::TEXT, 0, 0, [](S &s, const std::string &v, long long) { s.changes = v; }},
      {"--init", Option::TEXT, 0, 0, [](S &s, const std::string &v, long long) { s.init = v; }},
      {"--out", Option::TEXT, 0, 0, [](S &s, const std::string &v, long long) { s.out = v; }},
      {"--timing", Option::TEXT, 0, 0, [](S &s, const std::string &v, long long) { s.timing = v; }},
      {"--cache", Option::TEXT, 0, 0, [](S &s, const std::string &v, long long) { s.cache = v; }},
      {"-k", Option::INTEGER, 0, INT_MAX, [](S &s, const std::string &, long long x) { s.K = x; }},
      {"--source", Option::INTEGER, 0, INT_MAX,
       [](S &s, const std::string &, long long x) { s.source = x; }},
      {"--delta", Option::INTEGER, 1, LLONG_MAX,
       [](S &s, const std::string &, long long x) { s.delta = x; }},
      {"--comb-delta-scale", Option::INTEGER, 1, INT_MAX,
       [](S &s, const std::string &, long long x) { s.combScale = x; }},
      {"--delta-scale", Option::REAL, 0, 0,
       [](S &s, const std::string &v, long long) { s.deltaScale = std::stod(v); }},
      {"--pref", Option::LIST, 1, INT_MAX, [](S &, const std::string &, long long) {}},
      {"--canonicalize", Option::SWITCH, 0, 0, [](S &s, const std::string &, long long) { s.canonicalize = true; }},
      {"--no-output", Option::SWITCH, 0, 0, [](S &s, const std::string &, long long) { s.noOutput = true; }},
      {"--validate", Option::SWITCH, 0, 0, [](S &s, const std::string &, long long) { s.validate = true; }},
      {"--quiet", Option::SWITCH, 0, 0, [](S &s, const std::string &, long long) { s.quiet = true; }},
      {"--serial", Option::SWITCH, 0, 0, [](S &s, const std::string &, long long) { s.serial = true; }},
  };
  return table;
}

void usage() {
  std::cerr << "usage: mosp --graph <csrPrefix> --changes <dir> --init <dir> [options]\n";
  for (const Option &o : options()) {
    const char *arg = o.kind == Option::SWITCH ? "" : o.kind == Option::LIST ? " p1,..,pK"
                      : o.kind == Option::TEXT ? " <value>" : " <number>";
    std::cerr << "  " << o.name << arg << "\n";
  }
}

bool parse_list(const std::string &v, std::vector<int> &out) {
  out.clear();
  size_t from = 0;
  for (;;) {
    const size_t comma = v.find(',', from);
    long long x = 0;
    if (!parseInteger(v.substr(from, comma - from), 1, INT_MAX, x)) return false;
    out.push_back(static_cast<int>(x));
    if (comma == std::string::npos) return true;
    from = comma + 1;
  }
}

bool parse_command_line(int argc, char **argv, Settings &s) {
  for (int i = 1; i < argc; ++i) {
    const std::string name = argv[i];
    const auto it = std::find_if(options().begin(), options().end(),
                                 [&](const Option &o) { return name == o.name; });
    if (it == options().end()) {
      std::cerr << "unknown option: " << name << "\n";
      return false;
    }
    if (it->kind == Option::SWITCH) {
      it->set(s, "", 0);
      continue;
    }
    if (i + 1 >= argc) return false;
    const std::string value = argv[++i];
    long long x = 0;
    switch (it->kind) {
      case Option::INTEGER:
        if (!parseInteger(value, it->lo, it->hi, x)) {
          std::cerr << name << " needs an integer in [" << it->lo << ", " << it->hi
                    << "], got '" << value << "'\n";
          return false;
        }
        break;
      case Option::LIST:
        if (!parse_list(value, s.pref)) {
          std::cerr << name << " needs integers >= 1, got '" << value << "'\n";
          return false;
        }
        break;
      case Option::REAL: {
        double d = 0;
        const auto r = std::from_chars(value.data(), value.data() + value.size(), d);
        if (r.ec != std::errc() || r.ptr != value.data() + value.size() || !(d > 0)) {
          std::cerr << name << " needs a positive number, got '" << value << "'\n";
          return false;
        }
        break;
      }
      default:
        break;
    }
    it->set(s, value, x);
  }
  return !s.graph.empty() && !s.changes.empty() && !s.init.empty();
}

// ----------------------------------------------------------------- run --

struct Run {
  Settings opt;
  Clock::time_point start;
  double contextMs = 0, readMs = 0, canonMs = 0, writeMs = 0;
  CsrGraph original, updated;
  ChangeBatch batch;
  int n = 0, KG = 0, K = 0;
  std::vector<long long> distances;  // K * n, objective-major
  std::vector<int> parents;
  MospResult result;
};

template <typename T>
std::vector<T> objective(const std::vector<T> &all, int k, int n) {
  const auto at = all.begin() + static_cast<std::ptrdiff_t>(k) * n;
  return std::vector<T>(at, at + n);
}

std::string obj_dir(const std::string &root, int k) { return root + "/obj" + std::to_string(k); }

bool read_inputs(Run &r) {
  const auto t = Clock::now();
  {
    ScopedStage stage("read_graph");
    if (!loadCsrGraph(r.opt.graph, r.original, r.opt.cache)) return false;
  }
  r.n = r.original.numberOfNodes;
  r.KG = r.original.numberOfObjectives;
  r.K = r.opt.K > 0 ? static_cast<int>(std::min<long long>(r.opt.K, r.KG)) : r.KG;
  const size_t n = static_cast<size_t>(r.n);
  r.distances.resize(n * r.K);
  r.parents.resize(n * r.K);
  ScopedStage stage("read_changes_and_trees");
  // The batch and the 2K tree files are independent: one job each, every
  // tree read straight into its slice.
  std::vector<std::vector<long long>> d(r.K);
  std::vector<std::vector<int>> p(r.K);
  std::vector<std::function<bool()>> jobs;
  jobs.emplace_back([&] {
    return readChangeBatch(r.opt.changes + "/insert.txt", r.opt.changes + "/delete.txt", r.KG,
                           r.n, r.batch);
  });
  for (int k = 0; k < r.K; ++k) {
    const std::string dir = obj_dir(r.opt.init, k);
    jobs.emplace_back([&, k, dir] { return readDistances(dir + "/distancesOriginal.txt", r.n, d[k]); });
    jobs.emplace_back([&, k, dir] { return readParents(dir + "/SSSPTreeOriginal.txt", r.n, p[k]); });
  }
  if (!runConcurrently(jobs)) return false;
  for (int k = 0; k < r.K; ++k) {
    std::copy(d[k].begin(), d[k].end(), r.distances.begin() + static_cast<std::ptrdiff_t>(k * n));
    std::copy(p[k].begin(), p[k].end(), r.parents.begin() + static_cast<std::ptrdiff_t>(k * n));
  }
  r.readMs = ms_between(t, Clock::now());
  return true;
}

void canonicalize(Run &r) {
  const auto t = Clock::now();
  for (int k = 0; k < r.K; ++k) {
    canonicalizeTree(r.original, k, static_cast<int>(r.opt.source), objective(r.distances, k, r.n),
                     r.parents.data() + static_cast<size_t>(k) * r.n);
  }
  r.canonMs = ms_between(t, Clock::now());
}

bool update(Run &r) {
  MospOptions o;
  o.source = static_cast<int>(r.opt.source);
  o.numberOfObjectives = r.K;
  o.preferences = r.opt.pref;
  o.delta = r.opt.delta;
  o.deltaScale = r.opt.deltaScale;
  o.fused = !r.opt.serial;
  o.combinedDeltaScale = static_cast<int>(r.opt.combScale);
  return mospUpdate(r.original, r.batch, r.distances, r.parents, o, r.updated, r.result);
}

// mospCosts.txt: "v c1 .. cKG" per vertex (INF for unreachable ones).
bool write_costs(const std::string &path, const std::vector<long long> &costs, int KG) {
  const size_t n = KG > 0 ? costs.size() / KG : 0;
  std::string text;
  text.resize(n * (12 + 21 * static_cast<size_t>(KG)));
  char *o = &text[0];
  for (size_t v = 0; v < n; ++v) {
    o = std::to_chars(o, o + 12, static_cast<long long>(v)).ptr;
    for (int k = 0; k < KG; ++k) {
      *o++ = ' ';
      const long long c = costs[v * KG + k];
      if (c >= DISTANCE_INF / 2) {
        *o++ = 'I';
        *o++ = 'N';
        *o++ = 'F';
      } else {
        o = std::to_chars(o, o + 20, c).ptr;
      }
    }
    *o++ = '\n';
  }
  text.resize(static_cast<size_t>(o - text.data()));
  FILE *f = std::fopen(path.c_str(), "wb");
  if (f == nullptr) return false;
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  return std::fclose(f) == 0 && ok;
}

// MOSP path costs over all KG objectives of the updated graph, on the GPU.
bool path_costs(const Run &r, std::vector<long long> &costs) {
  DeviceGraph g;
  DeviceArray<int> parent;
  DeviceArray<long long> c(static_cast<size_t>(r.n) * r.KG);
  if (!uploadDeviceGraph(r.updated, g) || !parent.upload(r.result.combinedParent) ||
      !mosp_path_costs_gpu(g.handle, parent.data(), static_cast<int>(r.opt.source),
                           reinterpret_cast<int64_t *>(c.data()), nullptr) ||
      !c.download(costs)) {
    std::cerr << mosp_last_error() << "\n";
    return false;
  }
  return true;
}

bool write_outputs(Run &r) {
  const auto t = Clock::now();
  ScopedStage stage("write_outputs");
  namespace fs = std::filesystem;
  const std::string comb = r.opt.out + "/combinedGraph";
  std::error_code ec;
  for (int k = 0; k < r.K; ++k) fs::create_directories(obj_dir(r.opt.out, k), ec);
  fs::create_directories(comb, ec);
  std::vector<long long> costs;
  const bool costsOk = path_costs(r, costs);
  const MospResult &res = r.result;
  std::vector<std::function<bool()>> jobs;
  for (int k = 0; k < r.K; ++k) {
    const std::string dir = obj_dir(r.opt.out, k);
    jobs.emplace_back([&, k, dir] {
      return writeDistances(dir + "/distancesUpdated.txt", objective(res.distances, k, r.n));
    });
    jobs.emplace_back([&, k, dir] {
      return writeParents(dir + "/SSSPTreeUpdated.txt", objective(res.parents, k, r.n));
    });
  }
  jobs.emplace_back([&] { return writeDistances(comb + "/distancesCsr.txt", res.combinedDistances); });
  jobs.emplace_back([&] { return writeParents(comb + "/SSSPTreeCsr.txt", res.combinedParent); });
  jobs.emplace_back([&] { return costsOk && write_costs(comb + "/mospCosts.txt", costs, r.KG); });
  if (!runConcurrently(jobs)) {
    std::cerr << "cannot write the outputs to " << r.opt.out << "\n";
    return false;
  }
  r.writeMs = ms_between(t, Clock::now());
  return true;
}

void report(const Run &r) {
  const MospTimings &tm = r.result.timings;
  std::printf("graph  n=%d m=%d K=%d; batch %d inserts, %d deletes\n", r.n,
              r.updated.numberOfEdges(), r.K, r.batch.numberOfInserts(),
              r.batch.numberOfDeletes());
  std::printf("host   context %.1f ms, read inputs %.1f ms, canonicalize %.1f ms, upload %.1f ms, "
              "apply batch (GPU) %.1f ms, download %.1f ms, write %.1f ms\n",
              r.contextMs, r.readMs, r.canonMs, tm.upload, tm.applyBatch, tm.download, r.writeMs);
  for (int k = 0; k < r.K; ++k) {
    const SospStats &s = r.result.objectiveStats[k];
    std::printf("obj%d   SOSP update %.3f ms (invalidated %d, jump rounds %d, iterations %d, "
                "epochs %d, pushes %lld)\n",
                k, tm.objectives[k], s.invalidated, s.jumpRounds, s.iterations, s.epochs,
                s.pushes);
  }
  const CombineStats &c = r.result.combineStats;
  std::printf("comb   combined graph + SOSP %.3f ms (%d edges, L=%lld, delta %lld, iterations "
              "%d, pushes %lld)\n",
              tm.combined, c.numberOfEdges, c.scale, c.delta, c.search.iterations,
              c.search.pushes);
}

// ---------------------------------------------------------- validation --

struct Mismatch {
  long long distance = 0, parent = 0;
  bool none() const { return distance == 0 && parent == 0; }
};

Mismatch compare_trees(const std::vector<long long> &d, const std::vector<int> &p,
                       const std::vector<long long> &rd, const std::vector<int> &rp) {
  Mismatch m;
  for (size_t v = 0; v < d.size(); ++v) {
    const bool a = d[v] < DISTANCE_INF / 2, b = rd[v] < DISTANCE_INF / 2;
    m.distance += (a != b || (a && d[v] != rd[v])) ? 1 : 0;
    m.parent += p[v] != rp[v] ? 1 : 0;
  }
  return m;
}

void print_check(const std::string &what, const Mismatch &m) {
  std::cout << "VALIDATE " << what << " " << (m.none() ? "PASS" : "FAIL")
            << " (distance mismatches=" << m.distance << " parent mismatches=" << m.parent
            << ")\n";
}

// A from-scratch GPU search of objective k of g (default width, bound from
// the graph's largest weight).
bool scratch_tree(const CsrGraph &host, const DeviceGraph &g, int k, int source,
                  SospWorkspace &ws, std::vector<long long> &d, std::vector<int> &p) {
  long long maxW = 1, sum = 0;
  for (int e = 0; e < host.numberOfEdges(); ++e) {
    maxW = std::max<long long>(maxW, host.weight(e, k));
    sum += host.weight(e, k);
  }
  const int n = host.numberOfNodes;
  DeviceArray<long long> dd(n);
  DeviceArray<int> dp(n);
  return sospFromScratchGpu(g.out(k), source, defaultDelta(host.numberOfEdges(), n, sum), maxW,
                            ws, dd.data(), dp.data()) &&
         dd.download(d) && dp.download(p);
}

// The combined graph of the result's K trees, built on the host: for every
// vertex v and distinct parent p among its K parents, the edge p -> v with
// the weight of the trees that contain it.
CsrGraph host_combined_graph(const MospResult &res, int n, int source, const std::vector<int> &pref) {
  const int K = res.numberOfObjectives;
  const long long L = preferenceScale(pref, K);
  std::vector<std::vector<std::pair<int, int>>> rows(n);
  for (int v = 0; v < n; ++v) {
    if (v == source) continue;
    unsigned done = 0;
    for (int k = 0; k < K; ++k) {
      const int p = res.parents[static_cast<size_t>(k) * n + v];
      if (p < 0 || ((done >> k) & 1u)) continue;
      unsigned mask = 0;
      for (int j = k; j < K; ++j) {
        if (res.parents[static_cast<size_t>(j) * n + v] == p) mask |= 1u << j;
      }
      done |= mask;
      rows[p].emplace_back(v, static_cast<int>(combinedEdgeWeight(mask, pref, K, L)));
    }
  }
  CsrGraph c;
  c.numberOfNodes = n;
  c.numberOfObjectives = 1;
  c.rowPtr.assign(static_cast<size_t>(n) + 1, 0);
  for (int u = 0; u < n; ++u) {
    c.rowPtr[u + 1] = c.rowPtr[u] + static_cast<int>(rows[u].size());
    for (const auto &e : rows[u]) {
      c.colInd.push_back(e.first);
      c.weights.push_back(e.second);
    }
  }
  return c;
}

bool validate(const Run &r) {
  const int n = r.n, source = static_cast<int>(r.opt.source);
  SospWorkspace ws;
  DeviceGraph g;
  if (!uploadDeviceGraph(r.updated, g) || !ws.reserve(n)) return false;
  bool all = true;
  for (int k = 0; k < r.K; ++k) {
    std::vector<long long> d;
    std::vector<int> p;
    if (!scratch_tree(r.updated, g, k, source, ws, d, p)) return false;
    const Mismatch m = compare_trees(objective(r.result.distances, k, n),
                                     objective(r.result.parents, k, n), d, p);
    print_check("obj" + std::to_string(k), m);
    all = all && m.none();
  }
  const CsrGraph comb = host_combined_graph(r.result, n, source, r.opt.pref);
  DeviceGraph cg;
  std::vector<long long> d;
  std::vector<int> p;
  if (!uploadDeviceGraph(comb, cg) || !scratch_tree(comb, cg, 0, source, ws, d, p)) return false;
  const Mismatch m = compare_trees(r.result.combinedDistances, r.result.combinedParent, d, p);
  print_check("combined", m);
  return all && m.none();
}

}  // namespace

int main(int argc, char **argv) {
  Run r;
  if (!parse_command_line(argc, argv, r.opt)) {
    usage();
    return 2;
  }
  setInstrumentation(!r.opt.timing.empty());
  r.start = Clock::now();
  {
    ScopedStage stage("device_context_init");
    const auto t = Clock::now();
    (void)mosp_device_count();  // the runtime and device context, outside the timed region
    void *probe = mosp_dev_alloc(1);
    if (probe != nullptr) (void)mosp_dev_free(probe);
    r.contextMs = ms_between(t, Clock::now());
  }
  if (!read_inputs(r)) return 1;
  if (r.opt.canonicalize) canonicalize(r);
  if (!update(r)) return 1;
  if (!r.opt.noOutput && !write_outputs(r)) return 1;
  const double endToEnd = ms_between(r.start, Clock::now());
  if (!r.opt.quiet) report(r);
  if (!r.opt.timing.empty()) {
    recordStage("TOTAL_end_to_end", endToEnd);
    if (!writeInstrumentationCsv(r.opt.timing)) std::cerr << "cannot write " << r.opt.timing << "\n";
  }
  std::printf("RESULT gpu_compute_ms=%.3f end_to_end_ms=%.3f\n", r.result.timings.gpuCompute(),
              endToEnd);
  std::fflush(stdout);
  if (r.opt.validate && !validate(r)) return 1;
  return 0;
}
