// csr_graph.cpp -- host graph I/O, binary cache and batch application.
//
// Behaviour follows the reference's src/csrGraph.cu (readers 405-472,
// 626-725, 943-1030; writers 474-501, 1032-1064; binary cache 503-620;
// applyChangeBatch 727-894; transposeCsrGraph 896-922) so files, messages'
// meaning and results are interchangeable.  The code is this project's:
// a single cursor-based tokenizer over memory-mapped text, per-row rebuild of
// only the rows a batch touches, and std::thread for concurrent jobs.
#include "csrGraph.h"

#include <sys/stat.h>

#include <algorithm>
#include <atomic>
#include <charconv>
#include <chrono>
#include <cerrno>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

namespace fs = std::filesystem;

namespace {

bool slurp(const std::string &path, std::string &text) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (f == nullptr) {
    return false;
  }
  struct stat st;
  if (fstat(fileno(f), &st) != 0) {
    std::fclose(f);
    return false;
  }
  text.resize(static_cast<size_t>(st.st_size));
  const size_t got = text.empty() ? 0 : std::fread(&text[0], 1, text.size(), f);
  std::fclose(f);
  return got == text.size();
}

void make_parent_dirs(const std::string &path) {
  const fs::path parent = fs::path(path).parent_path();
  if (!parent.empty()) {
    std::error_code ec;
    fs::create_directories(parent, ec);
  }
}

long long line_of(const char *begin, const char *pos) {
  return 1 + static_cast<long long>(std::count(begin, pos, '\n'));
}

std::string at(const std::string &path, const char *begin, const char *pos) {
  return path + ":" + std::to_string(line_of(begin, pos));
}

// Cursor over one text buffer (or over [p, end) of it; `begin` is the
// buffer's start, for line numbers).
struct Cursor {
  const char *begin, *p, *end;
  explicit Cursor(const std::string &s)
      : begin(s.data()), p(s.data()), end(s.data() + s.size()) {}
  Cursor(const char *b, const char *from, const char *to) : begin(b), p(from), end(to) {}
  void blanks() {  // spaces, tabs, carriage returns
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  }
  void space() {  // all whitespace
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n')) ++p;
  }
  bool eol() const { return p >= end || *p == '\n'; }
  bool number(long long &v) {
    auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
  }
};

bool valid_weight(long long w) { return w >= 1 && w <= INT_MAX; }

// Large files are parsed in parallel: cut into parts at line boundaries,
// one thread per part, each into its own array, then concatenated.  Every
// part stops at its first error, so the first failing part's error is the
// file's first error, as in one sequential pass.
constexpr size_t kMinPart = size_t{4} << 20;

std::vector<const char *> line_cuts(const std::string &text) {
  const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const size_t parts = std::max<size_t>(1, std::min<size_t>(hw, text.size() / kMinPart));
  const char *const b = text.data(), *const e = b + text.size();
  std::vector<const char *> cuts{b};
  for (size_t i = 1; i < parts; ++i) {
    const char *c = std::max(cuts.back(), b + text.size() * i / parts);
    c = std::find(c, e, '\n');
    cuts.push_back(c < e ? c + 1 : e);
  }
  cuts.push_back(e);
  return cuts;
}

struct Part {
  std::vector<int> vals;
  std::string err;
  long long lines = 0;
  bool ok = true;
};

// Runs parse(part, from, to) over every part; false with the first error.
template <typename F>
bool parse_parts(const std::string &text, std::vector<Part> &parts, std::string &err, F parse) {
  const std::vector<const char *> cuts = line_cuts(text);
  parts.assign(cuts.size() - 1, Part{});
  std::vector<std::thread> threads;
  for (size_t i = 0; i + 1 < parts.size(); ++i) {
    threads.emplace_back([&, i] { parts[i].ok = parse(parts[i], cuts[i], cuts[i + 1]); });
  }
  parts.back().ok = parse(parts.back(), cuts[parts.size() - 1], cuts[parts.size()]);
  for (auto &t : threads) t.join();
  for (const Part &pt : parts) {
    if (!pt.ok) {
      err = pt.err;
      return false;
    }
  }
  return true;
}

// The parts' arrays in order into `out` (one copying thread per part).
void concatenate(std::vector<Part> &parts, std::vector<int> &out) {
  std::vector<size_t> at(parts.size() + 1, 0);
  for (size_t i = 0; i < parts.size(); ++i) at[i + 1] = at[i] + parts[i].vals.size();
  out.resize(at.back());
  auto copy = [&](size_t i) {
    std::copy(parts[i].vals.begin(), parts[i].vals.end(),
              out.begin() + static_cast<std::ptrdiff_t>(at[i]));
    std::vector<int>().swap(parts[i].vals);
  };
  std::vector<std::thread> threads;
  for (size_t i = 1; i < parts.size(); ++i) threads.emplace_back(copy, i);
  copy(0);
  for (auto &t : threads) t.join();
}

bool read_ids(const std::string &path, std::vector<int> &out, std::string &err) {
  std::string text;
  if (!slurp(path, text)) {
    err = "could not read " + path;
    return false;
  }
  std::vector<Part> parts;
  const char *const begin = text.data();
  const bool ok = parse_parts(text, parts, err, [&](Part &pt, const char *from, const char *to) {
    Cursor c(begin, from, to);
    pt.vals.reserve(static_cast<size_t>(to - from) / 4);
    c.space();
    while (c.p < c.end) {
      long long v = 0;
      if (!c.number(v)) {
        pt.err = "invalid token at " + at(path, c.begin, c.p);
        return false;
      }
      if (v < 0 || v > INT_MAX) {
        pt.err = "value " + std::to_string(v) + " out of range [0, " +
                 std::to_string(INT_MAX) + "] at " + at(path, c.begin, c.p);
        return false;
      }
      pt.vals.push_back(static_cast<int>(v));
      c.space();
    }
    return true;
  });
  if (!ok) {
    return false;
  }
  concatenate(parts, out);
  return true;
}

// Weights per line of [from, to): K of them (K == 0: the first non-empty
// line sets it, as it does for the whole file).
bool parse_values(const std::string &path, Cursor c, int &K, Part &pt) {
  for (;;) {
    c.blanks();
    int onLine = 0;
    while (!c.eol()) {
      long long v = 0;
      if (!c.number(v)) {
        pt.err = "Invalid token in CSR values file at " + at(path, c.begin, c.p) + ".";
        return false;
      }
      if (!valid_weight(v)) {
        pt.err = "CSR weight " + std::to_string(v) + " out of range [1, " +
                 std::to_string(INT_MAX) + "] at " + at(path, c.begin, c.p) + ".";
        return false;
      }
      pt.vals.push_back(static_cast<int>(v));
      ++onLine;
      c.blanks();
    }
    if (onLine > 0) {
      if (K == 0) {
        K = onLine;
      } else if (onLine != K) {
        pt.err = "Inconsistent number of objectives at " + at(path, c.begin, c.p) + ".";
        return false;
      }
      ++pt.lines;
    }
    if (c.p >= c.end) break;
    ++c.p;  // newline
  }
  return true;
}

bool read_values(const std::string &path, std::vector<int> &w, int &K,
                 long long &lines, std::string &err) {
  std::string text;
  if (!slurp(path, text)) {
    err = "Could not read CSR values file " + path + ".";
    return false;
  }
  // K from the first non-empty line, read alone (its errors are the file's
  // first ones); then every part checks its lines against it.
  K = 0;
  {
    Part first;
    const char *const e = text.data() + text.size();
    const char *p = text.data();
    while (p < e && K == 0) {
      const char *nl = std::find(p, e, '\n');
      if (!parse_values(path, Cursor(text.data(), p, nl), K, first)) {
        err = first.err;
        return false;
      }
      p = nl < e ? nl + 1 : e;
    }
  }
  std::vector<Part> parts;
  const char *const begin = text.data();
  const int k = K;
  const bool ok = parse_parts(text, parts, err, [&](Part &pt, const char *from, const char *to) {
    int partK = k;
    pt.vals.reserve(static_cast<size_t>(to - from) / 3);
    return parse_values(path, Cursor(begin, from, to), partK, pt);
  });
  if (!ok) {
    return false;
  }
  lines = 0;
  for (const Part &pt : parts) lines += pt.lines;
  concatenate(parts, w);
  return true;
}

// Buffered writer with std::to_chars formatting.
class Writer {
 public:
  explicit Writer(const std::string &path) {
    make_parent_dirs(path);
    f_ = std::fopen(path.c_str(), "wb");
    buf_.reserve(kChunk + 64);
  }
  ~Writer() { close(); }
  bool ok() const { return f_ != nullptr && !bad_; }
  void num(long long v) {
    char d[24];
    auto r = std::to_chars(d, d + sizeof(d), v);
    buf_.append(d, r.ptr);
    spill();
  }
  void str(const char *s) {
    buf_.append(s);
    spill();
  }
  void ch(char c) {
    buf_.push_back(c);
    spill();
  }
  bool close() {
    if (f_ == nullptr) return !bad_;
    flush();
    if (std::fclose(f_) != 0) bad_ = true;
    f_ = nullptr;
    return !bad_;
  }

 private:
  static constexpr size_t kChunk = 1 << 22;
  void spill() {
    if (buf_.size() >= kChunk) flush();
  }
  void flush() {
    if (f_ != nullptr && !buf_.empty() &&
        std::fwrite(buf_.data(), 1, buf_.size(), f_) != buf_.size()) {
      bad_ = true;
    }
    buf_.clear();
  }
  FILE *f_ = nullptr;
  std::string buf_;
  bool bad_ = false;
};

constexpr char kMagic[8] = {'M', 'O', 'S', 'P', 'C', 'S', 'R', '2'};

struct Identity {
  std::string prefix;
  int64_t size[3] = {0, 0, 0};
  int64_t mtime[3] = {0, 0, 0};
  bool operator==(const Identity &o) const {
    return prefix == o.prefix && std::equal(size, size + 3, o.size) &&
           std::equal(mtime, mtime + 3, o.mtime);
  }
};

bool identity_of(const std::string &prefix, Identity &id) {
  std::error_code ec;
  const fs::path canon = fs::weakly_canonical(prefix, ec);
  if (ec) return false;
  id.prefix = canon.string();
  const char *suffix[3] = {"RowPtr.txt", "ColInd.txt", "Values.txt"};
  for (int i = 0; i < 3; ++i) {
    const std::string file = prefix + suffix[i];
    const auto size = fs::file_size(file, ec);
    if (ec) return false;
    const auto time = fs::last_write_time(file, ec);
    if (ec) return false;
    id.size[i] = static_cast<int64_t>(size);
    id.mtime[i] = static_cast<int64_t>(
        std::chrono::duration_cast<std::chrono::nanoseconds>(
            time.time_since_epoch()).count());
  }
  return true;
}

bool valid_arrays(const CsrGraph &g) {
  const int n = g.numberOfNodes;
  const size_t m = g.colInd.size();
  if (g.rowPtr.size() != static_cast<size_t>(n) + 1 || g.rowPtr[0] != 0 ||
      static_cast<size_t>(g.rowPtr[n]) != m ||
      g.weights.size() != m * static_cast<size_t>(g.numberOfObjectives)) {
    return false;
  }
  for (int i = 0; i < n; ++i) {
    if (g.rowPtr[i + 1] < g.rowPtr[i]) return false;
  }
  for (int v : g.colInd) {
    if (v < 0 || v >= n) return false;
  }
  for (int w : g.weights) {
    if (w < 1) return false;
  }
  return true;
}

}  // namespace

bool runConcurrently(const std::vector<std::function<bool()>> &jobs) {
  std::vector<char> ok(jobs.size(), 0);
  std::vector<std::thread> threads;
  threads.reserve(jobs.size());
  for (size_t i = 0; i < jobs.size(); ++i) {
    threads.emplace_back([&, i] { ok[i] = jobs[i]() ? 1 : 0; });
  }
  for (auto &t : threads) t.join();
  return std::all_of(ok.begin(), ok.end(), [](char c) { return c != 0; });
}

bool parseInteger(const std::string &text, long long minimum, long long maximum,
                  long long &value) {
  if (text.empty() || text[0] == '+' || std::isspace(static_cast<unsigned char>(text[0]))) {
    return false;
  }
  const char *b = text.data(), *e = text.data() + text.size();
  auto r = std::from_chars(b, e, value);
  return r.ec == std::errc() && r.ptr == e && value >= minimum && value <= maximum;
}

bool readCsrGraph(const std::string &prefix, CsrGraph &graph) {
  graph = CsrGraph();
  bool okRows = false, okCols = false, okVals = false;
  std::string eRows, eCols, eVals;
  int K = 0;
  long long lines = 0;
  runConcurrently({
      [&] { return okRows = read_ids(prefix + "RowPtr.txt", graph.rowPtr, eRows); },
      [&] { return okCols = read_ids(prefix + "ColInd.txt", graph.colInd, eCols); },
      [&] { return okVals = read_values(prefix + "Values.txt", graph.weights, K, lines, eVals); },
  });
  if (!okRows) {
    std::cout << "Error: CSR row pointers: " << eRows << "\n";
    return false;
  }
  if (graph.rowPtr.size() < 2) {
    std::cout << "Error: Could not read CSR row pointers: " << prefix << "RowPtr.txt\n";
    return false;
  }
  graph.numberOfNodes = static_cast<int>(graph.rowPtr.size()) - 1;
  if (graph.rowPtr[0] != 0) {
    std::cout << "Error: CSR row pointers must start at 0.\n";
    return false;
  }
  for (int i = 0; i < graph.numberOfNodes; ++i) {
    if (graph.rowPtr[i + 1] < graph.rowPtr[i]) {
      std::cout << "Error: CSR row pointers are not monotone.\n";
      return false;
    }
  }
  if (!okCols) {
    std::cout << "Error: CSR column indices: " << eCols << "\n";
    return false;
  }
  if (static_cast<long long>(graph.colInd.size()) != graph.rowPtr.back()) {
    std::cout << "Error: CSR column index file missing or size mismatch.\n";
    return false;
  }
  for (int v : graph.colInd) {
    if (v < 0 || v >= graph.numberOfNodes) {
      std::cout << "Error: CSR column index out of range.\n";
      return false;
    }
  }
  if (!okVals) {
    std::cout << "Error: " << eVals << "\n";
    return false;
  }
  if (lines != graph.rowPtr.back()) {
    std::cout << "Error: values size mismatch.\n";
    return false;
  }
  graph.numberOfObjectives = K;
  return true;
}

bool writeCsrGraph(const std::string &prefix, const CsrGraph &graph) {
  Writer rows(prefix + "RowPtr.txt"), cols(prefix + "ColInd.txt"),
      vals(prefix + "Values.txt");
  if (!rows.ok() || !cols.ok() || !vals.ok()) {
    std::cout << "Error: Could not write CSR files: " << prefix << "\n";
    return false;
  }
  for (int x : graph.rowPtr) {
    rows.num(x);
    rows.ch('\n');
  }
  for (int x : graph.colInd) {
    cols.num(x);
    cols.ch('\n');
  }
  const int K = graph.numberOfObjectives;
  for (int e = 0; e < graph.numberOfEdges(); ++e) {
    for (int k = 0; k < K; ++k) {
      if (k > 0) vals.ch(' ');
      vals.num(graph.weight(e, k));
    }
    vals.ch('\n');
  }
  return rows.close() && cols.close() && vals.close();
}

bool saveCsrGraphBinary(const std::string &path, const CsrGraph &graph,
                        const std::string &sourcePrefix) {
  Identity id;
  if (!identity_of(sourcePrefix, id)) return false;
  make_parent_dirs(path);
  FILE *f = std::fopen(path.c_str(), "wb");
  if (f == nullptr) return false;
  const int32_t header[2] = {graph.numberOfNodes, graph.numberOfObjectives};
  const int64_t m = graph.numberOfEdges();
  const int64_t plen = static_cast<int64_t>(id.prefix.size());
  bool ok = std::fwrite(kMagic, 1, 8, f) == 8 &&
            std::fwrite(&plen, sizeof(plen), 1, f) == 1 &&
            std::fwrite(id.prefix.data(), 1, id.prefix.size(), f) == id.prefix.size() &&
            std::fwrite(id.size, sizeof(id.size), 1, f) == 1 &&
            std::fwrite(id.mtime, sizeof(id.mtime), 1, f) == 1 &&
            std::fwrite(header, sizeof(header), 1, f) == 1 &&
            std::fwrite(&m, sizeof(m), 1, f) == 1 &&
            std::fwrite(graph.rowPtr.data(), sizeof(int), graph.rowPtr.size(), f) ==
                graph.rowPtr.size() &&
            std::fwrite(graph.colInd.data(), sizeof(int), graph.colInd.size(), f) ==
                graph.colInd.size() &&
            std::fwrite(graph.weights.data(), sizeof(int), graph.weights.size(), f) ==
                graph.weights.size();
  ok = std::fclose(f) == 0 && ok;
  if (!ok) {
    std::error_code ec;
    fs::remove(path, ec);
  }
  return ok;
}

bool loadCsrGraphBinary(const std::string &path, CsrGraph &graph,
                        const std::string &sourcePrefix) {
  Identity want, got;
  if (!identity_of(sourcePrefix, want)) return false;
  FILE *f = std::fopen(path.c_str(), "rb");
  if (f == nullptr) return false;
  char magic[8];
  int32_t header[2] = {0, 0};
  int64_t m = 0, plen = 0;
  bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
            std::fread(&plen, sizeof(plen), 1, f) == 1 && plen >= 0 && plen <= (1 << 16);
  if (ok) {
    got.prefix.resize(static_cast<size_t>(plen));
    ok = std::fread(&got.prefix[0], 1, got.prefix.size(), f) == got.prefix.size() &&
         std::fread(got.size, sizeof(got.size), 1, f) == 1 &&
         std::fread(got.mtime, sizeof(got.mtime), 1, f) == 1 && got == want;
  }
  ok = ok && std::fread(header, sizeof(header), 1, f) == 1 &&
       std::fread(&m, sizeof(m), 1, f) == 1 && header[0] > 0 && header[1] > 0 &&
       header[1] <= 32 && m >= 0 && m <= INT_MAX;
  if (ok) {
    std::error_code ec;
    const uintmax_t fileSize = fs::file_size(path, ec);
    const long pos = std::ftell(f);
    const uint64_t bytes = sizeof(int) * (static_cast<uint64_t>(header[0]) + 1 +
                                          static_cast<uint64_t>(m) * (1 + header[1]));
    ok = !ec && pos >= 0 && fileSize == static_cast<uintmax_t>(pos) + bytes;
  }
  if (ok) {
    graph.numberOfNodes = header[0];
    graph.numberOfObjectives = header[1];
    graph.rowPtr.resize(static_cast<size_t>(header[0]) + 1);
    graph.colInd.resize(static_cast<size_t>(m));
    graph.weights.resize(static_cast<size_t>(m) * header[1]);
    ok = std::fread(graph.rowPtr.data(), sizeof(int), graph.rowPtr.size(), f) ==
             graph.rowPtr.size() &&
         std::fread(graph.colInd.data(), sizeof(int), graph.colInd.size(), f) ==
             graph.colInd.size() &&
         std::fread(graph.weights.data(), sizeof(int), graph.weights.size(), f) ==
             graph.weights.size() &&
         std::fgetc(f) == EOF && valid_arrays(graph);
  }
  std::fclose(f);
  if (!ok) graph = CsrGraph();
  return ok;
}

bool loadCsrGraph(const std::string &prefix, CsrGraph &graph,
                  const std::string &cachePath) {
  if (!cachePath.empty()) {
    if (loadCsrGraphBinary(cachePath, graph, prefix)) return true;
    std::error_code ec;
    if (fs::exists(cachePath, ec)) {
      std::cout << "Note: binary cache " << cachePath
                << " does not match the text graph " << prefix
                << " (another graph, changed files or an old format); "
                   "rebuilding it.\n";
    }
  }
  if (!readCsrGraph(prefix, graph)) return false;
  if (!cachePath.empty() && !saveCsrGraphBinary(cachePath, graph, prefix)) {
    std::cout << "Warning: could not write binary cache " << cachePath << "\n";
  }
  return true;
}

bool readChangeBatch(const std::string &insertPath, const std::string &deletePath,
                     int K, int n, ChangeBatch &batch) {
  batch = ChangeBatch();
  batch.numberOfObjectives = K;
  auto inRange = [&](long long x) { return x >= 0 && x < n; };
  std::string text;
  if (!slurp(insertPath, text)) {
    std::cout << "Error: Could not open insert file: " << insertPath << "\n";
    return false;
  }
  {
    Cursor c(text);
    std::vector<long long> tok;
    for (;;) {
      c.blanks();
      tok.clear();
      while (!c.eol()) {
        long long v = 0;
        if (!c.number(v)) {
          std::cout << "Error: Invalid insert line at " << at(insertPath, c.begin, c.p) << ".\n";
          return false;
        }
        tok.push_back(v);
        c.blanks();
      }
      if (!tok.empty()) {
        if (static_cast<int>(tok.size()) < 2 + K) {
          std::cout << "Error: Invalid insert line at " << at(insertPath, c.begin, c.p) << ".\n";
          return false;
        }
        if (!inRange(tok[0]) || !inRange(tok[1])) {
          std::cout << "Error: Inserted edge endpoint out of range at "
                    << at(insertPath, c.begin, c.p) << ".\n";
          return false;
        }
        for (int k = 0; k < K; ++k) {
          if (!valid_weight(tok[2 + k])) {
            std::cout << "Error: Inserted edge weight " << tok[2 + k] << " out of range [1, "
                      << INT_MAX << "] at " << at(insertPath, c.begin, c.p) << ".\n";
            return false;
          }
        }
        batch.insertFrom.push_back(static_cast<int>(tok[0]));
        batch.insertTo.push_back(static_cast<int>(tok[1]));
        for (int k = 0; k < K; ++k) batch.insertWeights.push_back(static_cast<int>(tok[2 + k]));
      }
      if (c.p >= c.end) break;
      ++c.p;
    }
  }
  if (!slurp(deletePath, text)) {
    std::cout << "Error: Could not open delete file: " << deletePath << "\n";
    return false;
  }
  Cursor c(text);
  for (;;) {
    c.blanks();
    long long pair[2] = {0, 0};
    int count = 0;
    while (!c.eol()) {
      long long v = 0;
      if (!c.number(v)) {
        std::cout << "Error: Invalid delete line at " << at(deletePath, c.begin, c.p) << ".\n";
        return false;
      }
      if (count < 2) pair[count] = v;
      ++count;
      c.blanks();
    }
    if (count >= 2) {  // shorter lines are ignored
      if (!inRange(pair[0]) || !inRange(pair[1])) {
        std::cout << "Error: Deleted edge endpoint out of range at "
                  << at(deletePath, c.begin, c.p) << ".\n";
        return false;
      }
      batch.deleteFrom.push_back(static_cast<int>(pair[0]));
      batch.deleteTo.push_back(static_cast<int>(pair[1]));
    }
    if (c.p >= c.end) break;
    ++c.p;
  }
  return true;
}

bool writeChangeBatch(const ChangeBatch &batch, const std::string &insertPath,
                      const std::string &deletePath) {
  Writer ins(insertPath), del(deletePath);
  if (!ins.ok() || !del.ok()) return false;
  const int K = batch.numberOfObjectives;
  for (int i = 0; i < batch.numberOfInserts(); ++i) {
    ins.num(batch.insertFrom[i]);
    ins.ch(' ');
    ins.num(batch.insertTo[i]);
    for (int k = 0; k < K; ++k) {
      ins.ch(' ');
      ins.num(batch.insertWeights[static_cast<size_t>(i) * K + k]);
    }
    ins.ch('\n');
  }
  for (int i = 0; i < batch.numberOfDeletes(); ++i) {
    del.num(batch.deleteFrom[i]);
    del.ch(' ');
    del.num(batch.deleteTo[i]);
    del.ch('\n');
  }
  return ins.close() && del.close();
}

bool applyChangeBatch(const CsrGraph &original, ChangeBatch &batch, CsrGraph &updated) {
  const int n = original.numberOfNodes;
  const int K = original.numberOfObjectives;
  if (batch.numberOfObjectives != K) {
    std::cout << "Error: change batch has " << batch.numberOfObjectives
              << " objectives, graph has " << K << ".\n";
    return false;
  }
  if (K > 32) {
    std::cout << "Error: at most 32 objectives are supported.\n";
    return false;
  }
  const int nIns = batch.numberOfInserts(), nDel = batch.numberOfDeletes();
  // Changes bucketed by source row (counting sort keeps batch order).
  std::vector<int> delHead(n + 1, 0), insHead(n + 1, 0);
  for (int i = 0; i < nDel; ++i) ++delHead[batch.deleteFrom[i] + 1];
  for (int i = 0; i < nIns; ++i) ++insHead[batch.insertFrom[i] + 1];
  for (int u = 0; u < n; ++u) {
    delHead[u + 1] += delHead[u];
    insHead[u + 1] += insHead[u];
  }
  std::vector<int> delIdx(nDel), insIdx(nIns);
  {
    std::vector<int> fillD(delHead.begin(), delHead.end() - 1);
    std::vector<int> fillI(insHead.begin(), insHead.end() - 1);
    for (int i = 0; i < nDel; ++i) delIdx[fillD[batch.deleteFrom[i]]++] = i;
    for (int i = 0; i < nIns; ++i) insIdx[fillI[batch.insertFrom[i]]++] = i;
  }
  // New contents of every touched row: (neighbour, weight source) where the
  // source is an original edge id (>= 0) or ~insertion index (< 0).
  struct Slot {
    int v;
    long long src;
  };
  std::vector<int> touched;
  std::vector<int> rowStart{0};
  std::vector<Slot> rows;
  std::vector<int> rowOf(n, -1);
  std::vector<Slot> work;
  std::vector<char> alive;
  for (int u = 0; u < n; ++u) {
    if (delHead[u] == delHead[u + 1] && insHead[u] == insHead[u + 1]) continue;
    work.clear();
    alive.clear();
    for (int e = original.rowPtr[u]; e < original.rowPtr[u + 1]; ++e) {
      work.push_back({original.colInd[e], e});
      alive.push_back(1);
    }
    for (int j = delHead[u]; j < delHead[u + 1]; ++j) {
      const int v = batch.deleteTo[delIdx[j]];
      for (size_t x = 0; x < work.size(); ++x) {
        if (alive[x] && work[x].v == v) {
          alive[x] = 0;
          break;
        }
      }
    }
    for (int j = insHead[u]; j < insHead[u + 1]; ++j) {
      const int i = insIdx[j];
      const int v = batch.insertTo[i];
      size_t x = 0;
      for (; x < work.size(); ++x) {
        if (alive[x] && work[x].v == v) break;
      }
      if (x < work.size()) {
        work[x].src = ~static_cast<long long>(i);
      } else {
        work.push_back({v, ~static_cast<long long>(i)});
        alive.push_back(1);
      }
    }
    rowOf[u] = static_cast<int>(touched.size());
    touched.push_back(u);
    for (size_t x = 0; x < work.size(); ++x) {
      if (alive[x]) rows.push_back(work[x]);
    }
    rowStart.push_back(static_cast<int>(rows.size()));
  }
  updated = CsrGraph();
  updated.numberOfNodes = n;
  updated.numberOfObjectives = K;
  updated.rowPtr.assign(n + 1, 0);
  for (int u = 0; u < n; ++u) {
    const int r = rowOf[u];
    const int deg = r >= 0 ? rowStart[r + 1] - rowStart[r]
                           : original.rowPtr[u + 1] - original.rowPtr[u];
    updated.rowPtr[u + 1] = updated.rowPtr[u] + deg;
  }
  const size_t m = static_cast<size_t>(updated.rowPtr[n]);
  updated.colInd.resize(m);
  updated.weights.resize(m * K);
  for (int u = 0; u < n; ++u) {
    size_t out = static_cast<size_t>(updated.rowPtr[u]);
    const int r = rowOf[u];
    if (r < 0) {
      const int b = original.rowPtr[u], e = original.rowPtr[u + 1];
      std::copy(original.colInd.begin() + b, original.colInd.begin() + e,
                updated.colInd.begin() + out);
      std::copy(original.weights.begin() + static_cast<size_t>(b) * K,
                original.weights.begin() + static_cast<size_t>(e) * K,
                updated.weights.begin() + out * K);
      continue;
    }
    for (int x = rowStart[r]; x < rowStart[r + 1]; ++x, ++out) {
      updated.colInd[out] = rows[x].v;
      const int *w = rows[x].src >= 0
                         ? &original.weights[static_cast<size_t>(rows[x].src) * K]
                         : &batch.insertWeights[static_cast<size_t>(~rows[x].src) * K];
      std::copy(w, w + K, updated.weights.begin() + out * K);
    }
  }
  // Weight increases: final first-occurrence weights of (u,v) against the
  // original first occurrence.
  batch.weightIncreaseMask.assign(nIns, 0u);
  for (int i = 0; i < nIns; ++i) {
    const int u = batch.insertFrom[i], v = batch.insertTo[i];
    int before = -1, after = -1;
    for (int e = original.rowPtr[u]; e < original.rowPtr[u + 1]; ++e) {
      if (original.colInd[e] == v) {
        before = e;
        break;
      }
    }
    if (before < 0) continue;
    for (int e = updated.rowPtr[u]; e < updated.rowPtr[u + 1]; ++e) {
      if (updated.colInd[e] == v) {
        after = e;
        break;
      }
    }
    if (after < 0) continue;
    unsigned mask = 0;
    for (int k = 0; k < K; ++k) {
      if (updated.weight(after, k) > original.weight(before, k)) mask |= 1u << k;
    }
    batch.weightIncreaseMask[i] = mask;
  }
  return true;
}

void transposeCsrGraph(const CsrGraph &g, CsrGraph &rev) {
  const int n = g.numberOfNodes, K = g.numberOfObjectives;
  const size_t m = static_cast<size_t>(g.numberOfEdges());
  rev = CsrGraph();
  rev.numberOfNodes = n;
  rev.numberOfObjectives = K;
  rev.rowPtr.assign(n + 1, 0);
  for (size_t e = 0; e < m; ++e) ++rev.rowPtr[g.colInd[e] + 1];
  for (int v = 0; v < n; ++v) rev.rowPtr[v + 1] += rev.rowPtr[v];
  rev.colInd.resize(m);
  rev.weights.resize(m * K);
  std::vector<int> fill(rev.rowPtr.begin(), rev.rowPtr.end() - 1);
  for (int u = 0; u < n; ++u) {
    for (int e = g.rowPtr[u]; e < g.rowPtr[u + 1]; ++e) {
      const size_t pos = static_cast<size_t>(fill[g.colInd[e]]++);
      rev.colInd[pos] = u;
      for (int k = 0; k < K; ++k) rev.weights[pos * K + k] = g.weight(e, k);
    }
  }
}

namespace {

[[gnu::cold]] bool bad_line(const char *what, const std::string &path,
                            const char *b, const char *p) {
  std::cout << "Error: " << what << " at " << at(path, b, p) << ".\n";
  return false;
}

}  // namespace

bool readDistances(const std::string &path, int n, std::vector<long long> &dist) {
  std::string text;
  if (!slurp(path, text)) {
    std::cout << "Error: Could not open distances file: " << path << "\n";
    return false;
  }
  dist.assign(n, -1);  // -1: not listed yet
  long long listed = 0;
  Cursor c(text);
  for (;;) {
    c.space();
    if (c.p >= c.end) break;
    long long v = 0, d = 0;
    if (!c.number(v)) return bad_line("Invalid line in distances file", path, c.begin, c.p);
    c.blanks();
    if (c.end - c.p >= 3 && std::memcmp(c.p, "INF", 3) == 0) {
      c.p += 3;
      d = DISTANCE_INF;
    } else if (!c.number(d)) {
      return bad_line("Invalid line in distances file", path, c.begin, c.p);
    }
    if (v < 0 || v >= n) return bad_line("Vertex ID out of range in distances file", path, c.begin, c.p);
    if (d < 0) return bad_line("Negative distance", path, c.begin, c.p);
    ++listed;
    dist[v] = d;
  }
  if (listed != n || std::find(dist.begin(), dist.end(), -1LL) != dist.end()) {
    std::cout << "Error: " << path << " must list each of the " << n
              << " vertices exactly once (" << listed << " lines).\n";
    return false;
  }
  return true;
}

bool readParents(const std::string &path, int n, std::vector<int> &parent) {
  std::string text;
  if (!slurp(path, text)) {
    std::cout << "Error: Could not open SSSP tree file: " << path << "\n";
    return false;
  }
  parent.assign(n, -2);  // -2: not listed yet
  long long listed = 0;
  Cursor c(text);
  for (;;) {
    c.space();
    if (c.p >= c.end) break;
    long long v = 0, p = 0;
    if (!c.number(v)) return bad_line("Invalid line in SSSP tree file", path, c.begin, c.p);
    c.blanks();
    if (!c.number(p)) return bad_line("Invalid line in SSSP tree file", path, c.begin, c.p);
    if (v < 0 || v >= n || p < -1 || p >= n) {
      return bad_line("Vertex ID out of range in SSSP tree file", path, c.begin, c.p);
    }
    ++listed;
    parent[v] = static_cast<int>(p);
  }
  if (listed != n || std::find(parent.begin(), parent.end(), -2) != parent.end()) {
    std::cout << "Error: " << path << " must list each of the " << n
              << " vertices exactly once (" << listed << " lines).\n";
    return false;
  }
  return true;
}

bool writeDistances(const std::string &path, const std::vector<long long> &dist) {
  Writer w(path);
  if (!w.ok()) {
    std::cout << "Error: Could not write distances file: " << path << "\n";
    return false;
  }
  for (size_t v = 0; v < dist.size(); ++v) {
    w.num(static_cast<long long>(v));
    w.ch(' ');
    if (dist[v] >= DISTANCE_INF / 2) {
      w.str("INF");
    } else {
      w.num(dist[v]);
    }
    w.ch('\n');
  }
  return w.close();
}

bool writeParents(const std::string &path, const std::vector<int> &parent) {
  Writer w(path);
  if (!w.ok()) {
    std::cout << "Error: Could not write SSSP tree file: " << path << "\n";
    return false;
  }
  for (size_t v = 0; v < parent.size(); ++v) {
    w.num(static_cast<long long>(v));
    w.ch(' ');
    w.num(parent[v]);
    w.ch('\n');
  }
  return w.close();
}
