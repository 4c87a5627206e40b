// csr_graph.cpp -- the host side of the file contract: CSR text, the MOSPCSR2
// binary cache, change batches, tree files, and the host batch application.
//
// Formats, the syntax accepted and the error messages are the reference's
// (src/csrGraph.cu: readers 186-472, 626-725, 943-1030; writers 474-501,
// 1032-1064; binary cache 503-620; applyChangeBatch 727-894), so that files
// and diagnostics are interchangeable.  The code is built from this
// project's own pieces:
//   * MappedFile: read-only mmap of an input, no copy into a string;
//   * every file is read line by line (memchr finds the line, LineTokens
//     takes its fields in order) and line numbers are counted only when an
//     error is reported;
//   * parse_in_parts: large files are cut at line boundaries and parsed by
//     one thread per part;
//   * write_lines: writers format consecutive blocks of lines on worker
//     threads into separate buffers, then write them in order;
//   * the binary cache is written from one assembled header plus the three
//     arrays, and read through a bounds-checked cursor over a mapping;
//   * applyChangeBatch works per (u, v) pair -- the formulation of the GPU
//     apply in device_graph.hip -- instead of replaying the changes one by
//     one.
#include "csrGraph.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <charconv>
#include <chrono>
#include <climits>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <iostream>
#include <string>
#include <thread>
#include <vector>

namespace fs = std::filesystem;

namespace {

// ---------------------------------------------------------------- inputs --

// A read-only mapping of a whole file ([begin, end) is empty for an empty
// file or one whose size fstat cannot report).
class MappedFile {
 public:
  MappedFile() = default;
  MappedFile(const MappedFile &) = delete;
  MappedFile &operator=(const MappedFile &) = delete;
  ~MappedFile() {
    if (base_ != nullptr) {
      ::munmap(base_, size_);
    }
  }

  bool open(const std::string &path) {
    const int fd = ::open(path.c_str(), O_RDONLY | O_CLOEXEC);
    if (fd < 0) {
      return false;
    }
    struct stat st {};
    bool ok = ::fstat(fd, &st) == 0;
    if (ok && st.st_size > 0) {
      size_ = static_cast<size_t>(st.st_size);
      void *p = ::mmap(nullptr, size_, PROT_READ, MAP_PRIVATE, fd, 0);
      if (p == MAP_FAILED) {
        size_ = 0;
        ok = false;
      } else {
        base_ = p;
        ::madvise(base_, size_, MADV_WILLNEED);
      }
    }
    ::close(fd);
    return ok;
  }

  const char *begin() const { return base_ != nullptr ? static_cast<const char *>(base_) : kEmpty; }
  const char *end() const { return begin() + size_; }
  size_t size() const { return size_; }

 private:
  static constexpr const char *kEmpty = "";
  void *base_ = nullptr;
  size_t size_ = 0;
};

// "<path>:<line>" of position pos of a file that starts at origin.
std::string where(const std::string &path, const char *origin, const char *pos) {
  return path + ":" + std::to_string(1 + std::count(origin, pos, '\n'));
}

bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r'; }

// The tokens of one line [p, eol), in order: next() stores the next integer
// (or DISTANCE_INF for the token "INF" where the caller allows it) and
// returns 1; 0 at the end of the line; -1 when a token starts no integer
// (bad() is its position).
class LineTokens {
 public:
  LineTokens(const char *p, const char *eol) : p_(p), eol_(eol) {}
  int next(long long &value, bool infAllowed = false) {
    while (p_ < eol_ && is_blank(*p_)) {
      ++p_;
    }
    if (p_ >= eol_) {
      return 0;
    }
    if (infAllowed && eol_ - p_ >= 3 && std::memcmp(p_, "INF", 3) == 0) {
      value = DISTANCE_INF;
      p_ += 3;
      return 1;
    }
    const auto r = std::from_chars(p_, eol_, value);
    if (r.ec != std::errc()) {
      return -1;
    }
    p_ = r.ptr;
    return 1;
  }
  const char *bad() const { return p_; }

 private:
  const char *p_;
  const char *eol_;
};

// Calls f(lineBegin, lineEnd) for every line of [from, to) (the last one may
// lack its newline); stops at the first false.
template <typename F>
bool each_line(const char *from, const char *to, F &&f) {
  while (from < to) {
    const void *nl = std::memchr(from, '\n', static_cast<size_t>(to - from));
    const char *eol = nl != nullptr ? static_cast<const char *>(nl) : to;
    if (!f(from, eol)) {
      return false;
    }
    from = eol + 1;
  }
  return true;
}

// ------------------------------------------------- parallel text parsing --

// Files above this size are parsed in parts, one thread each.
constexpr size_t kPartBytes = size_t{4} << 20;

unsigned worker_threads() {
  return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

// Cuts of [begin, end) into parts that end at line boundaries.
std::vector<const char *> part_cuts(const char *begin, const char *end) {
  const size_t bytes = static_cast<size_t>(end - begin);
  const size_t parts = std::max<size_t>(1, std::min<size_t>(worker_threads(), bytes / kPartBytes));
  std::vector<const char *> cuts(1, begin);
  for (size_t i = 1; i < parts; ++i) {
    const char *c = std::max(cuts.back(), begin + bytes * i / parts);
    const void *nl = std::memchr(c, '\n', static_cast<size_t>(end - c));
    cuts.push_back(nl != nullptr ? static_cast<const char *>(nl) + 1 : end);
  }
  cuts.push_back(end);
  return cuts;
}

struct PartResult {
  std::vector<int> values;
  long long lines = 0;   // non-empty lines
  std::string error;     // first error of the part ("" if none)
};

// parse(part, from, to) on every part in parallel, then the parts' values
// in order into `out`.  Each part stops at its first error, so the error of
// the first failing part is the file's first error, as in one pass.
template <typename F>
bool parse_in_parts(const MappedFile &file, std::vector<int> &out, long long &lines,
                    std::string &error, F &&parse) {
  const std::vector<const char *> cuts = part_cuts(file.begin(), file.end());
  const size_t count = cuts.size() - 1;
  std::vector<PartResult> parts(count);
  {
    std::vector<std::thread> threads;
    for (size_t i = 1; i < count; ++i) {
      threads.emplace_back([&, i] { parse(parts[i], cuts[i], cuts[i + 1]); });
    }
    parse(parts[0], cuts[0], cuts[1]);
    for (std::thread &t : threads) {
      t.join();
    }
  }
  for (const PartResult &p : parts) {
    if (!p.error.empty()) {
      error = p.error;
      return false;
    }
  }
  std::vector<size_t> offset(count + 1, 0);
  lines = 0;
  for (size_t i = 0; i < count; ++i) {
    offset[i + 1] = offset[i] + parts[i].values.size();
    lines += parts[i].lines;
  }
  out.resize(offset[count]);
  std::vector<std::thread> threads;
  for (size_t i = 0; i < count; ++i) {
    threads.emplace_back([&, i] {
      std::copy(parts[i].values.begin(), parts[i].values.end(),
                out.begin() + static_cast<std::ptrdiff_t>(offset[i]));
      std::vector<int>().swap(parts[i].values);
    });
  }
  for (std::thread &t : threads) {
    t.join();
  }
  return true;
}

bool valid_weight(long long w) { return w >= 1 && w <= INT_MAX; }

// RowPtr / ColInd: whitespace-separated integers in [0, INT_MAX].
bool read_id_file(const std::string &path, std::vector<int> &out, std::string &error) {
  MappedFile file;
  if (!file.open(path)) {
    error = "could not read " + path;
    return false;
  }
  const char *origin = file.begin();
  long long lines = 0;
  return parse_in_parts(file, out, lines, error, [&](PartResult &part, const char *from,
                                                     const char *to) {
    part.values.reserve(static_cast<size_t>(to - from) / 4);
    each_line(from, to, [&](const char *b, const char *e) {
      LineTokens tokens(b, e);
      long long value = 0;
      for (int got; (got = tokens.next(value)) != 0;) {
        if (got < 0) {
          part.error = "invalid token at " + where(path, origin, tokens.bad());
          return false;
        }
        if (value < 0 || value > INT_MAX) {
          part.error = "value " + std::to_string(value) + " out of range [0, " +
                       std::to_string(INT_MAX) + "] at " + where(path, origin, tokens.bad());
          return false;
        }
        part.values.push_back(static_cast<int>(value));
      }
      return true;
    });
  });
}

// Values: one line per edge with K weights in [1, INT_MAX]; K is the field
// count of the first non-empty line.
bool read_value_file(const std::string &path, std::vector<int> &weights, int &K, long long &lines,
                     std::string &error) {
  MappedFile file;
  if (!file.open(path)) {
    error = "Could not read CSR values file " + path + ".";
    return false;
  }
  const char *origin = file.begin();
  // One line: its weights are appended (and counted) as they are read; a
  // line with a field count other than k (0: this line sets it) fails.
  auto line = [&](PartResult &part, int &k, const char *b, const char *e) {
    LineTokens tokens(b, e);
    long long value = 0;
    int count = 0;
    for (int got; (got = tokens.next(value)) != 0; ++count) {
      if (got < 0) {
        part.error = "Invalid token in CSR values file at " + where(path, origin, tokens.bad()) + ".";
        return false;
      }
      if (!valid_weight(value)) {
        part.error = "CSR weight " + std::to_string(value) + " out of range [1, " +
                     std::to_string(INT_MAX) + "] at " + where(path, origin, tokens.bad()) + ".";
        return false;
      }
      part.values.push_back(static_cast<int>(value));
    }
    if (count == 0) {
      return true;
    }
    if (k == 0) {
      k = count;
    } else if (count != k) {
      part.error = "Inconsistent number of objectives at " + where(path, origin, e) + ".";
      return false;
    }
    ++part.lines;
    return true;
  };
  // K from the first non-empty line, read alone (its errors are the file's
  // first ones); every part then checks its lines against it.
  K = 0;
  {
    PartResult first;
    each_line(file.begin(), file.end(),
              [&](const char *b, const char *e) { return line(first, K, b, e) && K == 0; });
    if (!first.error.empty()) {
      error = first.error;
      return false;
    }
  }
  const int k0 = K;
  return parse_in_parts(file, weights, lines, error, [&](PartResult &part, const char *from,
                                                         const char *to) {
    part.values.reserve(static_cast<size_t>(to - from) / 3);
    int k = k0;
    each_line(from, to, [&](const char *b, const char *e) { return line(part, k, b, e); });
  });
}

// --------------------------------------------------------------- outputs --

bool make_parent_dirs(const std::string &path) {
  const fs::path parent = fs::path(path).parent_path();
  std::error_code ec;
  if (!parent.empty()) {
    fs::create_directories(parent, ec);
  }
  return !ec;
}

bool write_all(int fd, const char *p, size_t n, off_t at) {
  while (n > 0) {
    const ssize_t w = ::pwrite(fd, p, n, at);
    if (w <= 0) {
      return false;
    }
    p += w;
    n -= static_cast<size_t>(w);
    at += w;
  }
  return true;
}

// Writes lines [0, count) to `path`: format(i, out) stores line i (newline
// included, at most maxLine bytes) at out and returns its end.  In rounds of
// up to 64 MiB of text, worker threads format consecutive blocks of lines
// into their own buffers, which are then written in order.
template <typename F>
bool write_lines(const std::string &path, size_t count, size_t maxLine, F &&format) {
  make_parent_dirs(path);
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0) {
    return false;
  }
  constexpr size_t kMinLines = size_t{1} << 16;  // fewer: one thread
  const size_t threads = count >= 2 * kMinLines ? worker_threads() : 1;
  const size_t block = std::max<size_t>(
      1, std::min<size_t>(count, (size_t{64} << 20) / (threads * std::max<size_t>(maxLine, 1))));
  std::vector<std::vector<char>> buf(threads, std::vector<char>(block * maxLine));
  std::vector<size_t> used(threads, 0);
  bool ok = true;
  off_t at = 0;
  for (size_t base = 0; base < count && ok; base += block * threads) {
    auto fill = [&](size_t t) {
      const size_t lo = std::min(count, base + t * block), hi = std::min(count, lo + block);
      char *o = buf[t].data();
      for (size_t i = lo; i < hi; ++i) {
        o = format(i, o);
      }
      used[t] = static_cast<size_t>(o - buf[t].data());
    };
    std::vector<std::thread> workers;
    for (size_t t = 1; t < threads; ++t) workers.emplace_back(fill, t);
    fill(0);
    for (std::thread &w : workers) w.join();
    for (size_t t = 0; t < threads && ok; ++t) {
      ok = write_all(fd, buf[t].data(), used[t], at);
      at += static_cast<off_t>(used[t]);
    }
  }
  return ::close(fd) == 0 && ok;
}

char *put_int(char *out, long long v) { return std::to_chars(out, out + 24, v).ptr; }

// ----------------------------------------------------------- binary cache --

constexpr char kMagic[8] = {'M', 'O', 'S', 'P', 'C', 'S', 'R', '2'};
constexpr const char *kSuffixes[3] = {"RowPtr.txt", "ColInd.txt", "Values.txt"};

// What a cache was built from: the canonical text prefix, and the size and
// modification time (nanoseconds of the filesystem clock) of its three files.
struct SourceStamp {
  std::string prefix;
  int64_t size[3] = {0, 0, 0};
  int64_t mtime[3] = {0, 0, 0};
};

bool same_source(const SourceStamp &a, const SourceStamp &b) {
  return a.prefix == b.prefix && std::memcmp(a.size, b.size, sizeof(a.size)) == 0 &&
         std::memcmp(a.mtime, b.mtime, sizeof(a.mtime)) == 0;
}

bool stamp_of(const std::string &prefix, SourceStamp &s) {
  std::error_code ec;
  s.prefix = fs::weakly_canonical(prefix, ec).string();
  for (int i = 0; i < 3 && !ec; ++i) {
    const std::string file = prefix + kSuffixes[i];
    s.size[i] = static_cast<int64_t>(fs::file_size(file, ec));
    if (ec) break;
    const auto t = fs::last_write_time(file, ec);
    s.mtime[i] = static_cast<int64_t>(
        std::chrono::duration_cast<std::chrono::nanoseconds>(t.time_since_epoch()).count());
  }
  return !ec;
}

template <typename T>
void append_raw(std::string &buf, const T &value) {
  buf.append(reinterpret_cast<const char *>(&value), sizeof(T));
}

// Bounds-checked reads from a mapped byte range.
struct ByteCursor {
  const char *p, *end;
  bool take(void *dst, size_t bytes) {
    if (static_cast<size_t>(end - p) < bytes) return false;
    std::memcpy(dst, p, bytes);
    p += bytes;
    return true;
  }
  size_t left() const { return static_cast<size_t>(end - p); }
};

// The arrays of a cached graph obey the rules readCsrGraph enforces: row
// pointers from 0 to m and non-decreasing, columns in [0, n), weights >= 1.
bool arrays_valid(const CsrGraph &g) {
  const int n = g.numberOfNodes;
  const size_t m = g.colInd.size();
  if (g.rowPtr.size() != static_cast<size_t>(n) + 1 || g.rowPtr[0] != 0 ||
      static_cast<size_t>(g.rowPtr[n]) != m ||
      g.weights.size() != m * static_cast<size_t>(g.numberOfObjectives)) {
    return false;
  }
  const bool rows = std::is_sorted(g.rowPtr.begin(), g.rowPtr.end());
  const bool cols = std::all_of(g.colInd.begin(), g.colInd.end(),
                                [n](int v) { return static_cast<unsigned>(v) < static_cast<unsigned>(n); });
  const bool weights = std::all_of(g.weights.begin(), g.weights.end(), [](int w) { return w >= 1; });
  return rows && cols && weights;
}

// --------------------------------------------------------- tree readers --

[[gnu::cold]] bool line_error(const char *what, const std::string &path, const char *origin,
                              const char *pos) {
  std::cout << "Error: " << what << " at " << where(path, origin, pos) << ".\n";
  return false;
}

// "v x" pairs (any number per line): on_pair(v, x) checks and stores one.
// Each vertex must be listed exactly once; `unset` marks unlisted ones.
template <typename T, typename F>
bool read_pairs(const std::string &path, const char *openError, const char *lineError,
                bool infValues, int n, T unset, std::vector<T> &out, F &&on_pair) {
  MappedFile file;
  if (!file.open(path)) {
    std::cout << "Error: " << openError << path << "\n";
    return false;
  }
  out.assign(static_cast<size_t>(n), unset);
  long long listed = 0;
  const char *origin = file.begin();
  const bool ok = each_line(file.begin(), file.end(), [&](const char *b, const char *e) {
    // Vertex, then its value on the same line; any number of pairs a line.
    LineTokens tokens(b, e);
    for (;;) {
      long long v = 0, x = 0;
      const int first = tokens.next(v);
      if (first == 0) {
        return true;
      }
      if (first < 0 || tokens.next(x, infValues) != 1) {
        return line_error(lineError, path, origin, tokens.bad());
      }
      if (!on_pair(v, x, origin, tokens.bad())) {
        return false;
      }
      ++listed;
    }
  });
  if (!ok) {
    return false;
  }
  if (listed != n || std::find(out.begin(), out.end(), unset) != out.end()) {
    std::cout << "Error: " << path << " must list each of the " << n
              << " vertices exactly once (" << listed << " lines).\n";
    return false;
  }
  return true;
}

}  // namespace

// ================================================================ public ==

bool runConcurrently(const std::vector<std::function<bool()>> &jobs) {
  std::vector<unsigned char> ok(jobs.size(), 0);
  {
    std::vector<std::thread> threads;
    threads.reserve(jobs.size());
    for (size_t i = 0; i < jobs.size(); ++i) {
      threads.emplace_back([&ok, &jobs, i] { ok[i] = jobs[i]() ? 1 : 0; });
    }
    for (std::thread &t : threads) t.join();
  }
  return std::find(ok.begin(), ok.end(), 0) == ok.end();
}

bool parseInteger(const std::string &text, long long minimum, long long maximum,
                  long long &value) {
  // A bare decimal: '-' allowed, '+' and surrounding blanks not.
  if (text.empty() || !(text[0] == '-' || (text[0] >= '0' && text[0] <= '9'))) {
    return false;
  }
  const char *end = text.data() + text.size();
  const auto r = std::from_chars(text.data(), end, value);
  return r.ec == std::errc() && r.ptr == end && minimum <= value && value <= maximum;
}

bool readCsrGraph(const std::string &prefix, CsrGraph &graph) {
  graph = CsrGraph();
  std::string rowError, colError, valError;
  bool rowsRead = false, colsRead = false, valsRead = false;
  int K = 0;
  long long valueLines = 0;
  runConcurrently({
      [&] { return rowsRead = read_id_file(prefix + "RowPtr.txt", graph.rowPtr, rowError); },
      [&] { return colsRead = read_id_file(prefix + "ColInd.txt", graph.colInd, colError); },
      [&] {
        return valsRead = read_value_file(prefix + "Values.txt", graph.weights, K, valueLines,
                                          valError);
      },
  });
  // The checks and their order are the reference's (csrGraph.cu:405-472):
  // row pointers, then columns, then values.
  auto reject = [&](const std::string &message) {
    std::cout << "Error: " << message << "\n";
    graph = CsrGraph();
    return false;
  };
  if (!rowsRead) return reject("CSR row pointers: " + rowError);
  if (graph.rowPtr.size() < 2) return reject("Could not read CSR row pointers: " + prefix + "RowPtr.txt");
  const int n = static_cast<int>(graph.rowPtr.size()) - 1;
  if (graph.rowPtr.front() != 0) return reject("CSR row pointers must start at 0.");
  if (!std::is_sorted(graph.rowPtr.begin(), graph.rowPtr.end())) {
    return reject("CSR row pointers are not monotone.");
  }
  if (!colsRead) return reject("CSR column indices: " + colError);
  if (static_cast<long long>(graph.colInd.size()) != graph.rowPtr.back()) {
    return reject("CSR column index file missing or size mismatch.");
  }
  if (!std::all_of(graph.colInd.begin(), graph.colInd.end(), [n](int v) { return v < n; })) {
    return reject("CSR column index out of range.");  // ids are >= 0 by the reader
  }
  if (!valsRead) return reject(valError);
  if (valueLines != graph.rowPtr.back()) return reject("values size mismatch.");
  graph.numberOfNodes = n;
  graph.numberOfObjectives = K;
  return true;
}

bool writeCsrGraph(const std::string &prefix, const CsrGraph &graph) {
  const int K = graph.numberOfObjectives;
  const std::vector<int> &row = graph.rowPtr, &col = graph.colInd, &w = graph.weights;
  bool ok[3] = {false, false, false};
  runConcurrently({
      [&] {
        return ok[0] = write_lines(prefix + "RowPtr.txt", row.size(), 12, [&](size_t i, char *o) {
                 o = put_int(o, row[i]);
                 *o++ = '\n';
                 return o;
               });
      },
      [&] {
        return ok[1] = write_lines(prefix + "ColInd.txt", col.size(), 12, [&](size_t i, char *o) {
                 o = put_int(o, col[i]);
                 *o++ = '\n';
                 return o;
               });
      },
      [&] {
        return ok[2] = write_lines(prefix + "Values.txt", col.size(), 12 * static_cast<size_t>(K) + 1,
                                   [&](size_t e, char *o) {
                                     const int *x = &w[e * static_cast<size_t>(K)];
                                     for (int k = 0; k < K; ++k) {
                                       if (k > 0) *o++ = ' ';
                                       o = put_int(o, x[k]);
                                     }
                                     *o++ = '\n';
                                     return o;
                                   });
      },
  });
  if (!(ok[0] && ok[1] && ok[2])) {
    std::cout << "Error: Could not write CSR files: " << prefix << "\n";
    return false;
  }
  return true;
}

// Cache layout (MOSPCSR2, little-endian as written by the host):
//   magic[8] | int64 prefix length | prefix bytes | int64 size[3] |
//   int64 mtime[3] | int32 n, int32 K | int64 m | rowPtr[n+1] | colInd[m] |
//   weights[m*K] (edge-major)
bool saveCsrGraphBinary(const std::string &path, const CsrGraph &graph,
                        const std::string &sourcePrefix) {
  SourceStamp src;
  if (!stamp_of(sourcePrefix, src)) {
    return false;
  }
  std::string head(kMagic, sizeof(kMagic));
  append_raw(head, static_cast<int64_t>(src.prefix.size()));
  head += src.prefix;
  for (int64_t s : src.size) append_raw(head, s);
  for (int64_t t : src.mtime) append_raw(head, t);
  append_raw(head, static_cast<int32_t>(graph.numberOfNodes));
  append_raw(head, static_cast<int32_t>(graph.numberOfObjectives));
  append_raw(head, static_cast<int64_t>(graph.numberOfEdges()));
  make_parent_dirs(path);
  const int fd = ::open(path.c_str(), O_WRONLY | O_CREAT | O_TRUNC | O_CLOEXEC, 0644);
  if (fd < 0) {
    return false;
  }
  off_t at = 0;
  bool ok = true;
  auto put = [&](const void *p, size_t bytes) {
    ok = ok && write_all(fd, static_cast<const char *>(p), bytes, at);
    at += static_cast<off_t>(bytes);
  };
  put(head.data(), head.size());
  put(graph.rowPtr.data(), graph.rowPtr.size() * sizeof(int));
  put(graph.colInd.data(), graph.colInd.size() * sizeof(int));
  put(graph.weights.data(), graph.weights.size() * sizeof(int));
  ok = ::close(fd) == 0 && ok;
  if (!ok) {
    std::error_code ec;
    fs::remove(path, ec);
  }
  return ok;
}

bool loadCsrGraphBinary(const std::string &path, CsrGraph &graph,
                        const std::string &sourcePrefix) {
  graph = CsrGraph();
  SourceStamp want;
  MappedFile file;
  if (!stamp_of(sourcePrefix, want) || !file.open(path)) {
    return false;
  }
  ByteCursor in{file.begin(), file.end()};
  char magic[8];
  int64_t prefixLength = -1;
  if (!in.take(magic, 8) || std::memcmp(magic, kMagic, 8) != 0 ||
      !in.take(&prefixLength, 8) || prefixLength < 0 || prefixLength > (1 << 16) ||
      static_cast<size_t>(prefixLength) > in.left()) {
    return false;
  }
  SourceStamp got;
  got.prefix.assign(in.p, static_cast<size_t>(prefixLength));
  in.p += prefixLength;
  int32_t header[2] = {0, 0};
  int64_t m = -1;
  if (!in.take(got.size, sizeof(got.size)) || !in.take(got.mtime, sizeof(got.mtime)) ||
      !same_source(got, want) || !in.take(header, sizeof(header)) || !in.take(&m, 8)) {
    return false;
  }
  const int n = header[0], K = header[1];
  if (n <= 0 || K <= 0 || K > 32 || m < 0 || m > INT_MAX) {
    return false;
  }
  // The arrays must be exactly what remains: a damaged header cannot ask
  // for more memory than the file holds.
  const uint64_t words = static_cast<uint64_t>(n) + 1 + static_cast<uint64_t>(m) * (1 + K);
  if (in.left() != words * sizeof(int)) {
    return false;
  }
  graph.numberOfNodes = n;
  graph.numberOfObjectives = K;
  graph.rowPtr.resize(static_cast<size_t>(n) + 1);
  graph.colInd.resize(static_cast<size_t>(m));
  graph.weights.resize(static_cast<size_t>(m) * K);
  in.take(graph.rowPtr.data(), graph.rowPtr.size() * sizeof(int));
  in.take(graph.colInd.data(), graph.colInd.size() * sizeof(int));
  in.take(graph.weights.data(), graph.weights.size() * sizeof(int));
  if (!arrays_valid(graph)) {
    graph = CsrGraph();
    return false;
  }
  return true;
}

bool loadCsrGraph(const std::string &prefix, CsrGraph &graph, const std::string &cachePath) {
  const bool useCache = !cachePath.empty();
  if (useCache && loadCsrGraphBinary(cachePath, graph, prefix)) {
    return true;
  }
  std::error_code ec;
  if (useCache && fs::exists(cachePath, ec)) {
    std::cout << "Note: binary cache " << cachePath << " does not match the text graph "
              << prefix << " (another graph, changed files or an old format); rebuilding it.\n";
  }
  if (!readCsrGraph(prefix, graph)) {
    return false;
  }
  if (useCache && !saveCsrGraphBinary(cachePath, graph, prefix)) {
    std::cout << "Warning: could not write binary cache " << cachePath << "\n";
  }
  return true;
}

bool readChangeBatch(const std::string &insertPath, const std::string &deletePath, int K, int n,
                     ChangeBatch &batch) {
  batch = ChangeBatch();
  batch.numberOfObjectives = K;
  const auto vertex = [n](long long x) { return x >= 0 && x < n; };
  {
    MappedFile file;
    if (!file.open(insertPath)) {
      std::cout << "Error: Could not open insert file: " << insertPath << "\n";
      return false;
    }
    const char *origin = file.begin();
    // "u v w1 .. wK" per line (further fields ignored); blank lines skipped.
    std::vector<long long> fields;
    const bool ok = each_line(file.begin(), file.end(), [&](const char *b, const char *e) {
      LineTokens tokens(b, e);
      fields.clear();
      long long value = 0;
      for (int got; (got = tokens.next(value)) != 0;) {
        if (got < 0) return line_error("Invalid insert line", insertPath, origin, tokens.bad());
        fields.push_back(value);
      }
      if (fields.empty()) return true;
      if (static_cast<int>(fields.size()) < 2 + K) {
        return line_error("Invalid insert line", insertPath, origin, e);
      }
      if (!vertex(fields[0]) || !vertex(fields[1])) {
        return line_error("Inserted edge endpoint out of range", insertPath, origin, e);
      }
      for (int k = 0; k < K; ++k) {
        const long long w = fields[2 + k];
        if (!valid_weight(w)) {
          std::cout << "Error: Inserted edge weight " << w << " out of range [1, " << INT_MAX
                    << "] at " << where(insertPath, origin, e) << ".\n";
          return false;
        }
      }
      batch.insertFrom.push_back(static_cast<int>(fields[0]));
      batch.insertTo.push_back(static_cast<int>(fields[1]));
      for (int k = 0; k < K; ++k) batch.insertWeights.push_back(static_cast<int>(fields[2 + k]));
      return true;
    });
    if (!ok) return false;
  }
  MappedFile file;
  if (!file.open(deletePath)) {
    std::cout << "Error: Could not open delete file: " << deletePath << "\n";
    return false;
  }
  const char *origin = file.begin();
  // "u v" per line (further fields ignored); shorter lines are skipped.
  return each_line(file.begin(), file.end(), [&](const char *b, const char *e) {
    LineTokens tokens(b, e);
    long long pair[2] = {0, 0}, value = 0;
    int count = 0;
    for (int got; (got = tokens.next(value)) != 0; ++count) {
      if (got < 0) return line_error("Invalid delete line", deletePath, origin, tokens.bad());
      if (count < 2) pair[count] = value;
    }
    if (count < 2) return true;
    if (!vertex(pair[0]) || !vertex(pair[1])) {
      return line_error("Deleted edge endpoint out of range", deletePath, origin, e);
    }
    batch.deleteFrom.push_back(static_cast<int>(pair[0]));
    batch.deleteTo.push_back(static_cast<int>(pair[1]));
    return true;
  });
}

bool writeChangeBatch(const ChangeBatch &batch, const std::string &insertPath,
                      const std::string &deletePath) {
  const int K = batch.numberOfObjectives;
  const bool ins = write_lines(
      insertPath, static_cast<size_t>(batch.numberOfInserts()), 24 + 12 * static_cast<size_t>(K),
      [&](size_t i, char *o) {
        o = put_int(o, batch.insertFrom[i]);
        *o++ = ' ';
        o = put_int(o, batch.insertTo[i]);
        for (int k = 0; k < K; ++k) {
          *o++ = ' ';
          o = put_int(o, batch.insertWeights[i * static_cast<size_t>(K) + k]);
        }
        *o++ = '\n';
        return o;
      });
  const bool del = write_lines(deletePath, static_cast<size_t>(batch.numberOfDeletes()), 24,
                               [&](size_t i, char *o) {
                                 o = put_int(o, batch.deleteFrom[i]);
                                 *o++ = ' ';
                                 o = put_int(o, batch.deleteTo[i]);
                                 *o++ = '\n';
                                 return o;
                               });
  return ins && del;
}

// The batch's effect on one (u, v) pair, for D deletions and I insertions
// of it: the first D occurrences of v in row u die; with I > 0 the next one
// takes the last insertion's weights, or -- none left -- one entry carrying
// them is appended.  Pairs never interact, and a row's appended entries
// keep the order of their pairs' first insertions.  This is what replaying
// deletions first and then insertions in file order (each overwriting the
// first alive (u, v) or appending) produces (src/csrGraph.cu:743-859).
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
  const auto outside = [n](int x) { return x < 0 || x >= n; };
  for (int i = 0; i < nIns; ++i) {
    if (outside(batch.insertFrom[i]) || outside(batch.insertTo[i])) {
      std::cout << "Error: inserted edge " << i << " has an endpoint outside [0, " << n << ").\n";
      return false;
    }
  }
  for (int i = 0; i < nDel; ++i) {
    if (outside(batch.deleteFrom[i]) || outside(batch.deleteTo[i])) {
      std::cout << "Error: deleted edge " << i << " has an endpoint outside [0, " << n << ").\n";
      return false;
    }
  }

  // One record per change, sorted by (u, v), deletions before insertions,
  // file order within each kind.
  struct Change {
    uint64_t key;   // u << 32 | v
    int32_t index;  // deletion index, or nDel + insertion index
  };
  std::vector<Change> changes;
  changes.reserve(static_cast<size_t>(nIns) + nDel);
  const auto key = [](int u, int v) {
    return static_cast<uint64_t>(static_cast<uint32_t>(u)) << 32 | static_cast<uint32_t>(v);
  };
  for (int i = 0; i < nDel; ++i) changes.push_back({key(batch.deleteFrom[i], batch.deleteTo[i]), i});
  for (int i = 0; i < nIns; ++i) {
    changes.push_back({key(batch.insertFrom[i], batch.insertTo[i]), nDel + i});
  }
  std::sort(changes.begin(), changes.end(), [](const Change &a, const Change &b) {
    return a.key != b.key ? a.key < b.key : a.index < b.index;
  });

  struct Pair {
    int v;
    int deletions;      // D
    int firstInsert;    // first insertion index, -1: none
    int lastInsert;     // last insertion index (its weights win)
    int seen;           // occurrences of v met in the row so far
    bool placed;        // an original occurrence took the insertion
    size_t changes;     // its records: changes[changes, changes + count)
    size_t count;
  };
  struct Row {
    int u;
    size_t first, last;  // its pairs [first, last), sorted by v
  };
  std::vector<Pair> pairs;
  std::vector<Row> rows;
  for (size_t i = 0; i < changes.size();) {
    const uint64_t k = changes[i].key;
    Pair p{static_cast<int>(k & 0xffffffffu), 0, -1, -1, 0, false, i, 0};
    for (; i < changes.size() && changes[i].key == k; ++i, ++p.count) {
      const int idx = changes[i].index;
      if (idx < nDel) {
        ++p.deletions;
      } else {
        if (p.firstInsert < 0) p.firstInsert = idx - nDel;
        p.lastInsert = idx - nDel;
      }
    }
    const int u = static_cast<int>(k >> 32);
    if (rows.empty() || rows.back().u != u) rows.push_back({u, pairs.size(), pairs.size()});
    pairs.push_back(p);
    rows.back().last = pairs.size();
  }

  // New contents of each touched row, as (column, weight source) entries:
  // source >= 0 is an original edge, < 0 is ~insertion.
  struct Entry {
    int v;
    long long source;
  };
  std::vector<Entry> entries;
  std::vector<size_t> rowBegin;  // per touched row: its first entry
  std::vector<int> appendOrder;
  for (const Row &r : rows) {
    rowBegin.push_back(entries.size());
    const auto lo = pairs.begin() + static_cast<std::ptrdiff_t>(r.first);
    const auto hi = pairs.begin() + static_cast<std::ptrdiff_t>(r.last);
    for (int e = original.rowPtr[r.u]; e < original.rowPtr[r.u + 1]; ++e) {
      const int v = original.colInd[e];
      const auto it = std::lower_bound(lo, hi, v, [](const Pair &p, int x) { return p.v < x; });
      if (it == hi || it->v != v) {
        entries.push_back({v, e});
        continue;
      }
      const int occurrence = it->seen++;
      if (occurrence < it->deletions) {
        continue;  // deleted
      }
      if (occurrence == it->deletions && it->lastInsert >= 0) {
        it->placed = true;
        entries.push_back({v, ~static_cast<long long>(it->lastInsert)});
      } else {
        entries.push_back({v, e});
      }
    }
    appendOrder.clear();
    for (size_t p = r.first; p < r.last; ++p) {
      if (pairs[p].lastInsert >= 0 && !pairs[p].placed) appendOrder.push_back(static_cast<int>(p));
    }
    std::sort(appendOrder.begin(), appendOrder.end(),
              [&](int a, int b) { return pairs[a].firstInsert < pairs[b].firstInsert; });
    for (int p : appendOrder) {
      entries.push_back({pairs[p].v, ~static_cast<long long>(pairs[p].lastInsert)});
    }
  }
  rowBegin.push_back(entries.size());

  updated = CsrGraph();
  updated.numberOfNodes = n;
  updated.numberOfObjectives = K;
  updated.rowPtr.assign(static_cast<size_t>(n) + 1, 0);
  {
    size_t t = 0;  // next touched row
    for (int u = 0; u < n; ++u) {
      int degree = original.rowPtr[u + 1] - original.rowPtr[u];
      if (t < rows.size() && rows[t].u == u) {
        degree = static_cast<int>(rowBegin[t + 1] - rowBegin[t]);
        ++t;
      }
      updated.rowPtr[u + 1] = updated.rowPtr[u] + degree;
    }
  }
  const size_t m = static_cast<size_t>(updated.rowPtr[n]);
  updated.colInd.resize(m);
  updated.weights.resize(m * K);
  const size_t kk = static_cast<size_t>(K);
  size_t t = 0;
  for (int u = 0; u < n; ++u) {
    const size_t out = static_cast<size_t>(updated.rowPtr[u]);
    if (t < rows.size() && rows[t].u == u) {
      for (size_t x = rowBegin[t]; x < rowBegin[t + 1]; ++x) {
        const Entry &en = entries[x];
        const size_t at = out + (x - rowBegin[t]);
        updated.colInd[at] = en.v;
        const int *w = en.source >= 0 ? &original.weights[static_cast<size_t>(en.source) * kk]
                                      : &batch.insertWeights[static_cast<size_t>(~en.source) * kk];
        std::copy(w, w + K, updated.weights.begin() + static_cast<std::ptrdiff_t>(at * kk));
      }
      ++t;
      continue;
    }
    const int b = original.rowPtr[u], e = original.rowPtr[u + 1];
    std::copy(original.colInd.begin() + b, original.colInd.begin() + e,
              updated.colInd.begin() + static_cast<std::ptrdiff_t>(out));
    std::copy(original.weights.begin() + static_cast<std::ptrdiff_t>(b * kk),
              original.weights.begin() + static_cast<std::ptrdiff_t>(e * kk),
              updated.weights.begin() + static_cast<std::ptrdiff_t>(out * kk));
  }

  // Weight increases, per pair: the final first occurrence of (u, v) carries
  // the last insertion's weights (the surviving occurrence it overwrote, or
  // its appended entry); it is compared with the original first occurrence.
  batch.weightIncreaseMask.assign(static_cast<size_t>(nIns), 0u);
  for (const Row &r : rows) {
    for (size_t p = r.first; p < r.last; ++p) {
      const Pair &pr = pairs[p];
      if (pr.lastInsert < 0 || pr.seen == 0) {
        continue;  // no insertion, or no original occurrence
      }
      int before = original.rowPtr[r.u];
      while (original.colInd[before] != pr.v) ++before;
      const int *w0 = &original.weights[static_cast<size_t>(before) * kk];
      const int *w1 = &batch.insertWeights[static_cast<size_t>(pr.lastInsert) * kk];
      unsigned mask = 0;
      for (int k = 0; k < K; ++k) {
        mask |= (w1[k] > w0[k] ? 1u : 0u) << k;
      }
      // Every insertion of the pair gets the mask.
      for (size_t c = pr.changes; c < pr.changes + pr.count; ++c) {
        const int idx = changes[c].index;
        if (idx >= nDel) batch.weightIncreaseMask[static_cast<size_t>(idx - nDel)] = mask;
      }
    }
  }
  return true;
}

void transposeCsrGraph(const CsrGraph &g, CsrGraph &rev) {
  const int n = g.numberOfNodes, K = g.numberOfObjectives;
  const size_t m = static_cast<size_t>(g.numberOfEdges());
  rev = CsrGraph();
  rev.numberOfNodes = n;
  rev.numberOfObjectives = K;
  rev.rowPtr.assign(static_cast<size_t>(n) + 1, 0);
  for (size_t e = 0; e < m; ++e) ++rev.rowPtr[static_cast<size_t>(g.colInd[e]) + 1];
  for (int v = 0; v < n; ++v) rev.rowPtr[v + 1] += rev.rowPtr[v];
  rev.colInd.resize(m);
  rev.weights.resize(m * K);
  std::vector<int> next(rev.rowPtr.begin(), rev.rowPtr.end() - 1);
  for (int u = 0; u < n; ++u) {
    for (int e = g.rowPtr[u]; e < g.rowPtr[u + 1]; ++e) {
      const size_t at = static_cast<size_t>(next[g.colInd[e]]++);
      rev.colInd[at] = u;
      std::copy(g.weights.begin() + static_cast<std::ptrdiff_t>(static_cast<size_t>(e) * K),
                g.weights.begin() + static_cast<std::ptrdiff_t>(static_cast<size_t>(e + 1) * K),
                rev.weights.begin() + static_cast<std::ptrdiff_t>(at * K));
    }
  }
}

bool readDistances(const std::string &path, int n, std::vector<long long> &dist) {
  return read_pairs(path, "Could not open distances file: ", "Invalid line in distances file",
                    true, n, -1LL, dist,
                    [&](long long v, long long d, const char *origin, const char *at) {
                      if (v < 0 || v >= n) {
                        return line_error("Vertex ID out of range in distances file", path,
                                          origin, at);
                      }
                      if (d < 0) return line_error("Negative distance", path, origin, at);
                      dist[static_cast<size_t>(v)] = d;
                      return true;
                    });
}

bool readParents(const std::string &path, int n, std::vector<int> &parent) {
  return read_pairs(path, "Could not open SSSP tree file: ", "Invalid line in SSSP tree file",
                    false, n, -2, parent,
                    [&](long long v, long long p, const char *origin, const char *at) {
                      if (v < 0 || v >= n || p < -1 || p >= n) {
                        return line_error("Vertex ID out of range in SSSP tree file", path,
                                          origin, at);
                      }
                      parent[static_cast<size_t>(v)] = static_cast<int>(p);
                      return true;
                    });
}

bool writeDistances(const std::string &path, const std::vector<long long> &dist) {
  const bool ok = write_lines(path, dist.size(), 40, [&](size_t v, char *o) {
    o = put_int(o, static_cast<long long>(v));
    *o++ = ' ';
    if (dist[v] >= DISTANCE_INF / 2) {
      std::memcpy(o, "INF", 3);
      o += 3;
    } else {
      o = put_int(o, dist[v]);
    }
    *o++ = '\n';
    return o;
  });
  if (!ok) std::cout << "Error: Could not write distances file: " << path << "\n";
  return ok;
}

bool writeParents(const std::string &path, const std::vector<int> &parent) {
  const bool ok = write_lines(path, parent.size(), 24, [&](size_t v, char *o) {
    o = put_int(o, static_cast<long long>(v));
    *o++ = ' ';
    o = put_int(o, parent[v]);
    *o++ = '\n';
    return o;
  });
  if (!ok) std::cout << "Error: Could not write SSSP tree file: " << path << "\n";
  return ok;
}
