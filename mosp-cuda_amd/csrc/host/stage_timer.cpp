// stage_timer.cpp -- the instrumentation registry behind stageTimer.h
// (semantics of the reference's src/stageTimer.cu:21-134: stages in record
// order, counters summed per name and listed by name).
//
// One mutex guards the registry, so stages and counters may be recorded
// from concurrent jobs.  Profiler ranges go through the roctx library of
// rocprofiler-sdk, opened with dlopen the first time instrumentation is
// enabled: a process that never asks for timings does not load it, and a
// machine without it simply gets no ranges.
#include "stageTimer.h"

#include <dlfcn.h>

#include <cstdio>
#include <map>
#include <mutex>
#include <ostream>
#include <utility>
#include <vector>

#include "mosp_hip.h"

namespace {

struct Registry {
  std::mutex mu;
  bool enabled = false;
  std::string prefix;
  std::vector<std::pair<std::string, double>> stages;
  std::map<std::string, double> counters;
};

Registry &reg() {
  static Registry *r = new Registry();  // never destroyed: usable during exit
  return *r;
}

// roctxRangePushA / roctxRangePop, when the library can be opened.
struct Ranges {
  int (*push)(const char *) = nullptr;
  int (*pop)() = nullptr;
};

Ranges &ranges() {
  static Ranges r = [] {
    Ranges x;
    void *lib = dlopen("librocprofiler-sdk-roctx.so.1", RTLD_NOW | RTLD_LOCAL);
    if (lib == nullptr) lib = dlopen("librocprofiler-sdk-roctx.so", RTLD_NOW | RTLD_LOCAL);
    if (lib != nullptr) {
      x.push = reinterpret_cast<int (*)(const char *)>(dlsym(lib, "roctxRangePushA"));
      x.pop = reinterpret_cast<int (*)()>(dlsym(lib, "roctxRangePop"));
      if (x.push == nullptr || x.pop == nullptr) x = Ranges{};
    }
    return x;
  }();
  return r;
}

}  // namespace

void setInstrumentation(bool enabled) {
  {
    std::lock_guard<std::mutex> lock(reg().mu);
    reg().enabled = enabled;
  }
  if (enabled) (void)ranges();
}

bool instrumentationEnabled() {
  std::lock_guard<std::mutex> lock(reg().mu);
  return reg().enabled;
}

void setStagePrefix(const std::string &prefix) {
  std::lock_guard<std::mutex> lock(reg().mu);
  reg().prefix = prefix;
}

const std::string &stagePrefix() { return reg().prefix; }

void recordStage(const std::string &name, double milliseconds) {
  Registry &r = reg();
  std::lock_guard<std::mutex> lock(r.mu);
  if (r.enabled) r.stages.emplace_back(r.prefix + name, milliseconds);
}

void recordCounter(const std::string &name, double value) {
  Registry &r = reg();
  std::lock_guard<std::mutex> lock(r.mu);
  if (r.enabled) r.counters[r.prefix + name] += value;
}

double totalStageTime(const std::string &suffix) {
  Registry &r = reg();
  std::lock_guard<std::mutex> lock(r.mu);
  double total = 0;
  for (const auto &s : r.stages) {
    const std::string &name = s.first;
    if (name.size() >= suffix.size() &&
        name.compare(name.size() - suffix.size(), std::string::npos, suffix) == 0) {
      total += s.second;
    }
  }
  return total;
}

void clearInstrumentation() {
  Registry &r = reg();
  std::lock_guard<std::mutex> lock(r.mu);
  r.stages.clear();
  r.counters.clear();
}

void printInstrumentation(std::ostream &out) {
  Registry &r = reg();
  std::lock_guard<std::mutex> lock(r.mu);
  char line[320];
  for (const auto &s : r.stages) {
    std::snprintf(line, sizeof(line), "STAGE   %-52s %12.3f ms\n", s.first.c_str(), s.second);
    out << line;
  }
  for (const auto &c : r.counters) {
    std::snprintf(line, sizeof(line), "COUNTER %-52s %15.0f\n", c.first.c_str(), c.second);
    out << line;
  }
}

bool writeInstrumentationCsv(const std::string &path) {
  Registry &r = reg();
  std::string text = "kind,name,value\n";
  {
    std::lock_guard<std::mutex> lock(r.mu);
    char value[64];
    for (const auto &s : r.stages) {
      std::snprintf(value, sizeof(value), "%.6f", s.second);
      text += "stage," + s.first + "," + value + "\n";
    }
    for (const auto &c : r.counters) {
      std::snprintf(value, sizeof(value), "%.0f", c.second);
      text += "counter," + c.first + "," + value + "\n";
    }
  }
  FILE *f = std::fopen(path.c_str(), "w");
  if (f == nullptr) return false;
  const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
  return std::fclose(f) == 0 && ok;
}

void instrumentationSync() {
  if (instrumentationEnabled()) (void)mosp_device_synchronize();
}

ScopedStage::ScopedStage(const std::string &name, bool syncDevice)
    : name_(name), active_(instrumentationEnabled()), sync_(syncDevice) {
  if (active_) {
    if (ranges().push != nullptr) ranges().push((stagePrefix() + name_).c_str());
    start_ = std::chrono::steady_clock::now();
  }
}

void ScopedStage::stop() {
  if (!active_) return;
  active_ = false;
  if (sync_) (void)mosp_device_synchronize();
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - start_).count();
  if (ranges().pop != nullptr) ranges().pop();
  recordStage(name_, ms);
}
