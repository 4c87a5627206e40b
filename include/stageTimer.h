// stageTimer.h -- optional stage timings, counters and profiler ranges.
//
// Same functions and semantics as the reference's headers/stageTimer.cuh:
// while instrumentation is enabled, recordStage() appends a (prefix + name,
// ms) row and recordCounter() adds to a named counter; the CSV written by
// writeInstrumentationCsv() has the header "kind,name,value", then one
// "stage" row per recorded stage and one "counter" row per counter.
// ScopedStage times a scope (optionally after a device synchronisation) and,
// when the ROCm roctx library is present, also opens a roctx range named
// after the stage, which `rocprofv3 --marker-trace` records.  Unlike the
// reference's registry this one is safe to use from several threads.
#ifndef MOSP_STAGE_TIMER_H
#define MOSP_STAGE_TIMER_H

#include <chrono>
#include <iosfwd>
#include <string>

void setInstrumentation(bool enabled);
bool instrumentationEnabled();

// Prepended to every stage and counter name recorded afterwards.
void setStagePrefix(const std::string &prefix);
const std::string &stagePrefix();

void recordStage(const std::string &name, double milliseconds);
void recordCounter(const std::string &name, double value);

// Sum of the recorded stages whose name ends with `suffix`.
double totalStageTime(const std::string &suffix);

void clearInstrumentation();
// "STAGE   <name> <ms> ms" and "COUNTER <name> <value>" lines.
void printInstrumentation(std::ostream &out);
bool writeInstrumentationCsv(const std::string &path);

// Waits for the device when instrumentation is enabled (so a stage measures
// the GPU work it started).
void instrumentationSync();

class ScopedStage {
 public:
  explicit ScopedStage(const std::string &name, bool syncDevice = false);
  ~ScopedStage() { stop(); }
  ScopedStage(const ScopedStage &) = delete;
  ScopedStage &operator=(const ScopedStage &) = delete;
  // Records the stage now (at most once).
  void stop();

 private:
  std::string name_;
  bool active_;
  bool sync_;
  std::chrono::steady_clock::time_point start_;
};

#endif  // MOSP_STAGE_TIMER_H
