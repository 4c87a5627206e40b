// serve.cpp -- served updates (C ABI mosp_server_*): graphs resident in HBM,
// change batches arriving as requests, continuous batching over slots.
//
// * A graph (`Graph`) keeps, in HBM: its current and spare device graphs,
//   its K current trees and the buffers for the next ones, the combined
//   outputs, and the statistics the reference takes from each update's
//   original graph (per-objective max weight for the packing bound, m and the
//   weight sum for the near-far width; src/mospUpdate.cu:104-122, 177-181),
//   recomputed on the device after every apply.
// * Requests on one graph keep their order: one is on the GPU at a time, so
//   a request sees exactly the state its predecessor left and every result
//   is independent of the mix.
// * A slot is a worker thread with a reusable batched workspace
//   (mosp_batch_create_capacity / mosp_batch_reset: nothing is allocated per
//   request), its stream and a device arena for change lists.  A free slot
//   admits the oldest waiting requests of idle graphs -- as many as fit its
//   share of the HBM budget and max_coalesce -- and runs them in ONE batched
//   launch: per request the batch is applied to the graph in HBM
//   (mosp_device_graph_apply into the spare graph), the changed lists come
//   from the weight-increase mask (src/mospUpdate.cu:124-136), then the
//   launch runs every admitted update (K objectives + combined graph + its
//   search) and the graphs' current/spare buffers swap.
// * CUs: policy 0 gives each slot a fixed 1/slots of the CUs; policy 1 lets a
//   launch take the idle CUs, split evenly over the launches that could
//   start at that moment (at least min_workgroups, else the requests wait).
//   The workgroups of concurrent launches never exceed one per CU, so every
//   persistent workgroup is co-resident (plain launches, use_graph 2).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "engine_internal.hpp"
#include "mosp_hip.h"

using mosp::detail::clear_error;
using mosp::detail::fail;

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id = 0;
  int gid = 0;
  std::vector<int32_t> insFrom, insTo, insW, delFrom, delTo;
  double arrival = 0, admitted = -1, done = -1;
  int slot = -1, coalesced = 0;
  int status = 0;  // 0 waiting or running, 1 done, -1 failed
  std::vector<int64_t> dist, combDist;  // snapshots (keep_results)
  std::vector<int32_t> parents, combParent;
  int64_t changes() const { return static_cast<int64_t>(insFrom.size() + delFrom.size()); }
};

struct Graph {
  int n = 0, K = 0, KG = 0, source = 0;
  std::vector<int32_t> pref;
  mosp_device_graph *g[2] = {nullptr, nullptr};
  int cur = 0;
  int64_t *dist[2] = {nullptr, nullptr};
  int32_t *par[2] = {nullptr, nullptr};
  int64_t *combDist = nullptr;
  int32_t *combPar = nullptr;
  int64_t m = 0;
  std::vector<int64_t> maxw, wsum;  // KG entries, of the current graph
  std::deque<Request *> queue;
  bool busy = false;
  int64_t footprint = 0;  // slot scratch one update needs
  int64_t resident = 0;   // HBM held between requests (estimate)
};

struct Slot {
  int index = 0;
  mosp_batch *batch = nullptr;
  hipStream_t stream = nullptr;
  int32_t *arena = nullptr;
  size_t arenaInts = 0;
  int workgroups = 0;
  bool busy = false;
  double kernelMs = 0, applyS = 0, launchS = 0, postS = 0;
  int64_t launches = 0;
  std::thread thread;
};

double seconds(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

}  // namespace

struct mosp_server {
  mosp_server_config cfg{};
  int cus = 0;
  std::vector<std::unique_ptr<Graph>> graphs;
  std::vector<std::unique_ptr<Slot>> slots;
  std::unordered_map<int64_t, std::unique_ptr<Request>> requests;
  std::mutex mu;
  std::condition_variable cv;
  int64_t nextId = 0;
  int64_t pending = 0;
  bool started = false, stop = false;
  std::string failure;  // first worker error
  int freeCus = 0;
  int device = 0;
  int64_t slabBytes = 0;
  int64_t residentBytes = 0;
  Clock::time_point t0 = Clock::now();

  double now() const { return seconds(t0, Clock::now()); }
};

namespace {

void release_graph(Graph &G) {
  for (int i = 0; i < 2; ++i) {
    if (G.g[i] != nullptr) mosp_device_graph_free(G.g[i]);
    (void)hipFree(G.dist[i]);
    (void)hipFree(G.par[i]);
  }
  (void)hipFree(G.combDist);
  (void)hipFree(G.combPar);
}

// Oldest waiting requests of idle graphs that fit one slot, and their
// launch's workgroups (0: the slot's fixed share).  Caller holds the lock.
std::vector<Request *> admit(mosp_server *s, int &workgroups) {
  workgroups = 0;
  std::vector<Request *> heads;
  for (auto &G : s->graphs) {
    if (!G->queue.empty() && !G->busy) heads.push_back(G->queue.front());
  }
  if (heads.empty()) return {};
  if (s->cfg.cu_policy == 1) {
    int idle = 0;
    for (auto &slot : s->slots) idle += slot->busy ? 0 : 1;
    const int coalesce = std::max(1, s->cfg.max_coalesce);
    const int could = static_cast<int>((heads.size() + coalesce - 1) / coalesce);
    const int launches = std::max(1, std::min(idle, could));
    workgroups = s->freeCus / launches;
    if (workgroups < std::max(1, s->cfg.min_workgroups)) return {};
  }
  std::sort(heads.begin(), heads.end(), [](const Request *a, const Request *b) {
    return a->arrival != b->arrival ? a->arrival < b->arrival : a->id < b->id;
  });
  std::vector<Request *> chosen;
  int64_t used = 0;
  for (Request *r : heads) {
    const int64_t need = s->graphs[r->gid]->footprint;
    if (used + need > s->slabBytes || static_cast<int>(chosen.size()) >= s->cfg.max_coalesce) {
      continue;
    }
    chosen.push_back(r);
    used += need;
  }
  for (Request *r : chosen) {
    Graph &G = *s->graphs[r->gid];
    G.queue.pop_front();
    G.busy = true;
  }
  s->freeCus -= workgroups;
  return chosen;
}

#define SERVE_HIP(call)                                                                     \
  do {                                                                                      \
    const hipError_t err_ = (call);                                                         \
    if (err_ != hipSuccess) {                                                               \
      return fail(std::string("HIP error: ") + hipGetErrorString(err_) + " at " + __FILE__ + \
                  ":" + std::to_string(__LINE__));                                          \
    }                                                                                       \
  } while (0)

#define SERVE_TRY(call, what)                                                   \
  do {                                                                          \
    if (!(call)) return fail(std::string(what) + ": " + mosp_last_error());     \
  } while (0)

// Time and device time of one slot launch; the worker adds them to the
// slot's totals under the server's lock (the statistics calls read them).
struct LaunchTotals {
  double kernelMs = 0, applyS = 0, launchS = 0, postS = 0;
};

// One launch of the admitted requests on `slot` (its own thread, no lock).
int run_slot(mosp_server *s, Slot &slot, const std::vector<Request *> &reqs, int workgroups,
             LaunchTotals &tot) {
  const auto t0 = Clock::now();
  if (workgroups > 0 && workgroups != slot.workgroups) {
    SERVE_TRY(mosp_batch_configure(slot.batch, workgroups, -1, 1.0, 1), "configure");
    slot.workgroups = workgroups;
  }
  const size_t count = reqs.size();
  std::vector<mosp_batch_instance> inst(count);
  std::vector<std::vector<mosp_changes>> changes(count);
  std::vector<std::vector<int64_t>> delta(count), maxw(count);
  std::vector<std::vector<int32_t>> staged(count);
  size_t offset = 0;
  for (size_t i = 0; i < count; ++i) {
    Request &r = *reqs[i];
    Graph &G = *s->graphs[r.gid];
    const int n = G.n, K = G.K, KG = G.KG;
    const int nIns = static_cast<int>(r.insFrom.size());
    const int nDel = static_cast<int>(r.delFrom.size());
    mosp_device_graph *cur = G.g[G.cur], *spare = G.g[1 - G.cur];
    std::vector<uint32_t> mask(std::max(nIns, 1));
    SERVE_TRY(mosp_device_graph_apply(cur, nIns, r.insFrom.data(), r.insTo.data(), r.insW.data(),
                                      nDel, r.delFrom.data(), r.delTo.data(), spare, mask.data(),
                                      slot.stream),
              "apply");
    // Bounds and widths from the graph the batch was applied to
    // (src/mospUpdate.cu:104-122, 177-181).
    maxw[i].resize(K);
    delta[i].resize(K);
    for (int k = 0; k < K; ++k) {
      int64_t w = std::max<int64_t>(G.maxw[k], 1);
      for (int j = 0; j < nIns; ++j) w = std::max<int64_t>(w, r.insW[static_cast<size_t>(j) * KG + k]);
      maxw[i][k] = w;
      const double base = static_cast<double>(
          mosp_default_delta(std::max<int64_t>(G.m, 1), n, G.wsum[k]));
      delta[i][k] = std::max<int64_t>(1, static_cast<int64_t>(base * s->cfg.delta_scale));
    }
    // Arena layout per request: heads, then per objective (from, to) of its
    // changed list: every deletion + the insertions that raised objective k
    // (src/mospUpdate.cu:124-136).
    std::vector<int32_t> &h = staged[i];
    h.assign(r.insTo.begin(), r.insTo.end());
    std::vector<size_t> listAt(K), listLen(K);
    for (int k = 0; k < K; ++k) {
      std::vector<int32_t> from(r.delFrom), to(r.delTo);
      for (int j = 0; j < nIns; ++j) {
        if ((mask[j] >> k) & 1u) {
          from.push_back(r.insFrom[j]);
          to.push_back(r.insTo[j]);
        }
      }
      listAt[k] = h.size();
      listLen[k] = from.size();
      h.insert(h.end(), from.begin(), from.end());
      h.insert(h.end(), to.begin(), to.end());
    }
    if (offset + h.size() > slot.arenaInts) {
      return fail("the admitted requests exceed the slot's change arena (max_batch)");
    }
    int32_t *base = slot.arena + offset;
    if (!h.empty()) {
      SERVE_HIP(hipMemcpyAsync(base, h.data(), h.size() * sizeof(int32_t), hipMemcpyHostToDevice,
                               slot.stream));
    }
    changes[i].resize(K);
    for (int k = 0; k < K; ++k) {
      const int32_t *from = base + listAt[k];
      changes[i][k] = mosp_changes{from, from + listLen[k], static_cast<int32_t>(listLen[k]), base,
                                   nIns};
    }
    offset += h.size();
    const int next = 1 - G.cur;
    inst[i] = mosp_batch_instance{spare, K, changes[i].data(), G.source, delta[i].data(),
                                  maxw[i].data(), G.pref.data(),
                                  static_cast<int32_t>(G.pref.size()), G.dist[G.cur],
                                  G.par[G.cur], G.dist[next], G.par[next], G.combDist,
                                  G.combPar};
  }
  const auto t1 = Clock::now();
  SERVE_TRY(mosp_batch_reset(slot.batch, inst.data(), static_cast<int32_t>(count)), "reset");
  float ms = 0;
  SERVE_TRY(mosp_batch_run(slot.batch, 2, &ms), "run");
  SERVE_TRY(mosp_batch_check(slot.batch), "check");
  tot.kernelMs = ms;
  const auto t2 = Clock::now();
  for (Request *r : reqs) {
    Graph &G = *s->graphs[r->gid];
    G.cur = 1 - G.cur;
    mosp_device_graph *now = G.g[G.cur];
    G.m = mosp_device_graph_edges(now);
    SERVE_TRY(mosp_device_graph_weight_stats(now, G.maxw.data(), G.wsum.data(), slot.stream),
              "weight statistics");
    r->coalesced = static_cast<int>(count);
    if (s->cfg.keep_results) {
      const size_t kn = static_cast<size_t>(G.K) * G.n;
      r->dist.resize(kn);
      r->parents.resize(kn);
      r->combDist.resize(G.n);
      r->combParent.resize(G.n);
      SERVE_HIP(hipMemcpy(r->dist.data(), G.dist[G.cur], kn * 8, hipMemcpyDeviceToHost));
      SERVE_HIP(hipMemcpy(r->parents.data(), G.par[G.cur], kn * 4, hipMemcpyDeviceToHost));
      SERVE_HIP(hipMemcpy(r->combDist.data(), G.combDist, G.n * 8, hipMemcpyDeviceToHost));
      SERVE_HIP(hipMemcpy(r->combParent.data(), G.combPar, G.n * 4, hipMemcpyDeviceToHost));
    }
  }
  const auto t3 = Clock::now();
  tot.applyS = seconds(t0, t1);
  tot.launchS = seconds(t1, t2);
  tot.postS = seconds(t2, t3);
  return 1;
}

void worker(mosp_server *s, Slot *slot) {
  (void)hipSetDevice(s->device);
  while (true) {
    std::vector<Request *> reqs;
    int workgroups = 0;
    {
      std::unique_lock<std::mutex> lock(s->mu);
      while (!s->stop) {
        reqs = admit(s, workgroups);
        if (!reqs.empty()) break;
        s->cv.wait(lock);
      }
      if (s->stop) return;
      slot->busy = true;
      const double t = s->now();
      for (Request *r : reqs) {
        r->admitted = t;
        r->slot = slot->index;
      }
    }
    LaunchTotals tot;
    const int ok = run_slot(s, *slot, reqs, workgroups, tot);
    std::lock_guard<std::mutex> lock(s->mu);
    if (ok) {
      slot->kernelMs += tot.kernelMs;
      slot->launches += 1;
      slot->applyS += tot.applyS;
      slot->launchS += tot.launchS;
      slot->postS += tot.postS;
    }
    const double t = s->now();
    for (Request *r : reqs) {
      r->done = t;
      r->status = ok ? 1 : -1;
      s->graphs[r->gid]->busy = false;
    }
    s->pending -= static_cast<int64_t>(reqs.size());
    s->freeCus += workgroups;
    slot->busy = false;
    if (!ok) {
      if (s->failure.empty()) {
        s->failure = "slot " + std::to_string(slot->index) + ": " + mosp_last_error();
      }
      s->stop = true;
    }
    s->cv.notify_all();
    if (!ok) return;
  }
}

void stop_workers(mosp_server *s) {
  {
    std::lock_guard<std::mutex> lock(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  for (auto &slot : s->slots) {
    if (slot->thread.joinable()) slot->thread.join();
  }
}

}  // namespace

extern "C" mosp_server *mosp_server_create(const mosp_server_config *config) {
  clear_error();
  if (config == nullptr || config->slots < 1 || config->max_coalesce < 1 ||
      config->max_batch < 0 || !(config->delta_scale > 0) ||
      (config->cu_policy != 0 && config->cu_policy != 1)) {
    fail("Error: invalid server configuration.");
    return nullptr;
  }
  auto *s = new mosp_server();
  s->cfg = *config;
  s->cus = std::max(1, mosp_device_cus());
  if (hipGetDevice(&s->device) != hipSuccess) s->device = 0;
  return s;
}

extern "C" int32_t mosp_server_add_graph(mosp_server *s, int32_t n, int32_t kg,
                                         const int32_t *row_ptr, const int32_t *col_ind,
                                         const int32_t *weights, int32_t k, const int64_t *dist,
                                         const int32_t *parents, const int32_t *pref,
                                         int32_t n_pref, int32_t source) {
  clear_error();
  if (s == nullptr || s->started) {
    return fail("Error: graphs are added before mosp_server_start."), -1;
  }
  if (n <= 0 || kg <= 0 || k <= 0 || k > kg || source < 0 || source >= n || dist == nullptr ||
      parents == nullptr || (n_pref != 0 && n_pref != k)) {
    return fail("Error: invalid served graph."), -1;
  }
  auto G = std::make_unique<Graph>();
  G->n = n;
  G->K = k;
  G->KG = kg;
  G->source = source;
  if (n_pref > 0) G->pref.assign(pref, pref + n_pref);
  auto fail_graph = [&](const std::string &what) {
    fail(what + ": " + mosp_last_error());
    release_graph(*G);
    return -1;
  };
  G->g[0] = mosp_device_graph_upload(n, kg, row_ptr, col_ind, weights);
  if (G->g[0] == nullptr) return fail_graph("upload");
  G->g[1] = mosp_device_graph_create();
  const size_t kn = static_cast<size_t>(k) * n;
  for (int i = 0; i < 2; ++i) {
    if (hipMalloc(&G->dist[i], kn * 8) != hipSuccess ||
        hipMalloc(&G->par[i], kn * 4) != hipSuccess) {
      return fail_graph("tree buffers");
    }
  }
  if (hipMalloc(&G->combDist, static_cast<size_t>(n) * 8) != hipSuccess ||
      hipMalloc(&G->combPar, static_cast<size_t>(n) * 4) != hipSuccess ||
      hipMemcpy(G->dist[0], dist, kn * 8, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(G->par[0], parents, kn * 4, hipMemcpyHostToDevice) != hipSuccess) {
    return fail_graph("trees");
  }
  G->m = mosp_device_graph_edges(G->g[0]);
  G->maxw.assign(kg, 0);
  G->wsum.assign(kg, 0);
  if (!mosp_device_graph_weight_stats(G->g[0], G->maxw.data(), G->wsum.data(), nullptr)) {
    return fail_graph("weight statistics");
  }
  G->footprint = mosp_batch_instance_bytes(n, k);
  if (G->footprint < 0) return fail_graph("scratch size");
  const int64_t edges = static_cast<int64_t>(G->m * 1.125) + 4096;
  const int64_t graph = 2 * (4 * (static_cast<int64_t>(n) + 1) + 4 * edges * (1 + kg));
  G->resident = 2 * graph + 2 * 12 * static_cast<int64_t>(kn) + 12 * static_cast<int64_t>(n);
  std::lock_guard<std::mutex> lock(s->mu);
  s->graphs.push_back(std::move(G));
  return static_cast<int32_t>(s->graphs.size() - 1);
}

extern "C" int mosp_server_start(mosp_server *s) {
  clear_error();
  if (s == nullptr || s->started || s->graphs.empty()) {
    return fail("Error: a server starts once, with at least one graph.");
  }
  const int slots = s->cfg.slots;
  int64_t foot = 0, resident = 0;
  int kmax = 0;
  for (auto &G : s->graphs) {
    foot = std::max(foot, G->footprint);
    resident += G->resident;
    kmax = std::max(kmax, G->K);
  }
  int64_t budget = s->cfg.hbm_budget;
  if (budget <= 0) budget = resident + static_cast<int64_t>(slots) * foot * s->cfg.max_coalesce;
  const int64_t perSlot = (budget - resident) / slots;
  if (perSlot < foot) {
    return fail("Error: HBM budget " + std::to_string(budget) + " leaves " +
                std::to_string(perSlot) + " B per slot; the largest update needs " +
                std::to_string(foot) + " B.");
  }
  s->cfg.hbm_budget = budget;
  s->slabBytes = perSlot;
  s->residentBytes = resident;
  if (s->cfg.min_workgroups <= 0) s->cfg.min_workgroups = std::max(1, s->cus / (2 * slots));
  const int share = std::max(1, s->cus / slots);
  const size_t maxBatch = static_cast<size_t>(std::max(s->cfg.max_batch, 1));
  const size_t arena = static_cast<size_t>(s->cfg.max_coalesce) *
                       (static_cast<size_t>(kmax) * 4 * maxBatch + maxBatch);
  for (int i = 0; i < slots; ++i) {
    auto slot = std::make_unique<Slot>();
    slot->index = i;
    slot->batch = mosp_batch_create_capacity(s->cfg.max_coalesce, kmax * s->cfg.max_coalesce,
                                             perSlot);
    if (slot->batch == nullptr) return fail(std::string("slot: ") + mosp_last_error());
    slot->workgroups = share;
    if (!mosp_batch_configure(slot->batch, share, -1, 1.0, 1)) {
      return fail(std::string("slot: ") + mosp_last_error());
    }
    slot->stream = static_cast<hipStream_t>(mosp_batch_stream(slot->batch));
    slot->arenaInts = arena;
    SERVE_HIP(hipMalloc(&slot->arena, arena * sizeof(int32_t)));
    s->slots.push_back(std::move(slot));
  }
  s->freeCus = s->cus;
  s->t0 = Clock::now();
  s->started = true;
  for (auto &slot : s->slots) slot->thread = std::thread(worker, s, slot.get());
  return 1;
}

extern "C" int64_t mosp_server_submit(mosp_server *s, int32_t gid, int32_t n_ins,
                                      const int32_t *ins_from, const int32_t *ins_to,
                                      const int32_t *ins_w, int32_t n_del,
                                      const int32_t *del_from, const int32_t *del_to) {
  clear_error();
  if (s == nullptr || !s->started) return fail("Error: the server is not started."), -1;
  if (gid < 0 || gid >= static_cast<int32_t>(s->graphs.size()) || n_ins < 0 || n_del < 0 ||
      n_ins + n_del > std::max(s->cfg.max_batch, 0)) {
    return fail("Error: invalid request (graph id, or more changes than max_batch)."), -1;
  }
  // A malformed request is refused here, alone: once queued it would be
  // coalesced with others, and a failure inside a slot stops the server.
  const int n = s->graphs[gid]->n, KG0 = s->graphs[gid]->KG;
  if ((n_ins > 0 && (ins_from == nullptr || ins_to == nullptr || ins_w == nullptr)) ||
      (n_del > 0 && (del_from == nullptr || del_to == nullptr))) {
    return fail("Error: invalid request (missing change arrays)."), -1;
  }
  const auto outside = [n](int32_t v) { return v < 0 || v >= n; };
  for (int32_t i = 0; i < n_ins; ++i) {
    if (outside(ins_from[i]) || outside(ins_to[i])) {
      return fail("Error: invalid request: insertion " + std::to_string(i) +
                  " has a vertex outside [0, " + std::to_string(n) + ")."), -1;
    }
    for (int k = 0; k < KG0; ++k) {
      if (ins_w[static_cast<size_t>(i) * KG0 + k] < 1) {
        return fail("Error: invalid request: insertion " + std::to_string(i) +
                    " has a weight below 1."), -1;
      }
    }
  }
  for (int32_t i = 0; i < n_del; ++i) {
    if (outside(del_from[i]) || outside(del_to[i])) {
      return fail("Error: invalid request: deletion " + std::to_string(i) +
                  " has a vertex outside [0, " + std::to_string(n) + ")."), -1;
    }
  }
  auto r = std::make_unique<Request>();
  r->gid = gid;
  const int KG = KG0;
  r->insFrom.assign(ins_from, ins_from + n_ins);
  r->insTo.assign(ins_to, ins_to + n_ins);
  r->insW.assign(ins_w, ins_w + static_cast<size_t>(n_ins) * KG);
  r->delFrom.assign(del_from, del_from + n_del);
  r->delTo.assign(del_to, del_to + n_del);
  std::lock_guard<std::mutex> lock(s->mu);
  if (!s->failure.empty()) return fail("Error: " + s->failure), -1;
  r->id = s->nextId++;
  r->arrival = s->now();
  Request *raw = r.get();
  s->requests.emplace(raw->id, std::move(r));
  s->graphs[gid]->queue.push_back(raw);
  s->pending += 1;
  s->cv.notify_all();
  return raw->id;
}

extern "C" int mosp_server_wait(mosp_server *s, int64_t request, double timeout_s) {
  clear_error();
  if (s == nullptr) return fail("Error: no server.");
  std::unique_lock<std::mutex> lock(s->mu);
  auto it = s->requests.find(request);
  if (it == s->requests.end()) return fail("Error: unknown request.");
  Request *r = it->second.get();
  const auto until = Clock::now() + std::chrono::duration<double>(timeout_s);
  while (r->status == 0 && s->failure.empty()) {
    if (s->cv.wait_until(lock, until) == std::cv_status::timeout && r->status == 0) {
      return fail("Error: request not done within the timeout.");
    }
  }
  if (r->status != 1) return fail("Error: " + (s->failure.empty() ? "request failed" : s->failure));
  return 1;
}

extern "C" int mosp_server_drain(mosp_server *s, double timeout_s) {
  clear_error();
  if (s == nullptr) return fail("Error: no server.");
  std::unique_lock<std::mutex> lock(s->mu);
  const auto until = Clock::now() + std::chrono::duration<double>(timeout_s);
  while (s->pending > 0 && s->failure.empty()) {
    if (s->cv.wait_until(lock, until) == std::cv_status::timeout && s->pending > 0) {
      return fail("Error: " + std::to_string(s->pending) + " requests still pending.");
    }
  }
  if (!s->failure.empty()) return fail("Error: " + s->failure);
  return 1;
}

extern "C" int mosp_server_request(mosp_server *s, int64_t request, mosp_request_info *info) {
  clear_error();
  if (s == nullptr || info == nullptr) return fail("Error: no server.");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->requests.find(request);
  if (it == s->requests.end()) return fail("Error: unknown request.");
  const Request &r = *it->second;
  *info = mosp_request_info{r.gid, r.slot, r.coalesced, r.status, r.changes(),
                            r.arrival, r.admitted, r.done};
  return 1;
}

extern "C" int mosp_server_request_trees(mosp_server *s, int64_t request, int64_t *dist,
                                         int32_t *parents, int64_t *comb_dist,
                                         int32_t *comb_parent) {
  clear_error();
  if (s == nullptr) return fail("Error: no server.");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->requests.find(request);
  if (it == s->requests.end()) return fail("Error: unknown request.");
  const Request &r = *it->second;
  if (r.status != 1 || r.dist.empty()) {
    return fail("Error: no snapshot (request not done, or keep_results off).");
  }
  std::copy(r.dist.begin(), r.dist.end(), dist);
  std::copy(r.parents.begin(), r.parents.end(), parents);
  std::copy(r.combDist.begin(), r.combDist.end(), comb_dist);
  std::copy(r.combParent.begin(), r.combParent.end(), comb_parent);
  return 1;
}

extern "C" int mosp_server_release(mosp_server *s, int64_t request) {
  clear_error();
  if (s == nullptr) return fail("Error: no server.");
  std::lock_guard<std::mutex> lock(s->mu);
  auto it = s->requests.find(request);
  if (it == s->requests.end() || it->second->status == 0) {
    return fail("Error: unknown or unfinished request.");
  }
  s->requests.erase(it);
  return 1;
}

extern "C" int mosp_server_stats(mosp_server *s, mosp_server_counters *stats, double *slot_kernel_ms,
                                 int32_t n_slots) {
  clear_error();
  if (s == nullptr || stats == nullptr) return fail("Error: no server.");
  std::lock_guard<std::mutex> lock(s->mu);
  mosp_server_counters st{};
  st.slots = static_cast<int32_t>(s->slots.size());
  st.cus = s->cus;
  st.hbm_budget = s->cfg.hbm_budget;
  st.resident_bytes = s->residentBytes;
  st.slot_bytes = s->slabBytes;
  for (size_t i = 0; i < s->slots.size(); ++i) {
    const Slot &slot = *s->slots[i];
    st.launches += slot.launches;
    st.kernel_ms += slot.kernelMs;
    st.apply_s += slot.applyS;
    st.launch_s += slot.launchS;
    st.post_s += slot.postS;
    if (slot_kernel_ms != nullptr && static_cast<int32_t>(i) < n_slots) {
      slot_kernel_ms[i] = slot.kernelMs;
    }
  }
  *stats = st;
  return 1;
}

extern "C" int mosp_server_reset_stats(mosp_server *s) {
  clear_error();
  if (s == nullptr) return fail("Error: no server.");
  std::lock_guard<std::mutex> lock(s->mu);
  for (auto &slot : s->slots) {
    slot->kernelMs = slot->applyS = slot->launchS = slot->postS = 0;
    slot->launches = 0;
  }
  return 1;
}

extern "C" double mosp_server_clock(const mosp_server *s) { return s == nullptr ? 0 : s->now(); }

extern "C" void mosp_server_destroy(mosp_server *s) {
  if (s == nullptr) return;
  stop_workers(s);
  for (auto &slot : s->slots) {
    if (slot->batch != nullptr) mosp_batch_destroy(slot->batch);
    (void)hipFree(slot->arena);
  }
  for (auto &G : s->graphs) release_graph(*G);
  delete s;
}
