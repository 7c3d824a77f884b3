// Asynchronous ingestion over the C ABI (include/vdbgpu.h, vdbgpu_async_*):
// the reference's addDataToAccumulate / accumulation threads / integration
// thread (VDBMapping.hpp:355-370, :1383-1430, :133), as an opt-in engine per
// map handle.
//
// Per source: a latest-only input slot (a new cloud replaces one its worker
// has not taken yet, VDBMapping.hpp:366-368) and a worker thread that takes
// the slot, runs accumulateUpdate for it and then sleeps until its period
// (1 / max_rate of addInputSource, VDBMapping.hpp:1364-1371) since the start
// of that round has passed. One integration thread runs integrateUpdate
// every accumulation_period. Scans accumulated within one period OR-merge
// into the source's update set; replaced scans are dropped, as in the
// reference. The per-scan exact path (vdbgpu_insert, vdbgpu_server_*) is
// unaffected; the engine only calls the public ABI. Engines are held by
// shared pointers: a call that found an engine keeps it alive while
// vdbgpu_async_stop runs, and a stopped engine takes no more clouds.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vdbgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Slot {
  std::vector<float> xyz;  // 4 floats per point
  double origin[3];
  bool full = false;
};

struct Worker {
  std::string id;
  double period_ms = 0;
  Slot slot;
  bool busy = false;  // taken, accumulate running
  std::condition_variable cv;
  std::thread th;
};

struct Engine {
  vdbgpu_map* map = nullptr;
  double period_ms = 0;
  std::mutex mu;
  std::condition_variable cv_idle;   // a worker finished an accumulate
  std::condition_variable cv_integ;  // stop
  std::map<std::string, std::unique_ptr<Worker>> workers;
  std::thread integ;
  bool stop = false;
  int32_t first_error = VDBGPU_OK;
  int64_t submitted = 0, replaced = 0, accumulated = 0, integrations = 0;
};

std::mutex g_mu;
std::map<vdbgpu_map*, std::shared_ptr<Engine>> g_engines;

void note_error(Engine* e, int32_t rc) {
  if (rc != VDBGPU_OK && e->first_error == VDBGPU_OK) e->first_error = rc;
}

void worker_loop(Engine* e, Worker* w) {
  for (;;) {
    Slot taken;
    Clock::time_point round_start;
    {
      std::unique_lock<std::mutex> lk(e->mu);
      w->cv.wait(lk, [&] { return e->stop || w->slot.full; });
      if (e->stop) return;
      round_start = Clock::now();
      taken = std::move(w->slot);
      w->slot = Slot{};
      w->busy = true;
    }
    const int32_t rc = vdbgpu_accumulate(e->map, w->id.c_str(), taken.xyz.data(),
                                         (int64_t)taken.xyz.size() / 4, 16, taken.origin, 0);
    {
      std::lock_guard<std::mutex> lk(e->mu);
      w->busy = false;
      ++e->accumulated;
      note_error(e, rc);
    }
    e->cv_idle.notify_all();
    if (w->period_ms > 0) {
      std::unique_lock<std::mutex> lk(e->mu);
      const auto until = round_start + std::chrono::microseconds((int64_t)(w->period_ms * 1000.0));
      e->cv_integ.wait_until(lk, until, [&] { return e->stop; });
      if (e->stop) return;
    }
  }
}

void integration_loop(Engine* e) {
  for (;;) {
    const auto until = Clock::now() + std::chrono::microseconds((int64_t)(e->period_ms * 1000.0));
    const int32_t rc = vdbgpu_integrate(e->map);
    std::unique_lock<std::mutex> lk(e->mu);
    ++e->integrations;
    note_error(e, rc);
    if (e->period_ms > 0)
      e->cv_integ.wait_until(lk, until, [&] { return e->stop; });
    if (e->stop) return;
  }
}

std::shared_ptr<Engine> find_engine(vdbgpu_map* m) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_engines.find(m);
  return it == g_engines.end() ? nullptr : it->second;
}

}  // namespace

extern "C" {

int32_t vdbgpu_async_start(vdbgpu_map* m, double accumulation_period_s) {
  if (!m || !(accumulation_period_s >= 0)) return VDBGPU_BAD_ARGUMENT;
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_engines.count(m)) return VDBGPU_OK;
  auto e = std::make_shared<Engine>();
  e->map = m;
  e->period_ms = accumulation_period_s * 1000.0;
  Engine* ep = e.get();
  e->integ = std::thread(integration_loop, ep);
  g_engines[m] = std::move(e);
  return VDBGPU_OK;
}

int32_t vdbgpu_async_submit(vdbgpu_map* m, const char* source_id, const float* xyz, int64_t n,
                            int64_t stride_bytes, const double origin[3], double max_rate) {
  if (!m || !source_id || n < 0 || (n && !xyz) || stride_bytes < 12 || (stride_bytes & 3) || !origin)
    return VDBGPU_BAD_ARGUMENT;
  const std::shared_ptr<Engine> e = find_engine(m);
  if (!e) return VDBGPU_NOT_CONFIGURED;
  // addDataToAccumulate on an unknown source is rejected at once (VDBMapping.hpp:359-364)
  if (vdbgpu_has_source(m, source_id) == VDBGPU_UNKNOWN_SOURCE) return VDBGPU_UNKNOWN_SOURCE;
  Slot s;
  s.xyz.resize((size_t)n * 4);
  const char* src = reinterpret_cast<const char*>(xyz);
  for (int64_t i = 0; i < n; ++i) {
    std::memcpy(&s.xyz[(size_t)i * 4], src + i * stride_bytes, 12);
    s.xyz[(size_t)i * 4 + 3] = 1.0f;
  }
  std::memcpy(s.origin, origin, sizeof s.origin);
  s.full = true;
  Worker* w;
  {
    std::lock_guard<std::mutex> lk(e->mu);
    if (e->stop) return VDBGPU_NOT_CONFIGURED;  // stopped while this call ran
    auto it = e->workers.find(source_id);
    if (it == e->workers.end()) {
      auto nw = std::make_unique<Worker>();
      nw->id = source_id;
      nw->period_ms = max_rate > 0 ? 1000.0 / max_rate : 0.0;
      w = nw.get();
      e->workers[source_id] = std::move(nw);
      w->th = std::thread(worker_loop, e.get(), w);
    } else {
      w = it->second.get();
    }
    if (w->slot.full) ++e->replaced;  // latest only: the older cloud is dropped
    w->slot = std::move(s);
    ++e->submitted;
  }
  w->cv.notify_all();
  return VDBGPU_OK;
}

int32_t vdbgpu_async_flush(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  const std::shared_ptr<Engine> e = find_engine(m);
  if (!e) return VDBGPU_OK;
  {
    std::unique_lock<std::mutex> lk(e->mu);
    e->cv_idle.wait(lk, [&] {
      if (e->stop) return true;  // stopped: its workers take nothing more
      for (auto& kv : e->workers)
        if (kv.second->slot.full || kv.second->busy) return false;
      return true;
    });
    if (e->stop) return e->first_error;
  }
  const int32_t rc = vdbgpu_integrate(m);
  std::lock_guard<std::mutex> lk(e->mu);
  note_error(e.get(), rc);
  return e->first_error;
}

int32_t vdbgpu_async_counters(vdbgpu_map* m, int64_t out[4]) {
  if (!m || !out) return VDBGPU_BAD_ARGUMENT;
  const std::shared_ptr<Engine> e = find_engine(m);
  if (!e) {
    out[0] = out[1] = out[2] = out[3] = 0;
    return VDBGPU_OK;
  }
  std::lock_guard<std::mutex> lk(e->mu);
  out[0] = e->submitted;
  out[1] = e->replaced;
  out[2] = e->accumulated;
  out[3] = e->integrations;
  return VDBGPU_OK;
}

int32_t vdbgpu_async_stop(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  std::shared_ptr<Engine> e;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_engines.find(m);
    if (it == g_engines.end()) return VDBGPU_OK;
    e = std::move(it->second);
    g_engines.erase(it);
  }
  {
    std::lock_guard<std::mutex> lk(e->mu);
    e->stop = true;
  }
  for (auto& kv : e->workers) kv.second->cv.notify_all();
  e->cv_integ.notify_all();
  e->cv_idle.notify_all();  // a flush waiting on the workers returns
  for (auto& kv : e->workers)
    if (kv.second->th.joinable()) kv.second->th.join();
  if (e->integ.joinable()) e->integ.join();
  return e->first_error;
}

}  // extern "C"
