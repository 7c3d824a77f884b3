// Multi-map serving loop over the C ABI (include/vdbgpu.h, vdbgpu_server_*).
//
// Each robot stream owns a device map (its own vdbgpu_map: own HIP stream,
// own tables). A worker thread per map takes every request queued for its
// map when it is free and integrates them in one call: one scan through the
// plain insert path, several through vdbgpu_insert_batch, whose result equals
// inserting them one by one in arrival order. So requests that arrive while a
// map is busy are admitted as a batch at its next turn, and different maps'
// work overlaps on the device. Admission by HBM footprint: a map that already
// holds device memory gets new work only while the summed footprint of all
// maps is within the budget; otherwise the batch fails with
// VDBGPU_OVER_BUDGET (maps never shrink, so waiting could not help). Latency
// of a request = submit to completion (device work finished).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/vdbgpu.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t id;
  std::vector<float> xyz;  // 4 floats per point (x, y, z, pad)
  double origin[3];
  Clock::time_point t_submit;
};

struct Lane {
  vdbgpu_map* map = nullptr;
  std::deque<std::unique_ptr<Request>> queue;
  std::thread worker;
  int64_t device_bytes = 0;  // last known footprint
};

}  // namespace

struct vdbgpu_server {
  std::mutex mu;
  std::condition_variable cv_work;   // new requests / stop
  std::condition_variable cv_done;   // a request finished / memory freed
  std::vector<Lane> lanes;
  int64_t budget = 0;
  int32_t max_batch = 64;
  bool stop = false;
  int64_t next_id = 0;
  int64_t submitted = 0, completed = 0, in_flight = 0;
  int32_t first_error = VDBGPU_OK;
  std::vector<vdbgpu_request_record> records;
  int device = 0;
};

namespace {

int64_t footprint(vdbgpu_server* s) {
  int64_t t = 0;
  for (const Lane& l : s->lanes) t += l.device_bytes;
  return t;
}

void worker_loop(vdbgpu_server* s, int idx) {
  (void)hipSetDevice(s->device);
  Lane& lane = s->lanes[idx];
  std::vector<std::unique_ptr<Request>> batch;
  std::vector<float> cat;
  std::vector<int64_t> counts;
  std::vector<double> origins;
  bool over_budget = false;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(s->mu);
      s->cv_work.wait(lk, [&] { return s->stop || !lane.queue.empty(); });
      if (s->stop && lane.queue.empty()) return;
      const size_t take = std::min<size_t>(lane.queue.size(), (size_t)s->max_batch);
      for (size_t i = 0; i < take; ++i) {
        batch.push_back(std::move(lane.queue.front()));
        lane.queue.pop_front();
      }
      s->in_flight += (int64_t)batch.size();
      // HBM admission: a map that already holds memory may only grow while the
      // maps' summed footprint is within the budget (maps never shrink, so
      // waiting would not help: the batch is rejected instead)
      over_budget = lane.device_bytes > 0 && footprint(s) > s->budget;
    }
    int32_t rc;
    if (over_budget) {
      rc = VDBGPU_OVER_BUDGET;
    } else if (batch.size() == 1) {
      const Request& r = *batch[0];
      rc = vdbgpu_insert(lane.map, "stream", r.xyz.data(), (int64_t)r.xyz.size() / 4, 16, r.origin, 0);
    } else {
      cat.clear();
      counts.clear();
      origins.clear();
      for (const auto& r : batch) {
        cat.insert(cat.end(), r->xyz.begin(), r->xyz.end());
        counts.push_back((int64_t)r->xyz.size() / 4);
        origins.insert(origins.end(), r->origin, r->origin + 3);
      }
      rc = vdbgpu_insert_batch(lane.map, "stream", (int32_t)batch.size(), cat.data(), counts.data(), 16,
                               origins.data(), nullptr, 0);
    }
    if (rc == VDBGPU_OK && !over_budget) rc = vdbgpu_synchronize(lane.map);
    vdbgpu_stats st;
    const int64_t bytes = vdbgpu_get_stats(lane.map, &st) == VDBGPU_OK ? (int64_t)st.device_bytes : 0;
    const auto t_done = Clock::now();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      lane.device_bytes = bytes;
      for (const auto& r : batch) {
        vdbgpu_request_record rec;
        rec.request_id = r->id;
        rec.map_index = idx;
        rec.points = (int64_t)r->xyz.size() / 4;
        rec.batch_size = (int32_t)batch.size();
        rec.status = rc;
        rec.latency_ms = std::chrono::duration<double, std::milli>(t_done - r->t_submit).count();
        s->records.push_back(rec);
      }
      if (rc != VDBGPU_OK && s->first_error == VDBGPU_OK) s->first_error = rc;
      s->completed += (int64_t)batch.size();
      s->in_flight -= (int64_t)batch.size();
    }
    batch.clear();
    s->cv_done.notify_all();
    s->cv_work.notify_all();  // memory may have been freed for other lanes
  }
}

}  // namespace

extern "C" {

int32_t vdbgpu_server_create(int32_t n_maps, double resolution, const vdbgpu_config* config,
                             double max_range, int64_t hbm_budget_bytes, int32_t max_batch,
                             vdbgpu_server** out) {
  if (!out || n_maps <= 0 || !config || max_batch <= 0) return VDBGPU_BAD_ARGUMENT;
  *out = nullptr;
  auto* s = new vdbgpu_server();
  (void)hipGetDevice(&s->device);
  s->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : INT64_MAX;
  s->max_batch = std::min<int32_t>(max_batch, 1 << 20);
  s->lanes.resize(n_maps);
  for (int i = 0; i < n_maps; ++i) {
    int32_t rc = vdbgpu_create(resolution, &s->lanes[i].map);
    if (rc == VDBGPU_OK) rc = vdbgpu_set_config(s->lanes[i].map, config);
    if (rc == VDBGPU_OK) rc = vdbgpu_add_source(s->lanes[i].map, "stream", max_range, 0.0);
    if (rc == VDBGPU_OK) {
      // every request is synchronised on its own: no deferred scans, and two
      // record sets per map within the HBM budget
      vdbgpu_tuning t;
      rc = vdbgpu_get_tuning(s->lanes[i].map, &t);
      t.defer = 1;
      if (rc == VDBGPU_OK) rc = vdbgpu_set_tuning(s->lanes[i].map, &t);
    }
    if (rc != VDBGPU_OK) {
      for (int j = 0; j <= i; ++j) vdbgpu_destroy(s->lanes[j].map);
      delete s;
      return rc;
    }
  }
  for (int i = 0; i < n_maps; ++i) s->lanes[i].worker = std::thread(worker_loop, s, i);
  *out = s;
  return VDBGPU_OK;
}

int32_t vdbgpu_server_submit(vdbgpu_server* s, int32_t map_index, const float* xyz, int64_t n,
                             int64_t stride_bytes, const double origin[3], int64_t* request_id) {
  if (!s || map_index < 0 || map_index >= (int32_t)s->lanes.size() || n < 0 || (n && !xyz) ||
      stride_bytes < 12 || (stride_bytes & 3) || !origin)
    return VDBGPU_BAD_ARGUMENT;
  auto r = std::make_unique<Request>();
  r->xyz.resize((size_t)n * 4);
  const char* src = reinterpret_cast<const char*>(xyz);
  for (int64_t i = 0; i < n; ++i) {
    std::memcpy(&r->xyz[(size_t)i * 4], src + i * stride_bytes, 12);
    r->xyz[(size_t)i * 4 + 3] = 1.0f;
  }
  std::memcpy(r->origin, origin, sizeof r->origin);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    if (request_id) *request_id = r->id;
    r->t_submit = Clock::now();
    s->lanes[map_index].queue.push_back(std::move(r));
    ++s->submitted;
  }
  s->cv_work.notify_all();
  return VDBGPU_OK;
}

int32_t vdbgpu_server_drain(vdbgpu_server* s) {
  if (!s) return VDBGPU_BAD_ARGUMENT;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->completed == s->submitted; });
  return s->first_error;
}

int64_t vdbgpu_server_records(vdbgpu_server* s, vdbgpu_request_record* out, int64_t cap) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  const int64_t n = (int64_t)s->records.size();
  if (!out) return n;
  const int64_t k = std::min(n, cap);
  std::copy(s->records.begin(), s->records.begin() + k, out);
  return k;
}

vdbgpu_map* vdbgpu_server_map(vdbgpu_server* s, int32_t map_index) {
  if (!s || map_index < 0 || map_index >= (int32_t)s->lanes.size()) return nullptr;
  return s->lanes[map_index].map;
}

int64_t vdbgpu_server_device_bytes(vdbgpu_server* s) {
  if (!s) return -1;
  std::lock_guard<std::mutex> lk(s->mu);
  return footprint(s);
}

void vdbgpu_server_destroy(vdbgpu_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (Lane& l : s->lanes)
    if (l.worker.joinable()) l.worker.join();
  for (Lane& l : s->lanes) vdbgpu_destroy(l.map);
  delete s;
}

}  // extern "C"
