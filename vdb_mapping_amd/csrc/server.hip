// Multi-map serving loop over the C ABI (include/vdbgpu.h, vdbgpu_server_*).
//
// Each robot stream owns a device map (its own vdbgpu_map: own HIP streams,
// own tables). A worker thread per map takes every request queued for its
// map when it is free and integrates them in one call: one scan through the
// plain insert path, several through vdbgpu_insert_batch, whose result equals
// inserting them one by one in arrival order. So requests that arrive while a
// map is busy are admitted as a batch at its next turn, and different maps'
// work overlaps on the device.
//
// HBM admission. The footprint of a lane is its map's device memory: the map
// itself plus the per-scan working memory it keeps for reuse (scratch). A batch
// is admitted when the summed footprint, the reservations of the batches in
// flight and this batch's projected need fit the budget. The need is projected
// from the batch's largest scan and its total points: scratch (sized for one
// scan at a time, reused by the next) of s bytes per point of the largest scan,
// plus map growth of g bytes per point in total, less the scratch the lane
// already holds; s and g start at 2 KiB and 1 KiB, then follow what finished
// batches needed (totals over totals, with 25% / 50% headroom; s only from
// batches that sized their scratch afresh), so one allocation step of a small
// scan does not inflate them. A batch that does not fit waits: it gives back
// its own lane's scratch, asks idle lanes to give back theirs, and is admitted
// as batches finish and memory is released. When nothing is in flight and no
// lane holds scratch, waiting cannot help: the batch runs alone if the maps
// themselves are within the budget (the projection is an estimate), and fails
// with VDBGPU_OVER_BUDGET if they already fill it (maps never shrink; the
// caller may reset or destroy maps and resubmit). Latency of a request =
// submit to completion (device work finished).
#include <hip/hip_runtime.h>

#include <algorithm>
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
  int64_t map_bytes = 0;  // device memory of the map without its scratch (last known)
  int64_t scratch = 0;    // releasable working memory the map holds (last known)
  int64_t reserved = 0;   // admission reservation of the batch in flight
  bool busy = false;      // a batch of this lane is in flight
};

constexpr double kInitialScratchPerPoint = 2048.0;
constexpr double kInitialGrowthPerPoint = 1024.0;

}  // namespace

struct vdbgpu_server {
  std::mutex mu;
  std::condition_variable cv_work;  // new requests, memory wanted, stop
  std::condition_variable cv_mem;   // a batch finished or scratch was released
  std::condition_variable cv_done;  // a request finished
  std::vector<Lane> lanes;
  int64_t budget = 0;
  int32_t max_batch = 64;
  bool stop = false;
  int64_t next_id = 0;
  int64_t submitted = 0, completed = 0;
  int32_t in_flight = 0;   // batches on the device
  int32_t mem_waiters = 0; // batches waiting for memory
  double scratch_per_point = kInitialScratchPerPoint;  // of a batch's largest scan
  double growth_per_point = kInitialGrowthPerPoint;    // of all its points
  double seen_scratch = 0, seen_largest = 0;           // finished batches that sized scratch afresh
  double seen_growth = 0, seen_points = 0;             // all finished batches
  int32_t first_error = VDBGPU_OK;
  std::vector<vdbgpu_request_record> records;
  int device = 0;
};

namespace {

// (map bytes, scratch bytes) of a lane's map; callers hold no server lock.
void measure(Lane& lane, int64_t* map_bytes, int64_t* scratch) {
  vdbgpu_stats st;
  if (vdbgpu_get_stats(lane.map, &st) == VDBGPU_OK) {
    *scratch = (int64_t)st.scratch_bytes;
    *map_bytes = (int64_t)st.device_bytes - *scratch;
  }
}

int64_t held(const vdbgpu_server* s) {
  int64_t t = 0;
  for (const Lane& l : s->lanes) t += l.map_bytes + l.scratch + l.reserved;
  return t;
}

int64_t maps_only(const vdbgpu_server* s) {
  int64_t t = 0;
  for (const Lane& l : s->lanes) t += l.map_bytes;
  return t;
}

bool scratch_elsewhere(const vdbgpu_server* s, const Lane* self) {
  for (const Lane& l : s->lanes)
    if (&l != self && l.scratch > 0 && !l.busy) return true;
  return false;
}

// Give back a lane's scratch (lock released during the call).
void release(vdbgpu_server* s, Lane& lane, std::unique_lock<std::mutex>& lk) {
  lk.unlock();
  int64_t mb = lane.map_bytes, sc = 0;
  const bool ok = vdbgpu_release_scratch(lane.map) == VDBGPU_OK;
  if (ok) measure(lane, &mb, &sc);
  lk.lock();
  if (ok) {
    // what a release leaves behind (update sets with pending work) counts as map memory
    lane.map_bytes = mb + sc;
    lane.scratch = 0;
  }
  s->cv_mem.notify_all();
}

void worker_loop(vdbgpu_server* s, int idx) {
  (void)hipSetDevice(s->device);
  Lane& lane = s->lanes[idx];
  std::vector<std::unique_ptr<Request>> batch;
  std::vector<float> cat;
  std::vector<int64_t> counts;
  std::vector<double> origins;
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    s->cv_work.wait(lk, [&] {
      return s->stop || !lane.queue.empty() || (s->mem_waiters > 0 && lane.scratch > 0);
    });
    if (lane.queue.empty()) {
      if (s->stop) return;
      release(s, lane, lk);  // idle while other batches wait for memory
      continue;
    }
    const size_t take = std::min<size_t>(lane.queue.size(), (size_t)s->max_batch);
    int64_t points = 0, largest = 0;
    for (size_t i = 0; i < take; ++i) {
      const int64_t n = (int64_t)lane.queue.front()->xyz.size() / 4;
      points += n;
      largest = std::max(largest, n);
      batch.push_back(std::move(lane.queue.front()));
      lane.queue.pop_front();
    }
    // admission by projected footprint
    int32_t waits = 0;
    bool reject = false;
    for (;;) {
      const int64_t need = std::max<int64_t>(
          0, (int64_t)(s->scratch_per_point * (double)largest + s->growth_per_point * (double)points) -
                 lane.scratch);
      if (held(s) + need <= s->budget) {
        lane.reserved = need;
        break;
      }
      if (lane.scratch > 0) {  // waiting: give back this lane's own scratch first
        release(s, lane, lk);
        continue;
      }
      if (s->in_flight == 0 && (s->stop || !scratch_elsewhere(s, &lane))) {
        // nothing in flight and nothing left to free: waiting cannot help
        if (maps_only(s) < s->budget)
          lane.reserved = need;  // runs alone
        else
          reject = true;
        break;
      }
      ++waits;
      ++s->mem_waiters;
      s->cv_work.notify_all();  // idle lanes holding scratch release it
      s->cv_mem.wait(lk);
      --s->mem_waiters;
    }
    lane.busy = !reject;
    if (!reject) ++s->in_flight;
    const int64_t map_before = lane.map_bytes, scratch_before = lane.scratch;
    lk.unlock();

    int32_t rc;
    if (reject) {
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
    if (rc == VDBGPU_OK && !reject) rc = vdbgpu_synchronize(lane.map);
    int64_t mb = lane.map_bytes, sc = lane.scratch;
    if (!reject) measure(lane, &mb, &sc);
    const auto t_done = Clock::now();

    lk.lock();
    if (!reject) {
      lane.map_bytes = mb;
      lane.scratch = sc;
      lane.reserved = 0;
      lane.busy = false;
      --s->in_flight;
      if (rc == VDBGPU_OK && points > 0) {
        s->seen_growth += (double)std::max<int64_t>(0, mb - map_before);
        s->seen_points += (double)points;
        s->growth_per_point = 1.5 * s->seen_growth / s->seen_points;
        if (scratch_before == 0 && largest > 0) {  // the batch sized the lane's scratch afresh
          s->seen_scratch += (double)sc;
          s->seen_largest += (double)largest;
          s->scratch_per_point = 1.25 * s->seen_scratch / s->seen_largest;
        }
      }
    }
    for (const auto& r : batch) {
      vdbgpu_request_record rec;
      rec.request_id = r->id;
      rec.map_index = idx;
      rec.points = (int64_t)r->xyz.size() / 4;
      rec.batch_size = (int32_t)batch.size();
      rec.status = rc;
      rec.mem_waits = waits;
      rec.latency_ms = std::chrono::duration<double, std::milli>(t_done - r->t_submit).count();
      s->records.push_back(rec);
    }
    if (rc != VDBGPU_OK && s->first_error == VDBGPU_OK) s->first_error = rc;
    s->completed += (int64_t)batch.size();
    batch.clear();
    if (s->mem_waiters > 0 && lane.scratch > 0) release(s, lane, lk);  // memory is wanted elsewhere
    s->cv_mem.notify_all();
    s->cv_done.notify_all();
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
      // record sets per map
      vdbgpu_tuning t;
      rc = vdbgpu_get_tuning(s->lanes[i].map, &t);
      t.defer = 1;
      if (rc == VDBGPU_OK) rc = vdbgpu_set_tuning(s->lanes[i].map, &t);
    }
    if (rc == VDBGPU_OK) measure(s->lanes[i], &s->lanes[i].map_bytes, &s->lanes[i].scratch);
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
  int64_t t = 0;
  for (const Lane& l : s->lanes) t += l.map_bytes + l.scratch;
  return t;
}

void vdbgpu_server_destroy(vdbgpu_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_mem.notify_all();
  for (Lane& l : s->lanes)
    if (l.worker.joinable()) l.worker.join();
  for (Lane& l : s->lanes) vdbgpu_destroy(l.map);
  delete s;
}

}  // extern "C"
