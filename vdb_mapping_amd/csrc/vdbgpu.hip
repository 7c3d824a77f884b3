// Host runtime and C ABI of the gfx950 occupancy-integration path
// (declared in include/vdbgpu.h).
//
// One vdbgpu_map owns: a HIP stream (default priority), a mutex, the device
// leaf store + leaf hash, one update set per input source, per-scan scratch,
// the DDA crossing table and a pinned status block. The per-scan path is
//   accumulate: [H2D points] -> k_prep -> k_scan_blocks -> k_traverse -> k_compact
//               -> one status read (the only host sync of a scan)
//   integrate : k_map_insert -> k_apply per source, in source-id order (no sync)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/vdbgpu.h"
#include "device_common.h"
#include "kernels.h"

using namespace vdbgpu;

namespace {

constexpr uint32_t kMaxTableN = 8192;       // longest per-axis ray extent in voxels
constexpr uint32_t kMinUpdateCap = 1u << 12;
constexpr uint32_t kMaxRays = 1u << 30;
constexpr int kTraverseMaxBlock = 256;  // __launch_bounds__ of k_traverse

enum TimedKind { kPrep = 0, kTraverse = 1, kCompact = 2, kApply = 3, kRaytrace = 4, kWalk = 5 };

struct TimedPair {
  int kind;
  hipEvent_t a, b;
};

struct Source {
  std::string id;
  double max_range = 0;
  unsigned long long* keys = nullptr;
  unsigned long long* masks = nullptr;
  uint32_t* list = nullptr;
  uint32_t cap = 0;
  uint32_t pending = 0;         // listed update leaves not yet integrated
  bool has_isect = false;       // volume_ray_intersector set (fast mode); none after addInputSource
};

uint64_t next_pow2(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

}  // namespace

// A device array that grows in place: an address range reserved once and
// backed by physical chunks mapped at its end (HIP virtual memory).
struct GrowArray {
  char* base = nullptr;
  size_t reserved = 0, mapped = 0;
  std::vector<std::pair<hipMemGenericAllocationHandle_t, size_t>> chunks;
};

struct vdbgpu_map {
  std::mutex mu;
  int device = 0;
  int num_cu = 256;
  hipStream_t stream = nullptr;  // map stream: every map read / update, the hash path
  hipStream_t cast = nullptr;    // brick path casting (prep .. record list) of plain inserts
  double res = 0, inv = 0;

  // config (VDBMapping.hpp:1456-1469, OccupancyVDBMapping.hpp:59-89)
  bool config_set = false;
  double max_range = 0;
  bool fast_mode = false;
  float lo[6] = {0, 0, 0, 0, 0, 0};  // l_hit, l_miss, t_min, t_max, c_max, c_min
  std::map<std::string, Source> sources;

  // leaf store
  unsigned long long* mkeys = nullptr;
  uint32_t* midx = nullptr;
  uint32_t mcap = 0;
  float* values = nullptr;
  unsigned long long* active = nullptr;
  unsigned long long* leaf_keys = nullptr;
  uint8_t* dirty = nullptr;
  uint32_t lcap = 0;
  // the four leaf-pool arrays in reserved address ranges that grow in place
  // (physical chunks mapped at the end; vmm = 1), or plain allocations that
  // grow by copy (vmm = -1); 0 until the first leaf
  int vmm = 0;
  GrowArray g_values, g_active, g_keys, g_dirty;
  uint64_t vmm_leaves = 0;  // leaves the reserved ranges can hold
  size_t vmm_gran = 0;
  uint64_t leaf_count_ub = 0;  // upper bound of the device leaf count
  // Leaves that map-stream work enqueued since the last full sync may create,
  // as (sequence number, count) in enqueue order: a leaf-count snapshot taken
  // on the cast stream after it waited for an event recorded at sequence q
  // misses at most the entries after q.
  std::deque<std::pair<uint64_t, uint64_t>> ledger;
  uint64_t ledger_seq = 0;
  bool touched = false;        // any voxel modify since creation/reset
  // leaf-level tiles of the last pruneGrid (tile_kernels.hip): tile bit, state
  // bit and value per pool leaf below tcap, and the list of the leaves pruned
  uint32_t* tbits = nullptr;
  uint32_t* ton = nullptr;
  float* tval = nullptr;
  uint32_t* tlist = nullptr;
  uint32_t tcap = 0;    // leaves covered, a multiple of 32 (0: no tile arrays)
  uint32_t ntiles = 0;  // entries of tlist; 0 when the map holds no tile
  bool snap_dirty = true;      // the map changed since the last topology snapshot
  unsigned long long* noop4096 = nullptr;  // root children of updates that created no leaf
  static constexpr uint32_t kNoopCap = 4096;

  // raytrace topology snapshot (updateVolumeRayIntersectors)
  bool has_isect = false;
  RtView rt{};
  uint32_t rt_leaf_cap = 0, rt_n128_cap = 0, rt_n4096_cap = 0;
  int* d_bbox = nullptr;
  // raytrace query staging
  // batch: scan table, control block, fold groups
  ScanInfo* d_scans = nullptr;
  BatchCtl* d_ctl = nullptr;
  ScanInfo* h_scans = nullptr;
  BatchCtl* h_ctl = nullptr;
  unsigned long long* gkeys = nullptr;
  unsigned long long* gscans = nullptr;
  unsigned int* gneed = nullptr;
  uint32_t* glist = nullptr;
  uint32_t* gidx = nullptr;
  uint32_t gcap = 0;
  double* q_in = nullptr;
  uint8_t* q_succ = nullptr;
  double* q_end = nullptr;
  int64_t q_cap = 0;

  // per-scan scratch
  float* d_points = nullptr;
  size_t points_bytes = 0;
  int4* rayinfo = nullptr;
  uint32_t* pre = nullptr;
  uint32_t* bsum = nullptr;
  uint64_t* boff = nullptr;
  uint32_t ray_cap = 0;
  uint64_t* roff = nullptr;      // global visit offset of each ray
  uint32_t* first_ray = nullptr; // ray holding the first visit of each chunk
  uint64_t chunk_cap = 0;
  uint32_t* eidx = nullptr;
  uint32_t eidx_cap = 0;
  // brick path (plain inserts): brick table, segments, work items, shared
  // bitmaps of split bricks, update-leaf records
  unsigned long long* bkeys = nullptr;
  uint32_t* bcnt = nullptr;  // [0, bcap): segment counts, [bcap, 2 bcap): fill cursors
  uint64_t* bbsum = nullptr;
  uint64_t* bboff = nullptr;
  uint32_t bcap = 0;
  uint4* segs = nullptr;
  uint32_t seg_cap = 0;
  BrickItem* items = nullptr;
  uint32_t item_cap = 0;
  uint32_t* split_slot = nullptr;
  unsigned long long* shared_bits = nullptr;
  // per-workgroup brick tables of the segment count (kSegTab entries per
  // workgroup), reread by the segment fill
  unsigned long long* wkey = nullptr;
  unsigned long long* wsc = nullptr;
  uint64_t wtab_cap = 0;
  uint32_t split_cap = 0;
  // update-leaf records: the walk fills positional records (64 positions per
  // work item, cast stream only), listed and copied densely into one record
  // set of a ring; the cast of scan i+1 fills one set while the map stream
  // still applies earlier scans from the others (two sets, or 2 defer + 2
  // when scans are applied in groups)
  static constexpr int kRecSets = 2 * kDeferMax + 2;
  unsigned long long* rkeys = nullptr;
  unsigned long long* rmasks = nullptr;
  uint32_t* rlist = nullptr;  // filled record positions (dense)
  unsigned long long* dkeys[kRecSets] = {};   // record set: leaf keys
  unsigned long long* dmasks[kRecSets] = {};  // record set: 16 mask words per record
  uint32_t* reidx[kRecSets] = {};             // record set: map leaf index of each record
  uint32_t dcap[kRecSets] = {};
  uint32_t dense_hint = 0;     // record-set size for a first use (largest scan seen, +25%)
  uint32_t* ident = nullptr;   // 0, 1, 2, ...: the list of a dense record set
  uint32_t ident_cap = 0;
  hipEvent_t rec_free[kRecSets] = {};  // recorded on the map stream after the last read
  // leaf count after the map insert of each record set's scan (pinned, copied
  // on the map stream before ins_done), for a tight leaf-count bound
  hipEvent_t ins_done[kRecSets] = {};
  uint32_t* h_lc = nullptr;  // pinned; written by the device through d_lc
  uint32_t* d_lc = nullptr;
  uint64_t ins_seq[kRecSets] = {};  // ledger_seq after that insert
  std::deque<int> ins_fifo;         // sets with an insert snapshot in flight, oldest first
  uint64_t rec_seq[kRecSets] = {};     // ledger_seq when rec_free was recorded
  int rec_cur = 0;
  int rec_sets = kRecSets;  // 2 when tuning.defer <= 1, else 2 defer + 2
  // deferred scans (tuning.defer > 1): record sets whose map insert ran and
  // whose apply waits for the group pass, in insert order
  int npend = 0;
  // a scan whose cast is done and whose map-stream work (map insert, then the
  // apply or a place in the group) is enqueued by the next insert while that
  // insert's cast runs, or by any other call first
  bool staged = false;
  int staged_set = 0;
  uint32_t staged_recs = 0;
  int pend_set[kDeferMax] = {};
  uint32_t pend_recs[kDeferMax] = {};
  uint32_t* gslot = nullptr;  // per map leaf: kNone, or the record that opened its group
  uint32_t gslot_cap = 0;
  uint32_t* rgrp = nullptr;   // per (scan, record): the group it opened
  uint64_t rgrp_cap = 0;
  uint32_t* gl = nullptr;     // per group: map leaf index
  uint32_t* grec = nullptr;   // per group: record position of each scan (kDeferMax)
  uint64_t grp_cap = 0;
  unsigned int* ngroups = nullptr;
  uint32_t rec_cap = 0;        // record positions: 64 per work item and per split brick
  unsigned long long* leaf_mask = nullptr;  // filled positions per 64-position group
  uint64_t* lbsum = nullptr;
  uint64_t* lboff = nullptr;
  uint32_t group_cap = 0;      // groups: work items + split bricks
  double* tbl = nullptr;
  uint32_t tbl_n = 0;

  // direct edits (addPointsToGrid / removePointsFromGrid) and artificial areas
  Source edit;                 // marked voxels of one edit call (cleared after it)
  Source art, art_tmp;         // artificial grid (VM:132); art_tmp receives a new one
  uint32_t art_count = 0;      // listed leaves of the artificial grid (art.list)
  uint32_t* art_eidx = nullptr;
  uint32_t art_eidx_cap = 0;
  bool art_present = false;    // m_artificial_areas_present (VM:1529)
  bool art_leaves = false;     // every artificial leaf exists in the map
  std::vector<double> art_verts;  // walls added since the last restore: vertices (x, y, z)
  std::vector<int> art_walls;     // and {a, b, k0, k1} per wall

  // status: [0] scan, [1] map (leaf_count), [2] auxiliary, [3] cumulative counters
  Status* d_st = nullptr;
  Status* h_st = nullptr;
  // brick-path casts whose status the host has not read: two host slots
  Status* h_cst = nullptr;
  RecDest* d_dest = nullptr;  // the record set the next cast copies into
  hipEvent_t cst_done[2] = {};
  int cst_next = 0;
  uint64_t cast_seq = 0;    // brick casts enqueued (re-runs included)
  // brick casts of batched inserts replayed from captured graphs, one per
  // exact set of launch arguments (least recently used dropped first)
  struct CastGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
    uint64_t used;
  };
  static constexpr size_t kCastGraphs = 48;
  std::vector<CastGraph> cast_graphs;
  uint64_t cast_graph_clock = 0;
  bool graph_casts = false;  // set while vdbgpu_insert_batch streams scans (VDBGPU_GRAPH=1)
  uint64_t cast_epoch = 0;  // table overflows found (each clears the split-brick bitmaps)

  vdbgpu_tuning tuning{256, 32, 16, 0.5, 0, kDeferMax};
  bool profiling = false;
  std::vector<TimedPair> timed;
  std::vector<hipEvent_t> event_pool;
  vdbgpu_stats stats{};
  uint64_t device_bytes = 0;
  std::string last_error;
};

namespace {

#define CK(expr)                                                                 \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      m->last_error = std::string(#expr) + ": " + hipGetErrorString(e_);         \
      return VDBGPU_DEVICE_ERROR;                                                \
    }                                                                            \
  } while (0)

template <class T>
hipError_t dalloc(vdbgpu_map* m, T** p, size_t count) {
  *p = nullptr;
  if (count == 0) return hipSuccess;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e == hipSuccess) m->device_bytes += count * sizeof(T);
  return e;
}

template <class T>
void dfree(vdbgpu_map* m, T** p, size_t count) {
  if (*p) {
    (void)hipFree(*p);
    m->device_bytes -= count * sizeof(T);
  }
  *p = nullptr;
}

hipEvent_t get_event(vdbgpu_map* m) {
  if (!m->event_pool.empty()) {
    hipEvent_t e = m->event_pool.back();
    m->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

struct Timer {
  vdbgpu_map* m;
  int kind;
  hipStream_t s;
  hipEvent_t a = nullptr;
  Timer(vdbgpu_map* mm, int k, hipStream_t st = nullptr) : m(mm), kind(k), s(st ? st : mm->stream) {
    if (m->profiling) {
      a = get_event(m);
      (void)hipEventRecord(a, s);
    }
  }
  ~Timer() {
    if (a) {
      hipEvent_t b = get_event(m);
      (void)hipEventRecord(b, s);
      m->timed.push_back({kind, a, b});
    }
  }
};

// Fold finished event pairs into the stats (pairs still in flight on the
// other stream stay queued).
void collect_timing(vdbgpu_map* m) {
  std::vector<TimedPair> pending;
  for (auto& t : m->timed) {
    if (hipEventQuery(t.b) != hipSuccess) {
      pending.push_back(t);
      continue;
    }
    float ms = 0;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
      switch (t.kind) {
        case kPrep: m->stats.ms_prep += ms; m->stats.n_prep++; break;
        case kTraverse: m->stats.ms_traverse += ms; m->stats.n_traverse++; break;
        case kCompact: m->stats.ms_compact += ms; m->stats.n_compact++; break;
        case kRaytrace: m->stats.ms_raytrace += ms; m->stats.n_raytrace++; break;
        case kWalk: m->stats.ms_walk += ms; m->stats.n_walk++; break;
        default: m->stats.ms_apply += ms; m->stats.n_apply++; break;
      }
    }
    m->event_pool.push_back(t.a);
    m->event_pool.push_back(t.b);
  }
  m->timed.swap(pending);
}

int fail(vdbgpu_map* m, int code, const std::string& msg) {
  m->last_error = msg;
  return code;
}

void ledger_add(vdbgpu_map* m, uint64_t leaves) {
  if (leaves) m->ledger.emplace_back(++m->ledger_seq, leaves);
}

// Leaves of the ledger entries after sequence q.
uint64_t ledger_since(const vdbgpu_map* m, uint64_t q) {
  uint64_t n = 0;
  for (auto it = m->ledger.rbegin(); it != m->ledger.rend() && it->first > q; ++it) n += it->second;
  return n;
}

// Upper bound of the leaf count once every enqueued map insert has run, taken
// while the map stream may still be busy: the leaf count `cast_lc` that the
// cast stream read after waiting for an event at ledger position cast_q, or
// the newest map-insert snapshot already complete (no wait), each plus the
// leaves enqueued after it; the smaller bound is used. Both kinds of
// positions only grow, so ledger entries before both are dropped.
uint64_t leaf_bound(vdbgpu_map* m, uint32_t cast_lc, uint64_t cast_q) {
  uint64_t best = (uint64_t)cast_lc + ledger_since(m, cast_q);
  uint64_t q = cast_q;
  for (int k = (int)m->ins_fifo.size() - 1; k >= 0; --k) {
    const int Y = m->ins_fifo[k];
    if (hipEventQuery(m->ins_done[Y]) != hipSuccess) continue;
    best = std::min<uint64_t>(best, (uint64_t)m->h_lc[Y] + ledger_since(m, m->ins_seq[Y]));
    q = std::min(q, m->ins_seq[Y]);
    m->ins_fifo.erase(m->ins_fifo.begin(), m->ins_fifo.begin() + k);  // older snapshots are not needed
    break;
  }
  while (!m->ledger.empty() && m->ledger.front().first <= q) m->ledger.pop_front();
  return best;
}

int read_status(vdbgpu_map* m) {
  CK(hipMemcpyAsync(m->h_st, m->d_st, 4 * sizeof(Status), hipMemcpyDeviceToHost, m->stream));
  CK(hipStreamSynchronize(m->stream));
  collect_timing(m);
  m->leaf_count_ub = m->h_st[1].leaf_count;
  m->ledger.clear();
  m->ins_fifo.clear();
  if (m->h_st[1].map_full)
    return fail(m, VDBGPU_DEVICE_ERROR, "leaf pool overflow in a map insert (internal bound violated)");
  return VDBGPU_OK;
}

int grid_lanes(vdbgpu_map* m, int block) {
  const int lanes = m->num_cu * m->tuning.waves_per_cu * 64;
  return std::max(1, lanes / block);
}

// Traversal workgroups: LDS (36 KiB per workgroup) allows 4 resident per CU.
int traverse_grid(vdbgpu_map* m, int block) {
  const int per_cu = std::max(1, std::min(m->tuning.waves_per_cu * 64 / block, 4 * 256 / block));
  return m->num_cu * per_cu;
}

UpdView upd_view(const Source& s) { return UpdView{s.keys, s.masks, s.cap - 1}; }

MapView map_view(vdbgpu_map* m) {
  return MapView{m->mkeys, m->midx, m->mcap - 1, m->values, m->active, m->leaf_keys, m->dirty, m->lcap};
}

TileView tile_view(const vdbgpu_map* m) {
  return TileView{m->ntiles ? m->tbits : nullptr, m->ton, m->tval, m->tcap};
}

void free_tiles(vdbgpu_map* m) {
  dfree(m, &m->tbits, m->tcap / 32);
  dfree(m, &m->ton, m->tcap / 32);
  dfree(m, &m->tval, m->tcap);
  dfree(m, &m->tlist, m->tcap);
  m->tcap = 0;
  m->ntiles = 0;
}

// After an operation that wrote voxels: a tile whose voxels are no longer all
// (value, state) of the tile is a leaf again (pruneGrid tiles, OCC:173).
int check_tiles(vdbgpu_map* m) {
  if (!m->ntiles) return VDBGPU_OK;
  CK(launch_tile_check(m->stream, map_view(m), m->tlist, m->ntiles, tile_view(m)));
  return VDBGPU_OK;
}

// Host copies of the tile bits and state bits (empty when the map holds no
// tile); bit i stands for pool leaf i.
int tile_bits(vdbgpu_map* m, std::vector<uint32_t>& bits, std::vector<uint32_t>* on) {
  bits.clear();
  if (on) on->clear();
  if (!m->ntiles) return VDBGPU_OK;
  bits.resize(m->tcap / 32);
  CK(hipMemcpyAsync(bits.data(), m->tbits, m->tcap / 8, hipMemcpyDeviceToHost, m->stream));
  if (on) {
    on->resize(m->tcap / 32);
    CK(hipMemcpyAsync(on->data(), m->ton, m->tcap / 8, hipMemcpyDeviceToHost, m->stream));
  }
  CK(hipStreamSynchronize(m->stream));
  return VDBGPU_OK;
}

inline bool bit_at(const std::vector<uint32_t>& b, uint32_t i) {
  return (size_t)(i >> 5) < b.size() && ((b[i >> 5] >> (i & 31)) & 1u);
}

uint32_t count_bits(const std::vector<uint32_t>& b) {
  uint32_t n = 0;
  for (uint32_t w : b) n += (uint32_t)__builtin_popcount(w);
  return n;
}

// Leaves of the map (pool leaves that are not tiles); -1 on a device error.
int64_t live_leaves(vdbgpu_map* m) {
  std::vector<uint32_t> tb;
  if (read_status(m) || tile_bits(m, tb, nullptr)) return -1;
  return (int64_t)m->h_st[1].leaf_count - count_bits(tb);
}

// pruneGrid (OCC:173: tools::prune, tolerance 0) at the leaf level over the
// whole pool: every leaf of 512 equal values, all on or all off, is a tile.
int prune_tiles(vdbgpu_map* m) {
  int r;
  if ((r = read_status(m))) return r;
  const uint32_t L = m->h_st[1].leaf_count;
  const uint32_t cap = (uint32_t)(((uint64_t)L + 31) & ~31ull);
  if (cap != m->tcap) {
    free_tiles(m);
    if (cap == 0) return VDBGPU_OK;
    CK(dalloc(m, &m->tbits, cap / 32));
    CK(dalloc(m, &m->ton, cap / 32));
    CK(dalloc(m, &m->tval, cap));
    CK(dalloc(m, &m->tlist, cap));
    m->tcap = cap;
  }
  CK(hipMemsetAsync(m->tbits, 0, cap / 8, m->stream));
  CK(hipMemsetAsync(m->ton, 0, cap / 8, m->stream));
  CK(hipMemsetAsync(&m->d_st[2], 0, sizeof(Status), m->stream));
  CK(launch_prune(m->stream, map_view(m), L, TileView{m->tbits, m->ton, m->tval, cap}, m->tlist,
                  &m->d_st[2].entries));
  if ((r = read_status(m))) return r;
  m->ntiles = m->h_st[2].entries;
  if (!m->ntiles) free_tiles(m);
  m->snap_dirty = true;
  return VDBGPU_OK;
}

int alloc_update(vdbgpu_map* m, Source& s, uint32_t cap) {
  s.cap = cap;
  CK(dalloc(m, &s.keys, cap));
  CK(dalloc(m, &s.masks, (size_t)cap * 16));
  CK(dalloc(m, &s.list, cap));
  CK(launch_fill_u64(m->stream, s.keys, cap, kEmpty));
  CK(hipMemsetAsync(s.masks, 0, (size_t)cap * 16 * sizeof(unsigned long long), m->stream));
  return VDBGPU_OK;
}

void free_update(vdbgpu_map* m, Source& s) {
  dfree(m, &s.keys, s.cap);
  dfree(m, &s.masks, (size_t)s.cap * 16);
  dfree(m, &s.list, s.cap);
  s.cap = 0;
}

// Move a source's update set into a table of `cap` slots.
int grow_update(vdbgpu_map* m, Source& s, uint32_t cap) {
  Source t;
  int r = alloc_update(m, t, cap);
  if (r) return r;
  if (s.cap) {
    CK(launch_rehash_update(m->stream, std::min<int>(4096, (s.cap + 255) / 256), upd_view(s), s.cap,
                            upd_view(t), &m->d_st[2]));
  }
  free_update(m, s);  // hipFree waits for the rehash
  s.keys = t.keys;
  s.masks = t.masks;
  s.list = t.list;
  s.cap = t.cap;
  return VDBGPU_OK;
}

int ensure_rays(vdbgpu_map* m, uint32_t n) {
  if (n <= m->ray_cap) return VDBGPU_OK;
  const uint32_t cap = (uint32_t)std::min<uint64_t>(kMaxRays, next_pow2(n));
  const uint32_t nb_old = (m->ray_cap + kPrepBlock - 1) / kPrepBlock;
  dfree(m, &m->rayinfo, m->ray_cap);
  dfree(m, &m->pre, m->ray_cap);
  dfree(m, &m->roff, m->ray_cap);
  dfree(m, &m->bsum, nb_old);
  dfree(m, &m->boff, nb_old);
  const uint32_t nb = (cap + kPrepBlock - 1) / kPrepBlock;
  m->ray_cap = cap;
  CK(dalloc(m, &m->rayinfo, cap));
  CK(dalloc(m, &m->pre, cap));
  CK(dalloc(m, &m->roff, cap));
  CK(dalloc(m, &m->bsum, nb));
  CK(dalloc(m, &m->boff, nb));
  return VDBGPU_OK;
}

int ensure_chunks(vdbgpu_map* m, uint64_t need) {
  if (need <= m->chunk_cap) return VDBGPU_OK;
  dfree(m, &m->first_ray, m->chunk_cap);
  m->chunk_cap = std::max<uint64_t>(need, m->chunk_cap + m->chunk_cap / 2);
  CK(dalloc(m, &m->first_ray, m->chunk_cap));
  return VDBGPU_OK;
}

int ensure_table(vdbgpu_map* m, uint32_t need, hipStream_t stream) {
  need = std::min(need, kMaxTableN);
  if (need <= m->tbl_n) return VDBGPU_OK;
  const uint32_t n = std::min(kMaxTableN, (need + 255u) & ~255u);
  const uint64_t old_entries = (uint64_t)m->tbl_n * (m->tbl_n + 1) / 2;
  dfree(m, &m->tbl, old_entries);
  CK(dalloc(m, &m->tbl, (uint64_t)n * (n + 1) / 2));
  CK(launch_build_table(stream, m->tbl, 1, n));
  m->tbl_n = n;
  return VDBGPU_OK;
}

void ga_release(vdbgpu_map* m, GrowArray& g) {
  size_t off = 0;
  for (auto& c : g.chunks) {
    (void)hipMemUnmap(g.base + off, c.second);
    (void)hipMemRelease(c.first);
    off += c.second;
    m->device_bytes -= c.second;
  }
  g.chunks.clear();
  if (g.base) (void)hipMemAddressFree(g.base, g.reserved);
  g = GrowArray{};
}

// A refused call whose failure is handled (a fallback follows) leaves its error
// as the thread's last error, which the next launch check would report: it is
// cleared here.
bool refused() {
  (void)hipGetLastError();
  return false;
}

// Map physical memory behind g up to at least `bytes` (a multiple of gran),
// zero-filled on the map stream. Access is granted over the whole mapped
// range (the driver refuses some ranges that start inside a reservation).
// False, with the new chunk released, when the driver refuses.
bool ga_grow(vdbgpu_map* m, GrowArray& g, size_t bytes, size_t gran) {
  bytes = (bytes + gran - 1) / gran * gran;
  if (bytes <= g.mapped) return true;
  if (bytes > g.reserved) return false;
  const size_t sz = bytes - g.mapped;
  hipMemAllocationProp prop{};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = m->device;
  hipMemGenericAllocationHandle_t h;
  if (hipMemCreate(&h, sz, &prop, 0) != hipSuccess) return refused();
  if (hipMemMap(g.base + g.mapped, sz, 0, h, 0) != hipSuccess) {
    (void)hipMemRelease(h);
    return refused();
  }
  hipMemAccessDesc acc{};
  acc.location = prop.location;
  acc.flags = hipMemAccessFlagsProtReadWrite;
  if (hipMemSetAccess(g.base, bytes, &acc, 1) != hipSuccess ||
      hipMemsetAsync(g.base + g.mapped, 0, sz, m->stream) != hipSuccess) {
    (void)hipMemUnmap(g.base + g.mapped, sz);
    (void)hipMemRelease(h);
    return refused();
  }
  g.chunks.push_back({h, sz});
  g.mapped = bytes;
  m->device_bytes += sz;
  return true;
}

// Reserve the four pool ranges for as many leaves as the device's memory
// could hold; vmm = 1 on success, -1 (plain allocations) otherwise.
void vmm_init(vdbgpu_map* m) {
  m->vmm = -1;
  const char* plain = std::getenv("VDBGPU_PLAIN_POOL");  // "1": plain allocations (the fallback)
  if (plain && plain[0] == '1') return;
  int ok = 0;
  if (hipDeviceGetAttribute(&ok, hipDeviceAttributeVirtualMemoryManagementSupported, m->device) != hipSuccess || !ok) {
    refused();
    return;
  }
  hipMemAllocationProp prop{};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = m->device;
  size_t gran = 0, free_b = 0, total_b = 0;
  if (hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum) != hipSuccess || !gran ||
      hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
    refused();
    return;
  }
  const uint64_t leaves = std::min<uint64_t>(0xF0000000ull, total_b / 2048 + 1);
  GrowArray* arrs[4] = {&m->g_values, &m->g_active, &m->g_keys, &m->g_dirty};
  const size_t elem[4] = {2048, 64, 8, 1};
  for (int i = 0; i < 4; ++i) {
    const size_t bytes = (leaves * elem[i] + gran - 1) / gran * gran;
    void* p = nullptr;
    if (hipMemAddressReserve(&p, bytes, 0, nullptr, 0) != hipSuccess) {
      refused();
      for (int j = 0; j < i; ++j) ga_release(m, *arrs[j]);
      return;
    }
    arrs[i]->base = static_cast<char*>(p);
    arrs[i]->reserved = bytes;
  }
  m->vmm_gran = gran;
  m->vmm_leaves = leaves;
  m->vmm = 1;
}

// Leaf pool and leaf hash large enough for `need` leaves (hash load <= 0.5).
int ensure_map(vdbgpu_map* m, uint64_t need) {
  if (need > 0xF0000000ull) {
    m->last_error = "leaf pool would exceed 2^32 leaves";
    return VDBGPU_DEVICE_ERROR;
  }
  if (need > m->lcap) {
    const uint32_t cap = (uint32_t)std::max<uint64_t>(need, std::max<uint64_t>(4096, m->lcap + m->lcap / 2));
    if (m->vmm == 0) vmm_init(m);
    const size_t g = m->vmm_gran;
    const bool in_place = m->vmm == 1 && cap <= m->vmm_leaves && ga_grow(m, m->g_values, (size_t)cap * 2048, g) &&
                          ga_grow(m, m->g_active, (size_t)cap * 64, g) && ga_grow(m, m->g_keys, (size_t)cap * 8, g) &&
                          ga_grow(m, m->g_dirty, (size_t)cap, g);
    if (in_place) {  // grows in place: the pointers stay, nothing is copied
      m->values = reinterpret_cast<float*>(m->g_values.base);
      m->active = reinterpret_cast<unsigned long long*>(m->g_active.base);
      m->leaf_keys = reinterpret_cast<unsigned long long*>(m->g_keys.base);
      m->dirty = reinterpret_cast<uint8_t*>(m->g_dirty.base);
      // every array holds at least cap leaves (each was mapped to cap or more)
      m->lcap = (uint32_t)std::min<uint64_t>(
          std::min(m->g_values.mapped / 2048, m->g_active.mapped / 64),
          std::min<uint64_t>(std::min<uint64_t>(m->g_keys.mapped / 8, m->g_dirty.mapped), 0xF0000000ull));
    } else {
      float* v = nullptr;
      unsigned long long *a = nullptr, *k = nullptr;
      uint8_t* d = nullptr;
      CK(dalloc(m, &v, (size_t)cap * 512));
      CK(dalloc(m, &a, (size_t)cap * 8));
      CK(dalloc(m, &k, cap));
      CK(dalloc(m, &d, cap));
      CK(hipMemsetAsync(v, 0, (size_t)cap * 512 * sizeof(float), m->stream));
      CK(hipMemsetAsync(a, 0, (size_t)cap * 8 * sizeof(unsigned long long), m->stream));
      CK(hipMemsetAsync(d, 0, cap, m->stream));
      if (m->lcap) {
        CK(hipMemcpyAsync(v, m->values, (size_t)m->lcap * 512 * sizeof(float), hipMemcpyDeviceToDevice, m->stream));
        CK(hipMemcpyAsync(a, m->active, (size_t)m->lcap * 8 * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                          m->stream));
        CK(hipMemcpyAsync(k, m->leaf_keys, (size_t)m->lcap * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                          m->stream));
        CK(hipMemcpyAsync(d, m->dirty, m->lcap, hipMemcpyDeviceToDevice, m->stream));
      }
      if (m->vmm == 1) {  // the reserved ranges could not grow: plain allocations from now on
        CK(hipStreamSynchronize(m->stream));
        ga_release(m, m->g_values);
        ga_release(m, m->g_active);
        ga_release(m, m->g_keys);
        ga_release(m, m->g_dirty);
        m->vmm = -1;
      } else {
        dfree(m, &m->values, (size_t)m->lcap * 512);
        dfree(m, &m->active, (size_t)m->lcap * 8);
        dfree(m, &m->leaf_keys, m->lcap);
        dfree(m, &m->dirty, m->lcap);
      }
      m->values = v;
      m->active = a;
      m->leaf_keys = k;
      m->dirty = d;
      m->lcap = cap;
    }
  }
  if (2 * need > m->mcap) {
    const uint32_t cap = (uint32_t)std::min<uint64_t>(1ull << 31, std::max<uint64_t>(8192, next_pow2(4 * need)));
    dfree(m, &m->mkeys, m->mcap);
    dfree(m, &m->midx, m->mcap);
    m->mcap = cap;
    CK(dalloc(m, &m->mkeys, cap));
    CK(dalloc(m, &m->midx, cap));
    CK(launch_fill_u64(m->stream, m->mkeys, cap, kEmpty));
    // the leaf keys of the pool, as many as the map holds when the rehash runs
    CK(launch_rehash_map(m->stream, std::min<int>(4096, (int)((std::min<uint64_t>(need, m->lcap) + 255) / 256) + 1),
                         map_view(m), 0, &m->d_st[1]));
  }
  return VDBGPU_OK;
}

void create_flags(const vdbgpu_map* m, bool* ch, bool* cm);
int activate_artificial(vdbgpu_map* m);
int clear_grid(vdbgpu_map* m);
int flush_pending(vdbgpu_map* m);
int commit_staged(vdbgpu_map* m);

int ensure_groups(vdbgpu_map* m, uint32_t cap) {
  if (cap <= m->gcap) return VDBGPU_OK;
  dfree(m, &m->gkeys, m->gcap);
  dfree(m, &m->gscans, (size_t)m->gcap * 2);
  dfree(m, &m->gneed, m->gcap);
  dfree(m, &m->glist, m->gcap);
  dfree(m, &m->gidx, m->gcap);
  m->gcap = cap;
  CK(dalloc(m, &m->gkeys, cap));
  CK(dalloc(m, &m->gscans, (size_t)cap * 2));
  CK(dalloc(m, &m->gneed, cap));
  CK(dalloc(m, &m->glist, cap));
  CK(dalloc(m, &m->gidx, cap));
  CK(launch_fill_u64(m->stream, m->gkeys, cap, kEmpty));
  CK(hipMemsetAsync(m->gscans, 0, (size_t)cap * 16, m->stream));
  CK(hipMemsetAsync(m->gneed, 0, (size_t)cap * 4, m->stream));
  return VDBGPU_OK;
}

GroupView group_view(vdbgpu_map* m) { return GroupView{m->gkeys, m->gscans, m->gneed, m->gcap - 1}; }

int fold_grid(vdbgpu_map* m) { return m->num_cu * 8; }

// Fold half of a batch: group the (leaf, scan) entries by leaf, make room in
// the map, apply every scan in order, clear the update set.
int launch_fold_stage(vdbgpu_map* m, Source& src, bool from_groups) {
  bool ch, cm;
  create_flags(m, &ch, &cm);
  Status* st = &m->d_st[0];
  if (from_groups) {
    CK(launch_group_build(m->stream, fold_grid(m), upd_view(src), src.list, group_view(m), ch, cm, st));
    CK(launch_compact(m->stream, std::min<int>(4096, (int)((m->gcap + 255) / 256)),
                      UpdView{m->gkeys, nullptr, m->gcap - 1}, m->gcap, m->glist, &st->groups, st));
  }
  CK(launch_map_check(m->stream, st, &m->d_st[1], m->lcap, m->mcap));
  CK(launch_group_insert(m->stream, fold_grid(m), group_view(m), m->glist, map_view(m), m->gidx,
                         m->noop4096, vdbgpu_map::kNoopCap - 1, st, &m->d_st[1]));
  CK(launch_fold(m->stream, fold_grid(m), upd_view(src), group_view(m), m->glist, m->gidx, map_view(m),
                 m->lo, st, &m->d_st[3]));
  CK(launch_clear_keys(m->stream, fold_grid(m), upd_view(src), src.list, st));
  return VDBGPU_OK;
}

// Casts `scans` (rays concatenated in `xyz`) into `src`'s update set: prep,
// visit-count scan, ray offsets, traversal, compaction. A batch (batched =
// true, keys carry the scan index) continues with the scan-ordered fold into
// the map. One status read at the end; an update-set overflow grows the set
// and re-runs the traversal (marks are idempotent), a full map grows the leaf
// pool and re-runs the fold (nothing has been applied before either check).
// Stage the points and the scan table, run the per-point prep (rays, end
// voxels, visit counts and their block prefix, st[0].total). Returns the
// visit bound of the scans in *visits_bound.
int stage_scans(vdbgpu_map* m, hipStream_t stream, const float* xyz, int64_t stride, bool on_device,
                const std::vector<ScanInfo>& scans, uint32_t n_total, bool batched,
                uint64_t* visits_bound, bool clear_bricks = false) {
  int r;
  if ((r = ensure_rays(m, n_total))) return r;
  uint32_t ext_max = 1;
  uint64_t vb = 0;
  for (const ScanInfo& sc : scans) {
    const double ext = std::ceil(sc.R * m->inv) + 2.0;
    const uint32_t e = ext < (double)kMaxTableN ? (uint32_t)ext : kMaxTableN;
    ext_max = std::max(ext_max, e);
    vb += (uint64_t)sc.nrays * (1 + 3ull * e);  // visits per ray <= 1 + 3 * extent
  }
  *visits_bound = vb;
  if ((r = ensure_table(m, ext_max, stream))) return r;
  const float* pts = xyz;
  if (!on_device) {
    const size_t bytes = (size_t)(n_total - 1) * stride + 12;
    if (bytes > m->points_bytes) {
      dfree(m, &m->d_points, m->points_bytes / 4);
      const size_t cap = std::max(bytes, m->points_bytes * 3 / 2) + 16;
      CK(dalloc(m, &m->d_points, (cap + 3) / 4));
      m->points_bytes = ((cap + 3) / 4) * 4;
    }
    CK(hipMemcpyAsync(m->d_points, xyz, bytes, hipMemcpyHostToDevice, stream));
    pts = m->d_points;
  }
  const uint32_t nb = (n_total + kPrepBlock - 1) / kPrepBlock;
  const BatchCtl ctl{n_total, nb, (uint32_t)scans.size(), batched ? 1u : 0u};
  if (scans.size() == 1) {
    // one launch: scan record and sizes by value, status zeroed, brick table cleared
    CK(launch_stage_scan(stream, scans[0], ctl, m->d_scans, m->d_ctl, &m->d_st[0],
                         clear_bricks ? m->bkeys : nullptr,
                         clear_bricks ? reinterpret_cast<unsigned long long*>(m->bcnt) : nullptr,
                         clear_bricks ? m->bcap : 0));
  } else {
    std::copy(scans.begin(), scans.end(), m->h_scans);
    *m->h_ctl = ctl;
    CK(hipMemcpyAsync(m->d_scans, m->h_scans, scans.size() * sizeof(ScanInfo), hipMemcpyHostToDevice, stream));
    CK(hipMemcpyAsync(m->d_ctl, m->h_ctl, sizeof(BatchCtl), hipMemcpyHostToDevice, stream));
    CK(hipMemsetAsync(&m->d_st[0], 0, sizeof(Status), stream));
    if (clear_bricks) {
      CK(launch_fill_u64(stream, m->bkeys, m->bcap, kEmpty));
      CK(launch_fill_u64(stream, reinterpret_cast<unsigned long long*>(m->bcnt), m->bcap, 0));
    }
  }
  const bool vec4 = stride == 16 && (reinterpret_cast<uintptr_t>(pts) & 15) == 0;
  Timer t(m, kPrep, stream);
  CK(launch_prep(stream, (int)nb, pts, stride, vec4, m->d_scans, m->d_ctl, m->res, m->inv, m->tbl_n,
                 m->rayinfo, m->pre, m->bsum, &m->d_st[0]));
  CK(launch_scan_blocks(stream, m->bsum, m->boff, m->d_ctl, &m->d_st[0]));
  return VDBGPU_OK;
}

// Status code of the kernel error bits of a scan (0 = none).
int scan_error(vdbgpu_map* m, const Status& st, bool batched) {
  if (st.error & kErrRange)
    return fail(m, VDBGPU_OUT_OF_RANGE, batched ? "a voxel index is outside [-2^21, 2^21) (batch) or not finite; batch rejected"
                                                : "a voxel index is outside [-2^23, 2^23) or not finite; scan rejected");
  if (st.error & kErrTable) return fail(m, VDBGPU_UNSUPPORTED, "a ray spans more than 8192 voxels along one axis");
  if (st.error & kErrEntry) return fail(m, VDBGPU_DEVICE_ERROR, "mid-ray DDA entry did not converge");
  if (st.error & kErrBrick) return fail(m, VDBGPU_DEVICE_ERROR, "brick walk inconsistency (a segment left its brick)");
  if (st.error & kErrWalk) return fail(m, VDBGPU_DEVICE_ERROR, "fast-mode walk exceeded its step bound");
  return VDBGPU_OK;
}

// Casts `scans` (rays concatenated in `xyz`) into `src`'s update set: prep,
// visit-count scan, ray offsets, traversal, compaction. A batch (batched =
// true, keys carry the scan index) continues with the scan-ordered fold into
// the map. One status read at the end; an update-set overflow grows the set
// and re-runs the traversal (marks are idempotent), a full map grows the leaf
// pool and re-runs the fold (nothing has been applied before either check).
int run_pipeline(vdbgpu_map* m, Source& src, const float* xyz, int64_t stride, bool on_device,
                 const std::vector<ScanInfo>& scans, uint32_t n_total, bool batched) {
  int r;
  uint64_t visits_bound = 0;
  if ((r = stage_scans(m, m->stream, xyz, stride, on_device, scans, n_total, batched, &visits_bound))) return r;
  const uint32_t chunk = (uint32_t)m->tuning.steps_per_lane;
  const uint64_t max_chunks = (visits_bound + chunk - 1) / chunk + 1;
  if (max_chunks > (1ull << 29))
    return fail(m, VDBGPU_BAD_ARGUMENT, "steps_per_lane too small for this scan (chunk table > 2 GiB)");
  if ((r = ensure_chunks(m, max_chunks))) return r;
  if (src.cap == 0) {
    if ((r = alloc_update(m, src, (uint32_t)std::max<uint64_t>(kMinUpdateCap, next_pow2(4 * (uint64_t)n_total))))) return r;
  }
  if (batched && (r = ensure_groups(m, src.cap))) return r;
  const uint32_t nb = (n_total + kPrepBlock - 1) / kPrepBlock;
  {
    Timer t(m, kPrep);
    CK(launch_ray_offsets(m->stream, (int)nb, m->rayinfo, m->pre, m->boff, m->d_ctl, chunk, m->roff,
                          m->first_ray, m->chunk_cap, &m->d_st[0]));
  }
  const int block = m->tuning.block;
  bool fold_only = false;
  for (int attempt = 0;; ++attempt) {
    if (!fold_only) {
      {
        Timer t(m, kTraverse);
        CK(launch_traverse(m->stream, traverse_grid(m, block), block, m->rayinfo, m->roff, m->first_ray,
                           m->d_scans, m->d_ctl, m->tbl, upd_view(src), &m->d_st[0], chunk));
      }
      {
        Timer t(m, kCompact);
        CK(launch_compact(m->stream, std::min<int>(4096, (int)((src.cap + 255) / 256)), upd_view(src),
                          src.cap, src.list, &m->d_st[0].entries, &m->d_st[0]));
      }
    }
    if (batched) {
      Timer t(m, kApply);
      if ((r = launch_fold_stage(m, src, !fold_only))) return r;
    }
    if ((r = read_status(m))) return r;
    const Status& st = m->h_st[0];
    if ((r = scan_error(m, st, batched))) return r;
    if (attempt > 8) return fail(m, VDBGPU_DEVICE_ERROR, "update set or map cannot grow");
    if (st.overflow & kOvfHard) {
      if (src.cap >= (1u << 30)) return fail(m, VDBGPU_DEVICE_ERROR, "update set cannot grow");
      if ((r = grow_update(m, src, src.cap * 4))) return r;
      if (batched && (r = ensure_groups(m, src.cap))) return r;
      CK(hipMemsetAsync(&m->d_st[0].overflow, 0, sizeof(unsigned int), m->stream));
      CK(hipMemsetAsync(&m->d_st[0].entries, 0, sizeof(unsigned int), m->stream));
      CK(hipMemsetAsync(&m->d_st[0].groups, 0, sizeof(unsigned int), m->stream));
      fold_only = false;
      continue;
    }
    if (batched && st.map_full) {
      if ((r = ensure_map(m, (uint64_t)m->h_st[1].leaf_count + st.groups))) return r;
      CK(hipMemsetAsync(&m->d_st[0].map_full, 0, sizeof(unsigned int), m->stream));
      fold_only = true;
      continue;
    }
    break;
  }
  const Status& st = m->h_st[0];
  for (const ScanInfo& sc : scans) m->stats.scans += (sc.nrays > 0 && sc.R > 0) ? 1 : 0;
  m->stats.rays += st.rays;
  m->stats.visits += st.total;
  if (batched) {
    m->stats.update_leaves += st.entries;
    m->touched = m->touched || st.entries > 0;
    m->snap_dirty = m->snap_dirty || st.entries > 0;
    src.pending = 0;
    // the scans' updateMap calls each end by activating the artificial voxels
    // (VM:785-789); values do not depend on active states, so once after the
    // fold gives the same map
    if (st.entries > 0 && (r = activate_artificial(m))) return r;
  } else {
    src.pending = st.entries;
  }
  if ((double)st.entries > m->tuning.load_factor * src.cap && src.cap < (1u << 30)) {
    const uint32_t entries = st.entries;
    if ((r = grow_update(m, src, src.cap * 2))) return r;
    if (batched && (r = ensure_groups(m, src.cap))) return r;
    if (!batched) {
      CK(hipMemsetAsync(&m->d_st[0], 0, sizeof(Status), m->stream));
      CK(launch_compact(m->stream, std::min<int>(4096, (int)((src.cap + 255) / 256)), upd_view(src),
                        src.cap, src.list, &m->d_st[0].entries, &m->d_st[0]));
      if ((r = read_status(m))) return r;
      src.pending = m->h_st[0].entries;
      (void)entries;
    }
  }
  return VDBGPU_OK;
}

// raycastPointCloud in fast mode (VDBMapping.hpp:520-527, castRayIntoGridFast
// :577-602) into the source's update set: k_fast_cast marks the map-active
// voxels along each ray's runs of existing leaves (topology snapshot of the
// last fast-mode integrate) plus the endpoint hits; nothing is cast while the
// map is empty. An update-set overflow grows the set and re-runs (marks are
// idempotent).
int run_fast(vdbgpu_map* m, Source& src, const float* xyz, int64_t n, int64_t stride,
             const double origin[3], bool on_device) {
  const bool cast = m->touched;  // m_vdb_grid->empty() at the start of the raycast
  if (cast && !(src.has_isect && m->has_isect))
    return fail(m, VDBGPU_NO_INTERSECTOR,
                "fast mode: this source has no volume ray intersector yet (integrate in fast mode first)");
  ScanInfo sc{};
  for (int a = 0; a < 3; ++a) sc.o[a] = origin[a];
  sc.R = src.max_range;
  sc.ray0 = 0;
  sc.nrays = (uint32_t)n;
  int r;
  uint64_t vb = 0;
  if ((r = stage_scans(m, m->stream, xyz, stride, on_device, {sc}, (uint32_t)n, false, &vb))) return r;
  if (src.cap == 0) {
    if ((r = alloc_update(m, src, (uint32_t)std::max<uint64_t>(kMinUpdateCap, next_pow2(4 * (uint64_t)n))))) return r;
  }
  const int grid = std::min<int>(m->num_cu * 8, (int)((n + 255) / 256));
  for (int attempt = 0;; ++attempt) {
    {
      Timer t(m, kTraverse);
      CK(launch_fast_cast(m->stream, grid, m->rayinfo, m->d_scans, m->d_ctl, m->rt, map_view(m),
                          upd_view(src), cast, &m->d_st[0]));
    }
    {
      Timer t(m, kCompact);
      CK(launch_compact(m->stream, std::min<int>(4096, (int)((src.cap + 255) / 256)), upd_view(src),
                        src.cap, src.list, &m->d_st[0].entries, &m->d_st[0]));
    }
    if ((r = read_status(m))) return r;
    const Status& st = m->h_st[0];
    if ((r = scan_error(m, st, false))) return r;
    if (!(st.overflow & kOvfHard)) break;
    if (attempt > 8 || src.cap >= (1u << 30)) return fail(m, VDBGPU_DEVICE_ERROR, "update set cannot grow");
    if ((r = grow_update(m, src, src.cap * 4))) return r;
    CK(hipMemsetAsync(&m->d_st[0].overflow, 0, sizeof(unsigned int), m->stream));
    CK(hipMemsetAsync(&m->d_st[0].entries, 0, sizeof(unsigned int), m->stream));
  }
  const Status& st = m->h_st[0];
  m->stats.scans += 1;
  m->stats.rays += st.rays;
  src.pending = st.entries;
  return VDBGPU_OK;
}

int do_accumulate(vdbgpu_map* m, Source& src, const float* xyz, int64_t n, int64_t stride,
                  const double origin[3], bool on_device) {
  if (!m->config_set) return fail(m, VDBGPU_NOT_CONFIGURED, "Map not properly configured. Did you call setConfig method?");
  if (n < 0 || stride < 12 || (stride & 3) || n > (int64_t)kMaxRays || !origin || (n && !xyz))
    return fail(m, VDBGPU_BAD_ARGUMENT, "bad point buffer arguments");
  if (n == 0) return VDBGPU_OK;
  if (m->fast_mode) return run_fast(m, src, xyz, n, stride, origin, on_device);
  ScanInfo sc{};
  sc.o[0] = origin[0];
  sc.o[1] = origin[1];
  sc.o[2] = origin[2];
  sc.R = src.max_range;
  sc.ray0 = 0;
  sc.nrays = (uint32_t)n;
  return run_pipeline(m, src, xyz, stride, on_device, {sc}, (uint32_t)n, false);
}

// Brick path tables large enough for the given counts (grown, never shrunk).
int ensure_bricks(vdbgpu_map* m, uint32_t bcap, uint32_t seg_cap, uint32_t split_cap, uint32_t rec_cap) {
  if (bcap > m->bcap) {
    dfree(m, &m->bkeys, m->bcap);
    dfree(m, &m->bcnt, (size_t)m->bcap * 2);
    dfree(m, &m->bbsum, m->bcap / 256);
    dfree(m, &m->bboff, m->bcap / 256);
    m->bcap = bcap;
    CK(dalloc(m, &m->bkeys, bcap));
    CK(dalloc(m, &m->bcnt, (size_t)bcap * 2));
    CK(dalloc(m, &m->bbsum, bcap / 256));
    CK(dalloc(m, &m->bboff, bcap / 256));
  }
  if (seg_cap > m->seg_cap) {
    dfree(m, &m->segs, m->seg_cap);
    m->seg_cap = seg_cap;
    CK(dalloc(m, &m->segs, seg_cap));
  }
  // work items: at most one partial item per brick plus seg_cap / kItemSegs full ones
  const uint32_t item_need = m->bcap + m->seg_cap / kItemSegs + 1;
  if (item_need > m->item_cap) {
    dfree(m, &m->items, m->item_cap);
    m->item_cap = item_need;
    CK(dalloc(m, &m->items, item_need));
  }
  if (split_cap > m->split_cap) {
    dfree(m, &m->split_slot, m->split_cap);
    dfree(m, &m->shared_bits, (size_t)m->split_cap * 1024);
    m->split_cap = split_cap;
    CK(dalloc(m, &m->split_slot, split_cap));
    CK(dalloc(m, &m->shared_bits, (size_t)split_cap * 1024));
    CK(launch_fill_u64(m->cast, m->shared_bits, (uint64_t)split_cap * 1024, 0));
  }
  if (rec_cap > m->rec_cap) {  // positional records: read by the cast stream only
    dfree(m, &m->rkeys, m->rec_cap);
    dfree(m, &m->rmasks, (size_t)m->rec_cap * 16);
    dfree(m, &m->rlist, m->rec_cap);
    m->rec_cap = rec_cap;
    CK(dalloc(m, &m->rkeys, rec_cap));
    CK(dalloc(m, &m->rmasks, (size_t)rec_cap * 16));
    CK(dalloc(m, &m->rlist, rec_cap));
  }
  const uint32_t groups = (m->item_cap + m->split_cap + 255) & ~255u;
  if (groups > m->group_cap) {
    dfree(m, &m->leaf_mask, m->group_cap);
    dfree(m, &m->lbsum, m->group_cap / 256);
    dfree(m, &m->lboff, m->group_cap / 256);
    m->group_cap = groups;
    CK(dalloc(m, &m->leaf_mask, groups));
    CK(dalloc(m, &m->lbsum, groups / 256));
    CK(dalloc(m, &m->lboff, groups / 256));
  }
  return VDBGPU_OK;
}

uint32_t grow_cap(uint64_t need, uint32_t cur) {
  const uint64_t want = std::max<uint64_t>(need + need / 4, (uint64_t)cur * 2);
  return (uint32_t)std::min<uint64_t>(want, 0xC0000000ull);
}

void free_set(vdbgpu_map* m, int X) {
  dfree(m, &m->dkeys[X], m->dcap[X]);
  dfree(m, &m->dmasks[X], (size_t)m->dcap[X] * 16);
  dfree(m, &m->reidx[X], m->dcap[X]);
  m->dcap[X] = 0;
}

// Room for `cap` records in record set X (which no pending scan holds).
int grow_set(vdbgpu_map* m, int X, uint32_t cap) {
  m->dense_hint = std::max(m->dense_hint, cap);
  free_set(m, X);
  CK(dalloc(m, &m->dkeys[X], cap));
  CK(dalloc(m, &m->dmasks[X], (size_t)cap * 16));
  CK(dalloc(m, &m->reidx[X], cap));
  m->dcap[X] = cap;
  return VDBGPU_OK;
}

// Group pass of the deferred scans (insert order t = 0 .. npend-1): group
// their records by map leaf, apply every scan to each leaf in one read and
// write (k_apply_groups), activate the artificial voxels once (values do not
// depend on active states, so once after the last scan equals once after each)
// and free the record sets.
int flush_pending(vdbgpu_map* m) {
  if (m->staged) {
    const int r = commit_staged(m);
    if (r) return r;
  }
  const int P = m->npend;
  if (P == 0) return VDBGPU_OK;
  if (P == 1) {  // one scan: the per-scan apply (no grouping to pay for)
    const int X = m->pend_set[0];
    const uint32_t recs = m->pend_recs[0];
    const UpdView rec{m->dkeys[X], m->dmasks[X], 0};
    {
      Timer t(m, kApply);
      // a read or sync waits for this apply and no cast runs beside it: more
      // workgroups than the CUs hold at once (four leaves per wave, four waves
      // per SIMD), so the last ones fill in behind the slow ones (synchronous
      // p50 2.63 ms at 16 per CU against 2.65 at 8 and 4)
      CK(launch_apply(m->stream, std::min<int>(m->num_cu * 16, (int)((recs + 1) / 2)), rec, m->ident, m->reidx[X],
                      recs, map_view(m), m->lo, &m->d_st[3], false));
    }
    m->npend = 0;
    int r;
    if ((r = check_tiles(m)) || (r = activate_artificial(m))) return r;
    CK(hipEventRecord(m->rec_free[X], m->stream));
    m->rec_seq[X] = m->ledger_seq;
    return VDBGPU_OK;
  }
  uint64_t total = 0;
  uint32_t stride = 0;
  for (int p = 0; p < P; ++p) {
    total += m->pend_recs[p];
    stride = std::max(stride, m->pend_recs[p]);
  }
  const bool grow_slots = m->gslot_cap < m->lcap, grow_rgrp = (uint64_t)kDeferMax * stride > m->rgrp_cap,
             grow_grp = total > m->grp_cap;
  if (grow_slots || grow_rgrp || grow_grp) CK(hipStreamSynchronize(m->stream));  // earlier passes done
  if (grow_slots) {
    dfree(m, &m->gslot, m->gslot_cap);
    m->gslot_cap = m->lcap;
    CK(dalloc(m, &m->gslot, m->gslot_cap));
    CK(hipMemsetAsync(m->gslot, 0xFF, (size_t)m->gslot_cap * 4, m->stream));  // kNone
  }
  if (grow_rgrp) {
    dfree(m, &m->rgrp, m->rgrp_cap);
    m->rgrp_cap = (uint64_t)kDeferMax * next_pow2(stride);
    CK(dalloc(m, &m->rgrp, m->rgrp_cap));
  }
  if (grow_grp) {
    dfree(m, &m->gl, m->grp_cap);
    dfree(m, &m->grec, m->grp_cap * kDeferMax);
    m->grp_cap = std::max<uint64_t>(total, m->grp_cap * 2);
    CK(dalloc(m, &m->gl, m->grp_cap));
    CK(dalloc(m, &m->grec, m->grp_cap * kDeferMax));
  }
  if (!m->ngroups) CK(dalloc(m, &m->ngroups, 1));
  const uint32_t rs = (uint32_t)(m->rgrp_cap / kDeferMax);
  CK(hipMemsetAsync(m->ngroups, 0, sizeof(unsigned int), m->stream));
  GroupScans gs{};
  gs.nscan = P;
  {
    Timer t(m, kCompact);
    for (int p = 0; p < P; ++p) {  // in scan order: a scan's records join the groups earlier ones opened
      const int X = m->pend_set[p];
      CK(launch_group_add(m->stream, m->reidx[X], m->ident, m->pend_recs[p], (uint32_t)p, m->gslot, m->rgrp, rs,
                          m->gl, m->grec, m->ngroups));
      gs.masks[p] = m->dmasks[X];
    }
  }
  {
    Timer t(m, kApply);
    // six workgroups per CU (one wave per leaf, 64 VGPRs): on the outdoor
    // stream 2, 3, 4 and 8 per CU were slower (166.9-168.3, 169.9-170.5,
    // 167.7-170.3, 169.3-170.1 M rays/s against 170.8-171.0), the pass beside
    // the casts 7.5, 5.4, 4.2 and 4.0 ms against 3.3-3.5
    CK(launch_apply_groups(m->stream, m->num_cu * 6, gs, m->gl, m->grec, m->ngroups, m->gslot, map_view(m), m->lo,
                           &m->d_st[3]));
  }
  m->npend = 0;
  int r;
  if ((r = activate_artificial(m))) return r;
  for (int p = 0; p < P; ++p) {
    const int X = m->pend_set[p];
    CK(hipEventRecord(m->rec_free[X], m->stream));
    m->rec_seq[X] = m->ledger_seq;
  }
  return VDBGPU_OK;
}

// A brick-path scan whose cast is enqueued on the cast stream: what a re-run
// needs, its record set and the host slot its status is copied to.
struct BrickCast {
  const float* xyz = nullptr;
  int64_t n = 0, stride = 0;
  bool on_device = false;
  ScanInfo sc{};
  int X = 0;     // record set
  int slot = 0;  // host status slot: m->h_cst[2 slot] (scan), [2 slot + 1] (map snapshot)
  uint32_t cap = 0;    // capacity of set X when the records were copied into it
  uint64_t seq = 0;    // m->cast_seq of its latest enqueue
  uint64_t epoch = 0;  // m->cast_epoch then
};

// Tables of the brick path sized for a scan of n rays of range R (the first
// scan: from its visit bound; the per-block segment tables: for its blocks).
int brick_tables(vdbgpu_map* m, int64_t n, double R) {
  int r;
  if (m->bcap == 0) {  // first brick-path scan: tables sized from the scan's visit bound
    const double ext = std::min<double>(std::ceil(R * m->inv) + 2.0, (double)kMaxTableN);
    const uint64_t vb = (uint64_t)n * (1 + 3 * (uint64_t)ext);
    const uint64_t segs0 = std::min<uint64_t>(vb, 16ull * (uint64_t)n);
    if ((r = ensure_bricks(m, (uint32_t)std::max<uint64_t>(4096, next_pow2((uint64_t)n / 2)),
                           (uint32_t)std::max<uint64_t>(1u << 16, std::min<uint64_t>(segs0, 0xC0000000ull)), 256,
                           (uint32_t)std::max<uint64_t>(1u << 16, std::min<uint64_t>(8ull * n, 0xC0000000ull)))))
      return r;
  }
  const int nb = (int)((n + kPrepBlock - 1) / kPrepBlock);
  if ((uint64_t)nb * kSegTab > m->wtab_cap) {  // (hipFree waits for the kernels using them)
    dfree(m, &m->wkey, m->wtab_cap);
    dfree(m, &m->wsc, m->wtab_cap);
    // room for 262,144 rays at least, then powers of two: few growths
    m->wtab_cap = std::max<uint64_t>(1024, next_pow2((uint64_t)nb)) * kSegTab;
    CK(dalloc(m, &m->wkey, m->wtab_cap));
    CK(dalloc(m, &m->wsc, m->wtab_cap));
  }
  return VDBGPU_OK;
}

// The brick stages of a staged scan on the cast stream: cut, walk, record
// list and the copy into the record set named in m->d_dest. `nb` = workgroups
// of the cut (lanes past the scan's rays do nothing).
int cast_chain(vdbgpu_map* m, int nb) {
  hipStream_t cs = m->cast;
  const BrickView bv{m->bkeys, m->bcnt, m->bcnt + m->bcap, m->bcap - 1};
  const UpdView rec{m->rkeys, m->rmasks, 0};
  {
    Timer t(m, kTraverse, cs);
    CK(launch_seg_count(cs, nb, m->rayinfo, m->d_scans, m->d_ctl, m->tbl, bv, m->wkey, m->wsc, &m->d_st[0]));
    CK(launch_brick_offsets(cs, bv, m->bbsum, m->bboff, m->items, m->seg_cap, m->item_cap,
                            m->split_slot, m->split_cap, &m->d_st[0]));
    CK(launch_seg_fill(cs, nb, m->rayinfo, m->d_scans, m->d_ctl, m->tbl, bv, m->wkey, m->wsc, m->segs,
                       m->seg_cap, &m->d_st[0]));
  }
  {
    Timer t(m, kWalk, cs);
    // 64 work-groups per CU over the grid-stride item loop: bricks cost very
    // different amounts, and the finer split shortens the tail.
    CK(launch_brick_walk(cs, m->num_cu * 64, m->items, m->segs, m->rayinfo, m->d_scans, m->tbl, bv,
                         m->shared_bits, rec, m->leaf_mask, m->rec_cap, &m->d_st[0]));
  }
  {
    Timer t(m, kCompact, cs);
    CK(launch_split_emit(cs, m->num_cu, m->split_slot, m->shared_bits, bv, rec, m->leaf_mask,
                         m->rec_cap, &m->d_st[0]));
    CK(launch_rec_list(cs, m->leaf_mask, m->lbsum, m->lboff, m->group_cap / 256, m->rlist, &m->d_st[0]));
    CK(launch_rec_gather(cs, m->num_cu * 8, m->rlist, rec, m->d_dest, &m->d_st[0]));
  }
  return VDBGPU_OK;
}

void drop_cast_graphs(vdbgpu_map* m) {
  for (auto& g : m->cast_graphs) (void)hipGraphExecDestroy(g.exec);
  m->cast_graphs.clear();
}

// cast_chain replayed from a captured HIP graph: one instantiated graph per
// exact set of its launch arguments (every pointer and size the chain passes),
// so a graph is only replayed with the buffers it was captured with, and a
// grown table simply yields another graph. The record set comes from device
// memory and the cut's grid is rounded up to a power of two, so that every
// set of the ring and scans of different sizes share graphs.
int launch_cast_graph(vdbgpu_map* m, int nb) {
  hipStream_t cs = m->cast;
  const int grid = (int)next_pow2((uint64_t)nb);
  auto u = [](const void* p) { return (uint64_t)(uintptr_t)p; };
  const std::vector<uint64_t> key = {
      (uint64_t)grid, (uint64_t)m->num_cu, u(m->rayinfo), u(m->d_scans), u(m->d_ctl), u(m->tbl), u(m->bkeys),
      u(m->bcnt), m->bcap, u(m->wkey), u(m->wsc), m->wtab_cap, u(m->segs), m->seg_cap, u(m->bbsum), u(m->bboff),
      u(m->items), m->item_cap, u(m->split_slot), m->split_cap, u(m->shared_bits), u(m->rkeys), u(m->rmasks),
      u(m->leaf_mask), m->rec_cap, u(m->lbsum), u(m->lboff), m->group_cap, u(m->rlist), u(m->d_dest),
      u(m->d_st)};
  vdbgpu_map::CastGraph* hit = nullptr;
  for (auto& g : m->cast_graphs)
    if (g.key == key) hit = &g;
  if (!hit) {
    if (m->cast_graphs.size() >= vdbgpu_map::kCastGraphs) {  // drop the least recently used
      size_t lru = 0;
      for (size_t i = 1; i < m->cast_graphs.size(); ++i)
        if (m->cast_graphs[i].used < m->cast_graphs[lru].used) lru = i;
      CK(hipStreamSynchronize(cs));  // (no launch of it is still running)
      (void)hipGraphExecDestroy(m->cast_graphs[lru].exec);
      m->cast_graphs.erase(m->cast_graphs.begin() + (long)lru);
    }
    CK(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
    const int rc = cast_chain(m, grid);
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(cs, &graph);
    if (rc || ec != hipSuccess || !graph) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      CK(ec);
      return fail(m, VDBGPU_DEVICE_ERROR, "stream capture of a brick cast returned no graph");
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    CK(ei);
    m->cast_graphs.push_back({key, exec, 0});
    hit = &m->cast_graphs.back();
  }
  hit->used = ++m->cast_graph_clock;
  CK(hipGraphLaunch(hit->exec, cs));
  return VDBGPU_OK;
}

// Enqueues the cast of c on the cast stream: staging (prep, brick table
// cleared), record set c.X named for the copy, brick cut, walk, record list
// and the copy (launched, or replayed from a graph), then the copy of its
// status (and a snapshot of the map's leaf count) into host slot c.slot. A first attempt takes the next record set of the ring and
// waits until the map stream has applied the scan that used it last; a re-run
// (attempt > 0, tables grown) stages the scan again, since a later cast may
// have used the shared tables meanwhile.
int brick_enqueue(vdbgpu_map* m, BrickCast& c, int attempt) {
  int r;
  hipStream_t cs = m->cast;
  const int nb = (int)((c.n + kPrepBlock - 1) / kPrepBlock);
  uint64_t vb = 0;
  if ((r = stage_scans(m, cs, c.xyz, c.stride, c.on_device, {c.sc}, (uint32_t)c.n, false, &vb, true))) return r;
  if (attempt == 0) {
    // record set X: free once the map stream has applied the scan that last used it
    c.X = m->rec_cur;
    m->rec_cur = (c.X + 1) % m->rec_sets;
    c.slot = m->cst_next;
    m->cst_next ^= 1;
    CK(hipStreamWaitEvent(cs, m->rec_free[c.X], 0));
    if (m->dcap[c.X] < m->dense_hint) {  // a first use of set X: size it like the others
      CK(hipStreamSynchronize(cs));      // (the map stream is done with it)
      if ((r = grow_set(m, c.X, m->dense_hint))) return r;
      // the ring's sets never used yet are sized now as well, so the next
      // scans of the ring do not each stop the cast stream for an allocation
      for (int y = 0; y < m->rec_sets; ++y)
        if (m->dcap[y] == 0 && (r = grow_set(m, y, m->dense_hint))) return r;
    }
  }
  c.cap = m->dcap[c.X];  // (a later first use may size an empty set X before this is checked)
  CK(launch_put_dest(cs, m->d_dest, RecDest{m->dkeys[c.X], m->dmasks[c.X], c.cap, 0}));
  if (m->graph_casts && !m->profiling) {
    if ((r = launch_cast_graph(m, nb))) return r;
  } else if ((r = cast_chain(m, nb))) {
    return r;
  }
  // the scan's status and a snapshot of the leaf count: the map stream keeps
  // applying earlier scans meanwhile
  CK(hipMemcpyAsync(m->h_cst + 2 * c.slot, m->d_st, 2 * sizeof(Status), hipMemcpyDeviceToHost, cs));
  CK(hipEventRecord(m->cst_done[c.slot], cs));
  c.seq = ++m->cast_seq;
  c.epoch = m->cast_epoch;
  return VDBGPU_OK;
}

// Waits for the status of the cast of c and checks it: errors fail the scan,
// table overflows grow the tables and cast the scan again (nothing of it has
// been applied), a record set too small grows and takes the records again.
// A cast enqueued before an overflow of an earlier cast was found started
// from split-brick bitmaps holding that cast's partial ORs: it runs again.
// When another cast has run since c's, the shared tables no longer hold c's
// records, so a set too small also means casting again. A scan with records
// is left staged for the map stream (commit_staged).
int brick_check(vdbgpu_map* m, BrickCast& c) {
  int r;
  const int X = c.X;
  for (int attempt = 1;; ++attempt) {
    CK(hipEventSynchronize(m->cst_done[c.slot]));
    collect_timing(m);
    if (attempt > 10) return fail(m, VDBGPU_DEVICE_ERROR, "brick tables cannot grow");
    if (c.epoch != m->cast_epoch) {
      if ((r = brick_enqueue(m, c, attempt))) return r;
      continue;
    }
    const bool later = c.seq != m->cast_seq;
    const Status* h = m->h_cst + 2 * c.slot;
    m->leaf_count_ub = leaf_bound(m, h[1].leaf_count, m->rec_seq[X]);
    if (h[1].map_full)
      return fail(m, VDBGPU_DEVICE_ERROR, "leaf pool overflow in a map insert (internal bound violated)");
    const Status st = h[0];
    if ((r = scan_error(m, st, false))) return r;
    if (st.overflow == 0 && (st.records <= c.cap || !later)) break;
    if (st.overflow) {
      const uint32_t bcap = (st.overflow & kOvfBricks) ? (uint32_t)std::min<uint64_t>(1u << 30, (uint64_t)m->bcap * 4) : m->bcap;
      const uint32_t seg_cap = st.segs > m->seg_cap ? grow_cap(st.segs, m->seg_cap) : m->seg_cap;
      const uint32_t split_cap = st.splits > m->split_cap ? grow_cap(st.splits, m->split_cap) : m->split_cap;
      const uint64_t positions = (uint64_t)(st.items + st.splits) * 64;
      const uint32_t rec_cap = positions > m->rec_cap ? grow_cap(positions, m->rec_cap) : m->rec_cap;
      if ((st.overflow & kOvfBricks) && bcap == m->bcap) return fail(m, VDBGPU_DEVICE_ERROR, "brick table cannot grow");
      if ((r = ensure_bricks(m, bcap, seg_cap, split_cap, rec_cap))) return r;
      // split-brick bitmaps may hold partial ORs of the stopped run: cleared,
      // and any cast enqueued since runs again
      CK(launch_fill_u64(m->cast, m->shared_bits, (uint64_t)m->split_cap * 1024, 0));
      ++m->cast_epoch;
    } else if (st.records > m->dcap[X] &&
               (r = grow_set(m, X, (uint32_t)std::min<uint64_t>(st.records + (uint64_t)st.records / 4,
                                                                0xC0000000ull)))) {
      return r;
    }
    if ((r = brick_enqueue(m, c, attempt))) return r;
  }
  const Status& st = m->h_cst[2 * c.slot];
  const uint32_t recs = st.records;  // listed in m->rlist, copied to set X when it had room
  m->stats.scans += 1;
  m->stats.rays += st.rays;
  m->stats.visits += st.total;
  if (recs == 0) {
    CK(hipEventRecord(m->rec_free[X], m->stream));  // orders the next scan's snapshot as usual
    m->rec_seq[X] = m->ledger_seq;
    return VDBGPU_OK;
  }
  if (recs > c.cap) {  // set X (no pending scan holds it) grows; the copy runs again
    if (recs > m->dcap[X] &&
        (r = grow_set(m, X, (uint32_t)std::min<uint64_t>(recs + (uint64_t)recs / 4, 0xC0000000ull))))
      return r;
    c.cap = m->dcap[X];
    CK(launch_put_dest(m->cast, m->d_dest, RecDest{m->dkeys[X], m->dmasks[X], c.cap, 0}));
    CK(launch_rec_gather(m->cast, m->num_cu * 8, m->rlist, UpdView{m->rkeys, m->rmasks, 0}, m->d_dest,
                         &m->d_st[0]));
    CK(hipStreamSynchronize(m->cast));
  }
  m->staged = true;
  m->staged_set = X;
  m->staged_recs = recs;
  return VDBGPU_OK;
}

BrickCast brick_cast(const Source& src, const float* xyz, int64_t n, int64_t stride, const double origin[3],
                     bool on_device) {
  BrickCast c;
  c.xyz = xyz;
  c.n = n;
  c.stride = stride;
  c.on_device = on_device;
  for (int a = 0; a < 3; ++a) c.sc.o[a] = origin[a];
  c.sc.R = src.max_range;
  c.sc.ray0 = 0;
  c.sc.nrays = (uint32_t)n;
  return c;
}

// insertPointCloud of one scan when no source holds pending updates: the
// brick path casts the scan into leaf records and applies them at once
// (equal to accumulateUpdate + integrateUpdate). While the cast runs, the
// previous scan's map-stream work (and a full group of deferred scans) is
// enqueued; the call returns once the scan's status is checked.
int run_brick(vdbgpu_map* m, Source& src, const float* xyz, int64_t n, int64_t stride,
              const double origin[3], bool on_device) {
  int r;
  if ((r = brick_tables(m, n, src.max_range))) return r;
  BrickCast c = brick_cast(src, xyz, n, stride, origin, on_device);
  if ((r = brick_enqueue(m, c, 0))) return r;
  if ((r = commit_staged(m)) || (m->npend >= m->tuning.defer && (r = flush_pending(m)))) return r;
  return brick_check(m, c);
}

// The map-stream work of the staged scan (set X, recs records): map insert
// and leaf-count snapshot, then a place among the deferred scans, or (one
// scan at a time) the apply.
int commit_staged(vdbgpu_map* m) {
  if (!m->staged) return VDBGPU_OK;
  m->staged = false;
  const int X = m->staged_set;
  const uint32_t recs = m->staged_recs;
  int r;
  const uint32_t art_new = m->art_leaves ? 0 : m->art_count;  // leaves the activation may create
  if ((r = ensure_map(m, m->leaf_count_ub + recs + art_new))) return r;
  if (recs > m->ident_cap) {
    dfree(m, &m->ident, m->ident_cap);
    m->ident_cap = m->dcap[X];
    CK(dalloc(m, &m->ident, m->ident_cap));
    CK(launch_iota_u32(m->stream, m->ident, m->ident_cap));
  }
  bool ch, cm;
  create_flags(m, &ch, &cm);
  const UpdView rec{m->dkeys[X], m->dmasks[X], 0};
  const int g1 = std::min<int>(4096, (int)((recs + 255) / 256));
  {
    Timer t(m, kCompact);
    CK(launch_map_insert(m->stream, g1, rec, m->ident, recs, map_view(m), m->reidx[X], ch, cm, m->noop4096,
                         vdbgpu_map::kNoopCap - 1, &m->d_st[1], &m->d_st[3]));
  }
  m->leaf_count_ub += recs;
  ledger_add(m, recs);
  CK(launch_snap_count(m->stream, &m->d_st[1], &m->d_lc[X]));
  CK(hipEventRecord(m->ins_done[X], m->stream));
  m->ins_seq[X] = m->ledger_seq;
  for (auto it = m->ins_fifo.begin(); it != m->ins_fifo.end(); ++it)
    if (*it == X) {  // its earlier snapshot slot is being rewritten
      m->ins_fifo.erase(it);
      break;
    }
  m->ins_fifo.push_back(X);
  m->touched = true;
  m->snap_dirty = true;
  // a tile's state is checked after every scan, so scans with tiles are not deferred
  if (m->rec_sets > 2 && recs <= kGroupRecMask && !m->ntiles) {
    // applied with the next group pass: when defer scans are waiting (the
    // next insert enqueues it while its cast runs) or the map is read or
    // changed otherwise
    if (m->npend >= kDeferMax && (r = flush_pending(m))) return r;
    m->pend_set[m->npend] = X;
    m->pend_recs[m->npend] = recs;
    ++m->npend;
    return VDBGPU_OK;
  }
  if ((r = flush_pending(m))) return r;  // earlier scans first
  // 128 lanes per leaf; three workgroups per CU (each lane keeps four words in
  // flight) leave the CUs to the next scan's cast: fewer apply waves slow the
  // apply less than they speed up the overlapped cast (measured 3-6 per CU)
  const int g2 = std::min<int>(m->num_cu * 3, (int)((recs + 1) / 2));
  {
    Timer t(m, kApply);
    CK(launch_apply(m->stream, g2, rec, m->ident, m->reidx[X], recs, map_view(m), m->lo, &m->d_st[3], false));
  }
  if ((r = check_tiles(m))) return r;
  if ((r = activate_artificial(m))) return r;  // enters its new leaves in the ledger
  CK(hipEventRecord(m->rec_free[X], m->stream));
  m->rec_seq[X] = m->ledger_seq;
  return VDBGPU_OK;
}

// The raycast half of insertPointCloud: the brick path when it applies
// (no pending updates anywhere, tuning.traversal == 0), else the update-set
// traversal (accumulateUpdate); the caller integrates afterwards.
int insert_cast(vdbgpu_map* m, Source& src, const float* xyz, int64_t n, int64_t stride,
                const double origin[3], bool on_device) {
  bool pending = false;
  for (auto& kv : m->sources) pending = pending || kv.second.pending > 0;
  const bool args_ok = !(n < 0 || stride < 12 || (stride & 3) || n > (int64_t)kMaxRays || !origin || (n && !xyz));
  if (m->tuning.traversal == 0 && !pending && m->config_set && !m->fast_mode && args_ok && n > 0 &&
      src.max_range > 0)
    return run_brick(m, src, xyz, n, stride, origin, on_device);
  return do_accumulate(m, src, xyz, n, stride, origin, on_device);
}

// Creation rule for a missing leaf: the op must change a background voxel
// (value 0, inactive).
void create_flags(const vdbgpu_map* m, bool* ch, bool* cm) {
  float x = 0.0f + m->lo[0];
  bool a = false;
  if (x > m->lo[3]) {
    a = true;
    if (x > m->lo[4]) x = m->lo[4];
  }
  *ch = (x != 0.0f) || a;
  x = 0.0f + m->lo[1];
  if (x < m->lo[2]) {
    if (x < m->lo[5]) x = m->lo[5];
  }
  *cm = x != 0.0f;
}

// updateVolumeRayIntersectors (VDBMapping.hpp:1436-1449): in fast mode, a
// non-empty map gets a fresh topology snapshot after every integrate.
int build_snapshot(vdbgpu_map* m) {
  int r = read_status(m);
  if (r) return r;
  const uint32_t L = m->h_st[1].leaf_count;
  const uint32_t cap = (uint32_t)std::max<uint64_t>(1024, next_pow2(2ull * L + 2));
  const uint32_t cap_root = (uint32_t)std::max<uint64_t>(1024, next_pow2(2ull * L + 2 * vdbgpu_map::kNoopCap));
  if (cap > m->rt_leaf_cap) {
    dfree(m, &m->rt.leaf, m->rt_leaf_cap);
    dfree(m, &m->rt.n128, m->rt_n128_cap);
    m->rt_leaf_cap = m->rt_n128_cap = cap;
    CK(dalloc(m, &m->rt.leaf, cap));
    CK(dalloc(m, &m->rt.n128, cap));
  }
  if (cap_root > m->rt_n4096_cap) {
    dfree(m, &m->rt.n4096, m->rt_n4096_cap);
    m->rt_n4096_cap = cap_root;
    CK(dalloc(m, &m->rt.n4096, cap_root));
  }
  m->rt.leaf_mask = m->rt_leaf_cap - 1;
  m->rt.n128_mask = m->rt_n128_cap - 1;
  m->rt.n4096_mask = m->rt_n4096_cap - 1;
  CK(launch_fill_u64(m->stream, m->rt.leaf, m->rt_leaf_cap, kEmpty));
  CK(launch_fill_u64(m->stream, m->rt.n128, m->rt_n128_cap, kEmpty));
  CK(launch_fill_u64(m->stream, m->rt.n4096, m->rt_n4096_cap, kEmpty));
  if (!m->d_bbox) CK(dalloc(m, &m->d_bbox, 6));
  int hb[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  CK(hipMemcpyAsync(m->d_bbox, hb, sizeof hb, hipMemcpyHostToDevice, m->stream));
  if (L) CK(launch_snapshot(m->stream, map_view(m), L, m->rt, m->d_bbox, tile_view(m)));
  CK(launch_copy_set(m->stream, m->noop4096, vdbgpu_map::kNoopCap, m->rt.n4096, m->rt.n4096_mask));
  CK(hipMemcpyAsync(hb, m->d_bbox, sizeof hb, hipMemcpyDeviceToHost, m->stream));
  CK(hipStreamSynchronize(m->stream));
  for (int a = 0; a < 3; ++a) {
    m->rt.bmin[a] = hb[a];
    m->rt.bmax[a] = hb[3 + a] + 1;  // mBBox.max().offset(1)
  }
  m->has_isect = true;
  m->snap_dirty = false;
  for (auto& kv : m->sources) kv.second.has_isect = true;  // every source gets a copy (VDBMapping.hpp:1441-1446)
  return VDBGPU_OK;
}

// updateVolumeRayIntersectors after an integrate (VDBMapping.hpp:1436-1449):
// in fast mode with a non-empty map every source gets the intersector; the
// snapshot is rebuilt only when the map changed since the last one (an
// identical topology copy otherwise).
int refresh_snapshot(vdbgpu_map* m) {
  if (!m->fast_mode || !m->touched) return VDBGPU_OK;
  if (m->snap_dirty || !m->has_isect) return build_snapshot(m);
  for (auto& kv : m->sources) kv.second.has_isect = true;
  return VDBGPU_OK;
}

int do_integrate(vdbgpu_map* m) {
  bool ch, cm;
  create_flags(m, &ch, &cm);
  for (auto& kv : m->sources)
    if (kv.second.pending) {
      const int r = flush_pending(m);  // deferred brick-path scans come first
      if (r) return r;
      break;
    }
  bool applied = false;
  for (auto& kv : m->sources) {
    Source& s = kv.second;
    if (s.pending == 0) continue;
    int r;
    if ((r = ensure_map(m, m->leaf_count_ub + s.pending))) return r;
    if (s.pending > m->eidx_cap) {
      dfree(m, &m->eidx, m->eidx_cap);
      m->eidx_cap = (uint32_t)next_pow2(s.pending);
      CK(dalloc(m, &m->eidx, m->eidx_cap));
    }
    const int g1 = std::min<int>(4096, (int)((s.pending + 255) / 256));
    const int g2 = std::min<int>(m->num_cu * 16, (int)((s.pending + 1) / 2));  // 128 lanes per leaf
    {
      Timer t(m, kCompact);
      CK(launch_map_insert(m->stream, g1, upd_view(s), s.list, s.pending, map_view(m), m->eidx, ch, cm,
                           m->noop4096, vdbgpu_map::kNoopCap - 1, &m->d_st[1], &m->d_st[3]));
    }
    {
      Timer t(m, kApply);
      CK(launch_apply(m->stream, g2, upd_view(s), s.list, m->eidx, s.pending, map_view(m), m->lo,
                      &m->d_st[3], true));
    }
    m->leaf_count_ub += s.pending;
    ledger_add(m, s.pending);
    m->touched = true;
    m->snap_dirty = true;
    s.pending = 0;
    applied = true;
    if ((r = check_tiles(m))) return r;
    // with tiles the activation runs per source: a tile checked in between
    // may turn into a leaf that the activation would have left uniform
    if (m->ntiles && (r = activate_artificial(m))) return r;
  }
  // every non-empty updateMap ends by activating the artificial voxels
  // (VM:785-789); values do not depend on active states, so once after the
  // last source gives the same map as after each one
  if (applied && !m->ntiles) {
    const int r = activate_artificial(m);
    if (r) return r;
  }
  return refresh_snapshot(m);
}

// ---------------------------------------------------------- direct edits --

// Empty an edit set (keys EMPTY, masks 0).
int clear_set(vdbgpu_map* m, Source& s) {
  if (!s.cap) return VDBGPU_OK;
  CK(launch_fill_u64(m->stream, s.keys, s.cap, kEmpty));
  CK(hipMemsetAsync(s.masks, 0, (size_t)s.cap * 16 * sizeof(unsigned long long), m->stream));
  return VDBGPU_OK;
}

// Status of the kernel error bits of an edit call.
int edit_error(vdbgpu_map* m, const Status& st) {
  if (st.error & kErrRange)
    return fail(m, VDBGPU_OUT_OF_RANGE, "a voxel index is outside [-2^23, 2^23) or not finite; edit rejected");
  if (st.error & kErrTable) return fail(m, VDBGPU_UNSUPPORTED, "a wall spans more than 8192 voxels along one axis");
  return st.error ? fail(m, VDBGPU_DEVICE_ERROR, "edit kernel error") : VDBGPU_OK;
}

// Mark an edit set with `mark` (idempotent marks), list it, read the status
// (the call's only sync). A full set is cleared, grown 4x and marked again; a
// kernel error clears the set and returns its status. *count = listed leaves.
template <class Mark>
int mark_set(vdbgpu_map* m, Source& s, uint32_t cap0, Mark&& mark, uint32_t* count) {
  int r;
  if (s.cap < cap0) {
    free_update(m, s);
    if ((r = alloc_update(m, s, cap0))) return r;
  }
  for (int attempt = 0;; ++attempt) {
    CK(hipMemsetAsync(&m->d_st[2], 0, sizeof(Status), m->stream));
    if ((r = mark())) return r;
    CK(launch_compact(m->stream, std::min<int>(4096, (int)((s.cap + 255) / 256)), upd_view(s), s.cap, s.list,
                      &m->d_st[2].entries, &m->d_st[2]));
    if ((r = read_status(m))) return r;
    const Status st = m->h_st[2];
    if (st.error) {
      if ((r = clear_set(m, s))) return r;
      return edit_error(m, st);
    }
    if (!(st.overflow & kOvfHard)) {
      *count = st.entries;
      return VDBGPU_OK;
    }
    if (attempt > 6 || s.cap >= (1u << 30)) return fail(m, VDBGPU_DEVICE_ERROR, "edit set cannot grow");
    const uint32_t cap = s.cap * 4;
    free_update(m, s);
    if ((r = alloc_update(m, s, cap))) return r;
  }
}

// Find or create the map leaf of each listed edit-set leaf (created only when
// `need`: the op changes a background voxel; a no-op modify still gives the
// root its 4096^3 child), then run `op` on the marked voxels.
int apply_edit(vdbgpu_map* m, Source& s, uint32_t count, uint32_t* eidx, int op, float val, bool need,
               bool clear) {
  if (count == 0) return VDBGPU_OK;
  const int g1 = std::min<int>(4096, (int)((count + 255) / 256));
  CK(launch_map_insert(m->stream, g1, upd_view(s), s.list, count, map_view(m), eidx, need, need, m->noop4096,
                       vdbgpu_map::kNoopCap - 1, &m->d_st[1], &m->d_st[2]));
  CK(launch_set_voxels(m->stream, upd_view(s), s.list, eidx, count, map_view(m), op, val, m->lo[3], clear));
  m->touched = true;
  m->snap_dirty = true;
  return check_tiles(m);
}

int ensure_eidx(vdbgpu_map* m, uint32_t n) {
  if (n <= m->eidx_cap) return VDBGPU_OK;
  dfree(m, &m->eidx, m->eidx_cap);
  m->eidx_cap = (uint32_t)next_pow2(n);
  CK(dalloc(m, &m->eidx, m->eidx_cap));
  return VDBGPU_OK;
}

// Last step of a non-empty updateMap (VM:785-789): setActiveState(on) for
// every artificial voxel (a missing leaf is created with value 0). The
// artificial leaves are created once; later activations create none.
int activate_artificial(vdbgpu_map* m) {
  if (m->art_count == 0) return VDBGPU_OK;
  int r;
  const uint32_t fresh = m->art_leaves ? 0 : m->art_count;
  if (fresh && (r = ensure_map(m, m->leaf_count_ub + fresh))) return r;
  if ((r = apply_edit(m, m->art, m->art_count, m->art_eidx, kEditActivate, 0.0f, true, false))) return r;
  m->leaf_count_ub += fresh;
  ledger_add(m, fresh);
  m->art_leaves = true;
  return VDBGPU_OK;
}

// restoreMapIntegrity (VM:1152-1166): setNodeState through
// modifyValueAndActiveState on every artificial voxel, then the grid is cleared.
int restore_integrity(vdbgpu_map* m) {
  int r;
  if (m->art_count) {
    const bool need = 0.0f > m->lo[3];  // would a background voxel change?
    if (need && (r = ensure_map(m, m->leaf_count_ub + m->art_count))) return r;
    if ((r = apply_edit(m, m->art, m->art_count, m->art_eidx, kEditState, 0.0f, need, true))) return r;
    if (need) {
      m->leaf_count_ub += m->art_count;
      ledger_add(m, m->art_count);
    }
  }
  m->art_count = 0;
  m->art_present = false;
  m->art_leaves = false;
  m->art_verts.clear();
  m->art_walls.clear();
  return VDBGPU_OK;
}

// Stage `bytes` of host data in a fresh device buffer (freed by the caller).
template <class T>
int upload(vdbgpu_map* m, T** dst, const T* src, size_t count) {
  CK(dalloc(m, dst, count));
  if (count) CK(hipMemcpyAsync(*dst, src, count * sizeof(T), hipMemcpyHostToDevice, m->stream));
  return VDBGPU_OK;
}

// Validate section leaf origins: multiples of 8, inside the key range, distinct.
int check_origins(vdbgpu_map* m, int64_t n, const int32_t* origins) {
  std::vector<unsigned long long> keys((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    if ((o[0] | o[1] | o[2]) & 7) return fail(m, VDBGPU_BAD_ARGUMENT, "section leaf origin is not a multiple of 8");
    for (int a = 0; a < 3; ++a)
      if (o[a] < -kCoordLimit || o[a] >= kCoordLimit)
        return fail(m, VDBGPU_OUT_OF_RANGE, "section leaf origin outside [-2^23, 2^23)");
    keys[(size_t)i] = leaf_key(o[0], o[1], o[2]);
  }
  std::sort(keys.begin(), keys.end());
  if (std::adjacent_find(keys.begin(), keys.end()) != keys.end())
    return fail(m, VDBGPU_BAD_ARGUMENT, "section leaf origins repeat");
  return VDBGPU_OK;
}

// Lock of one ABI call. Every call but a plain insert first applies the
// deferred scans, so whatever it reads or changes sees the map of every
// earlier insert; when that fails the call returns the failure (last_error
// says what it was) and does nothing else.
// The per-scan working memory of the insert, batch, edit and query paths:
// point staging, per-ray arrays, brick tables, segments, work items, records,
// record sets, group-pass and fold tables, query staging, and the update sets
// of sources with nothing pending. Each buffer regrows on its next use, as
// after vdbgpu_create. `release` frees them (the caller has waited for both
// streams); otherwise their bytes are summed.
uint64_t walk_scratch(vdbgpu_map* m, bool release) {
  uint64_t bytes = 0;
  auto one = [&](auto*& p, size_t count) {
    if (!p) return;
    bytes += count * sizeof(*p);
    if (release) dfree(m, &p, count);
  };
  const uint32_t nb = (m->ray_cap + kPrepBlock - 1) / kPrepBlock;
  one(m->d_points, m->points_bytes / 4);
  one(m->rayinfo, m->ray_cap);
  one(m->pre, m->ray_cap);
  one(m->roff, m->ray_cap);
  one(m->bsum, nb);
  one(m->boff, nb);
  one(m->first_ray, m->chunk_cap);
  one(m->eidx, m->eidx_cap);
  one(m->bkeys, m->bcap);
  one(m->bcnt, (size_t)m->bcap * 2);
  one(m->bbsum, m->bcap / 256);
  one(m->bboff, m->bcap / 256);
  one(m->segs, m->seg_cap);
  one(m->items, m->item_cap);
  one(m->split_slot, m->split_cap);
  one(m->shared_bits, (size_t)m->split_cap * 1024);
  one(m->wkey, m->wtab_cap);
  one(m->wsc, m->wtab_cap);
  one(m->rkeys, m->rec_cap);
  one(m->rmasks, (size_t)m->rec_cap * 16);
  one(m->rlist, m->rec_cap);
  one(m->leaf_mask, m->group_cap);
  one(m->lbsum, m->group_cap / 256);
  one(m->lboff, m->group_cap / 256);
  for (int X = 0; X < vdbgpu_map::kRecSets; ++X) {
    one(m->dkeys[X], m->dcap[X]);
    one(m->dmasks[X], (size_t)m->dcap[X] * 16);
    one(m->reidx[X], m->dcap[X]);
  }
  one(m->ident, m->ident_cap);
  one(m->gslot, m->gslot_cap);
  one(m->rgrp, m->rgrp_cap);
  one(m->gl, m->grp_cap);
  one(m->grec, m->grp_cap * kDeferMax);
  one(m->gkeys, m->gcap);
  one(m->gscans, (size_t)m->gcap * 2);
  one(m->gneed, m->gcap);
  one(m->glist, m->gcap);
  one(m->gidx, m->gcap);
  one(m->q_in, (size_t)m->q_cap * 7);
  one(m->q_succ, (size_t)m->q_cap);
  one(m->q_end, (size_t)m->q_cap * 3);
  auto upd = [&](Source& u) {
    if (u.pending) return;
    one(u.keys, u.cap);
    one(u.masks, (size_t)u.cap * 16);
    one(u.list, u.cap);
    if (release) u.cap = 0;
  };
  for (auto& kv : m->sources) upd(kv.second);
  upd(m->edit);
  if (release) {
    m->points_bytes = 0;
    m->ray_cap = 0;
    m->chunk_cap = 0;
    m->eidx_cap = 0;
    m->bcap = m->seg_cap = m->item_cap = m->split_cap = m->rec_cap = m->group_cap = 0;
    m->wtab_cap = 0;
    for (int X = 0; X < vdbgpu_map::kRecSets; ++X) m->dcap[X] = 0;
    m->ident_cap = 0;
    m->gslot_cap = 0;
    m->rgrp_cap = 0;
    m->grp_cap = 0;
    m->gcap = 0;
    m->q_cap = 0;
  }
  return bytes;
}

struct Guard {
  vdbgpu_map* m;
  std::lock_guard<std::mutex> lk;
  int rc = VDBGPU_OK;  // status of applying the deferred scans: the call returns it when not OK
  explicit Guard(vdbgpu_map* mm, bool keep_deferred = false) : m(mm), lk(mm->mu) {
    (void)hipSetDevice(m->device);
    if (!keep_deferred) rc = flush_pending(m);
  }
};

int64_t export_sorted(vdbgpu_map* m, std::vector<unsigned long long>& keys, std::vector<float>& vals,
                      std::vector<unsigned long long>& act, int32_t* origins, float* values,
                      uint64_t* masks, int64_t cap, bool with_vals) {
  const int64_t n = (int64_t)keys.size();
  if (!origins && !values && !masks) return n;
  if (n > cap) return -n;
  std::vector<std::pair<std::array<int, 3>, uint32_t>> order(n);
  for (int64_t i = 0; i < n; ++i) {
    std::array<int, 3> o;
    key_origin(keys[i], o.data());
    order[i] = {o, (uint32_t)i};
  }
  std::sort(order.begin(), order.end());
  for (int64_t j = 0; j < n; ++j) {
    const uint32_t i = order[j].second;
    if (origins) std::memcpy(origins + 3 * j, order[j].first.data(), 12);
    if (values && with_vals) std::memcpy(values + 512 * j, &vals[(size_t)i * 512], 2048);
    if (masks) std::memcpy(masks + 8 * j, &act[(size_t)i * 8], 64);
  }
  (void)m;
  return n;
}

}  // namespace

extern "C" {

int32_t vdbgpu_version(void) { return 100; }

const char* vdbgpu_status_string(int32_t s) {
  switch (s) {
    case VDBGPU_OK: return "ok";
    case VDBGPU_NOT_CONFIGURED: return "not configured";
    case VDBGPU_UNKNOWN_SOURCE: return "unknown source";
    case VDBGPU_BAD_MAX_RANGE: return "max_range < 0";
    case VDBGPU_BAD_PROB_MISS: return "prob_miss > 0.5";
    case VDBGPU_BAD_PROB_HIT: return "prob_hit < 0.5";
    case VDBGPU_OUT_OF_RANGE: return "voxel index out of range";
    case VDBGPU_UNSUPPORTED: return "unsupported";
    case VDBGPU_DEVICE_ERROR: return "device error";
    case VDBGPU_BAD_ARGUMENT: return "bad argument";
    case VDBGPU_NO_INTERSECTOR: return "no intersector";
    case VDBGPU_OVER_BUDGET: return "over HBM budget";
    default: return "unknown status";
  }
}

int32_t vdbgpu_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

static bool create_rec_events(vdbgpu_map* m) {
  for (int i = 0; i < vdbgpu_map::kRecSets; ++i)
    if (hipEventCreateWithFlags(&m->rec_free[i], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&m->ins_done[i], hipEventDisableTiming) != hipSuccess)
      return false;
  return hipHostMalloc((void**)&m->h_lc, vdbgpu_map::kRecSets * sizeof(uint32_t), hipHostMallocDefault) ==
             hipSuccess &&
         hipHostGetDevicePointer((void**)&m->d_lc, m->h_lc, 0) == hipSuccess;
}

int32_t vdbgpu_create(double resolution, vdbgpu_map** out) {
  if (!out) return VDBGPU_BAD_ARGUMENT;
  *out = nullptr;
  if (!(resolution > 0) || !std::isfinite(resolution)) return VDBGPU_BAD_ARGUMENT;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return VDBGPU_DEVICE_ERROR;
  vdbgpu_map* m = new vdbgpu_map();
  m->res = resolution;
  m->inv = 1.0 / resolution;  // ScaleMap keeps 1/scale and multiplies (worldToIndex)
  if (hipGetDevice(&m->device) != hipSuccess) {
    delete m;
    return VDBGPU_DEVICE_ERROR;
  }
  int cu = 0;
  if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, m->device) == hipSuccess && cu > 0)
    m->num_cu = cu;
  if (hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&m->cast, hipStreamNonBlocking) != hipSuccess ||
      !create_rec_events(m) ||
      hipMalloc((void**)&m->d_st, 4 * sizeof(Status)) != hipSuccess ||
      hipHostMalloc((void**)&m->h_st, 4 * sizeof(Status), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&m->h_cst, 4 * sizeof(Status), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&m->d_dest, sizeof(RecDest)) != hipSuccess ||
      hipEventCreateWithFlags(&m->cst_done[0], hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&m->cst_done[1], hipEventDisableTiming) != hipSuccess ||
      hipMemsetAsync(m->d_st, 0, 4 * sizeof(Status), m->stream) != hipSuccess) {
    vdbgpu_destroy(m);
    return VDBGPU_DEVICE_ERROR;
  }
  memset(m->h_st, 0, 4 * sizeof(Status));
  memset(m->h_cst, 0, 4 * sizeof(Status));
  if (hipMalloc((void**)&m->d_scans, kMaxBatchScans * sizeof(ScanInfo)) != hipSuccess ||
      hipMalloc((void**)&m->d_ctl, sizeof(BatchCtl)) != hipSuccess ||
      hipHostMalloc((void**)&m->h_scans, kMaxBatchScans * sizeof(ScanInfo), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&m->h_ctl, sizeof(BatchCtl), hipHostMallocDefault) != hipSuccess ||
      hipMalloc((void**)&m->noop4096, vdbgpu_map::kNoopCap * 8) != hipSuccess ||
      launch_fill_u64(m->stream, m->noop4096, vdbgpu_map::kNoopCap, kEmpty) != hipSuccess) {
    vdbgpu_destroy(m);
    return VDBGPU_DEVICE_ERROR;
  }
  *out = m;
  return VDBGPU_OK;
}

void vdbgpu_destroy(vdbgpu_map* m) {
  if (!m) return;
  (void)vdbgpu_async_stop(m);  // the ingestion threads call into this handle
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  if (m->cast) (void)hipStreamSynchronize(m->cast);
  for (auto& kv : m->sources) free_update(m, kv.second);
  free_update(m, m->edit);
  free_update(m, m->art);
  free_update(m, m->art_tmp);
  dfree(m, &m->art_eidx, m->art_eidx_cap);
  dfree(m, &m->mkeys, m->mcap);
  dfree(m, &m->midx, m->mcap);
  if (m->vmm == 1) {
    ga_release(m, m->g_values);
    ga_release(m, m->g_active);
    ga_release(m, m->g_keys);
    ga_release(m, m->g_dirty);
  } else {
    dfree(m, &m->values, (size_t)m->lcap * 512);
    dfree(m, &m->active, (size_t)m->lcap * 8);
    dfree(m, &m->leaf_keys, m->lcap);
    dfree(m, &m->dirty, m->lcap);
  }
  free_tiles(m);
  dfree(m, &m->d_points, m->points_bytes / 4);
  const uint32_t nb = (m->ray_cap + kPrepBlock - 1) / kPrepBlock;
  dfree(m, &m->rayinfo, m->ray_cap);
  dfree(m, &m->pre, m->ray_cap);
  dfree(m, &m->bsum, nb);
  dfree(m, &m->boff, nb);
  dfree(m, &m->eidx, m->eidx_cap);
  dfree(m, &m->bkeys, m->bcap);
  dfree(m, &m->bcnt, (size_t)m->bcap * 2);
  dfree(m, &m->bbsum, m->bcap / 256);
  dfree(m, &m->bboff, m->bcap / 256);
  dfree(m, &m->segs, m->seg_cap);
  dfree(m, &m->items, m->item_cap);
  dfree(m, &m->split_slot, m->split_cap);
  dfree(m, &m->shared_bits, (size_t)m->split_cap * 1024);
  dfree(m, &m->wkey, m->wtab_cap);
  dfree(m, &m->wsc, m->wtab_cap);
  dfree(m, &m->rkeys, m->rec_cap);
  dfree(m, &m->rmasks, (size_t)m->rec_cap * 16);
  dfree(m, &m->rlist, m->rec_cap);
  for (int i = 0; i < vdbgpu_map::kRecSets; ++i) {
    free_set(m, i);
    if (m->rec_free[i]) (void)hipEventDestroy(m->rec_free[i]);
    if (m->ins_done[i]) (void)hipEventDestroy(m->ins_done[i]);
  }
  if (m->h_lc) (void)hipHostFree(m->h_lc);
  dfree(m, &m->ident, m->ident_cap);
  dfree(m, &m->gslot, m->gslot_cap);
  dfree(m, &m->rgrp, m->rgrp_cap);
  dfree(m, &m->gl, m->grp_cap);
  dfree(m, &m->grec, m->grp_cap * kDeferMax);
  dfree(m, &m->ngroups, 1);
  dfree(m, &m->leaf_mask, m->group_cap);
  dfree(m, &m->lbsum, m->group_cap / 256);
  dfree(m, &m->lboff, m->group_cap / 256);
  dfree(m, &m->roff, m->ray_cap);
  dfree(m, &m->first_ray, m->chunk_cap);
  dfree(m, &m->tbl, (uint64_t)m->tbl_n * (m->tbl_n + 1) / 2);
  dfree(m, &m->rt.leaf, m->rt_leaf_cap);
  dfree(m, &m->rt.n128, m->rt_n128_cap);
  dfree(m, &m->rt.n4096, m->rt_n4096_cap);
  dfree(m, &m->d_bbox, 6);
  dfree(m, &m->q_in, (size_t)m->q_cap * 7);
  dfree(m, &m->q_succ, (size_t)m->q_cap);
  dfree(m, &m->q_end, (size_t)m->q_cap * 3);
  if (m->noop4096) (void)hipFree(m->noop4096);
  if (m->d_scans) (void)hipFree(m->d_scans);
  if (m->d_ctl) (void)hipFree(m->d_ctl);
  if (m->h_scans) (void)hipHostFree(m->h_scans);
  if (m->h_ctl) (void)hipHostFree(m->h_ctl);
  dfree(m, &m->gkeys, m->gcap);
  dfree(m, &m->gscans, (size_t)m->gcap * 2);
  dfree(m, &m->gneed, m->gcap);
  dfree(m, &m->glist, m->gcap);
  dfree(m, &m->gidx, m->gcap);
  if (m->d_st) (void)hipFree(m->d_st);
  if (m->h_st) (void)hipHostFree(m->h_st);
  drop_cast_graphs(m);
  if (m->d_dest) (void)hipFree(m->d_dest);
  if (m->h_cst) (void)hipHostFree(m->h_cst);
  for (hipEvent_t e : m->cst_done)
    if (e) (void)hipEventDestroy(e);
  for (auto& t : m->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : m->event_pool) (void)hipEventDestroy(e);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  if (m->cast) (void)hipStreamDestroy(m->cast);
  delete m;
}

int32_t vdbgpu_set_config(vdbgpu_map* m, const vdbgpu_config* c) {
  if (!m || !c) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  int status = VDBGPU_OK;
  if (c->max_range < 0.0) {
    status = VDBGPU_BAD_MAX_RANGE;
  } else {
    m->max_range = c->max_range;
    m->fast_mode = c->fast_mode != 0;
    m->config_set = true;
  }
  if (c->prob_miss > 0.5) return VDBGPU_BAD_PROB_MISS;
  if (c->prob_hit < 0.5) return VDBGPU_BAD_PROB_HIT;
  m->lo[1] = (float)(std::log(c->prob_miss) - std::log(1 - c->prob_miss));
  m->lo[0] = (float)(std::log(c->prob_hit) - std::log(1 - c->prob_hit));
  m->lo[2] = (float)(std::log(c->prob_thres_min) - std::log(1 - c->prob_thres_min));
  m->lo[3] = (float)(std::log(c->prob_thres_max) - std::log(1 - c->prob_thres_max));
  m->lo[4] = (float)(std::log(0.99) - std::log(0.01));
  m->lo[5] = (float)(std::log(0.01) - std::log(0.99));
  m->config_set = true;
  return status;
}

int32_t vdbgpu_logodds(vdbgpu_map* m, float out[6]) {
  if (!m || !out) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  std::memcpy(out, m->lo, sizeof m->lo);
  return VDBGPU_OK;
}

int32_t vdbgpu_add_source(vdbgpu_map* m, const char* id, double max_range, double max_rate) {
  (void)max_rate;
  if (!m || !id) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  Source& s = m->sources[id];
  s.id = id;
  s.max_range = (max_range == 0) ? m->max_range : max_range;
  if (s.cap) {  // re-adding a source starts it with an empty update grid
    free_update(m, s);
    s.pending = 0;
  }
  s.has_isect = false;  // a new InputSource has no intersector (VDBMapping.hpp:1352-1375)
  return VDBGPU_OK;
}

int32_t vdbgpu_remove_source(vdbgpu_map* m, const char* id) {
  if (!m || !id) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  auto it = m->sources.find(id);
  if (it == m->sources.end()) return fail(m, VDBGPU_UNKNOWN_SOURCE, std::string("no source ") + id);
  free_update(m, it->second);
  m->sources.erase(it);
  return VDBGPU_OK;
}

int32_t vdbgpu_has_source(vdbgpu_map* m, const char* id) {
  if (!m || !id) return VDBGPU_BAD_ARGUMENT;
  Guard g(m, true);  // a lookup: waiting scans stay waiting
  return m->sources.count(id) ? VDBGPU_OK : VDBGPU_UNKNOWN_SOURCE;
}

int32_t vdbgpu_accumulate(vdbgpu_map* m, const char* id, const float* xyz, int64_t n, int64_t stride,
                          const double origin[3], int32_t on_device) {
  if (!m || !id) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  auto it = m->sources.find(id);
  if (it == m->sources.end()) return fail(m, VDBGPU_UNKNOWN_SOURCE, std::string("Tried to accumulate update for ") + id + ". Source not available");
  if (!(it->second.max_range > 0)) return VDBGPU_OK;
  return do_accumulate(m, it->second, xyz, n, stride, origin, on_device != 0);
}

int32_t vdbgpu_integrate(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  return do_integrate(m);
}

int32_t vdbgpu_insert(vdbgpu_map* m, const char* id, const float* xyz, int64_t n, int64_t stride,
                      const double origin[3], int32_t on_device) {
  if (!m || !id) return VDBGPU_BAD_ARGUMENT;
  Guard g(m, true);  // a brick-path scan joins the deferred ones
  int status = VDBGPU_OK;
  auto it = m->sources.find(id);
  if (it == m->sources.end())
    status = fail(m, VDBGPU_UNKNOWN_SOURCE, std::string("Tried to accumulate update for ") + id + ". Source not available");
  else if (it->second.max_range > 0)
    status = insert_cast(m, it->second, xyz, n, stride, origin, on_device != 0);
  const int r = do_integrate(m);
  return status ? status : r;
}

int32_t vdbgpu_insert_batch(vdbgpu_map* m, const char* id, int32_t nscans, const float* xyz,
                            const int64_t* counts, int64_t stride, const double* origins,
                            const double* ranges, int32_t on_device) {
  if (!m || !id || nscans < 0 || (nscans && (!counts || !origins))) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  auto it = m->sources.find(id);
  bool any_range = false;
  if (it != m->sources.end())
    for (int i = 0; i < nscans && !any_range; ++i) any_range = (ranges ? ranges[i] : it->second.max_range) > 0;
  if (it == m->sources.end() || !m->config_set || !any_range || nscans == 0) {
    // every insert would only integrate (nothing cast): one integrate does the same
    int status = VDBGPU_OK;
    if (nscans > 0) {
      if (it == m->sources.end())
        status = fail(m, VDBGPU_UNKNOWN_SOURCE, std::string("Tried to accumulate update for ") + id + ". Source not available");
      else if (!m->config_set)
        status = fail(m, VDBGPU_NOT_CONFIGURED, "Map not properly configured. Did you call setConfig method?");
    }
    const int r = do_integrate(m);
    return status ? status : r;
  }
  Source& src = it->second;
  if (stride < 12 || (stride & 3) || (!xyz && nscans)) return fail(m, VDBGPU_BAD_ARGUMENT, "bad point buffer arguments");
  std::vector<int64_t> off(nscans + 1, 0);
  for (int i = 0; i < nscans; ++i) {
    if (counts[i] < 0) return fail(m, VDBGPU_BAD_ARGUMENT, "negative point count");
    off[i + 1] = off[i] + counts[i];
  }
  auto scan_range = [&](int i) { return ranges ? ranges[i] : src.max_range; };
  auto scan_ptr = [&](int i) {
    return reinterpret_cast<const float*>(reinterpret_cast<const char*>(xyz) + off[i] * stride);
  };
  // one scan through the plain path: accumulate (with the scan's range) + integrate
  auto plain_insert = [&](int i) -> int {
    const double keep = src.max_range;
    src.max_range = scan_range(i);
    int st = VDBGPU_OK;
    if (src.max_range > 0 && counts[i] > 0)
      st = insert_cast(m, src, scan_ptr(i), counts[i], stride, origins + 3 * i, on_device != 0);
    src.max_range = keep;
    const int r = do_integrate(m);
    return st ? st : r;
  };
  int i0 = 0;
  bool pending = false;
  for (auto& kv : m->sources) pending = pending || kv.second.pending > 0;
  if (m->tuning.traversal == 0 && !pending && !m->fast_mode && !m->ntiles) {
    // Every scan through the brick path, one scan of lookahead: the cast of
    // scan i is enqueued before the status of scan i - 1 is read, so the cast
    // stream does not wait for the host between scans; scan i - 1's map-stream
    // work is enqueued once its status is checked (overlapping scan i's cast),
    // in scan order. A scan with nothing to cast takes the plain path after
    // the scan before it is through.
    // VDBGPU_GRAPH=1: the casts are replayed from captured graphs instead of
    // launched one by one (the same results; measured no faster: the GPU, not
    // the host, sets the pace); the flag ends with the call
    struct GraphCasts {
      vdbgpu_map* m;
      explicit GraphCasts(vdbgpu_map* mm) : m(mm) {
        const char* on = std::getenv("VDBGPU_GRAPH");
        m->graph_casts = on && on[0] == '1';
      }
      ~GraphCasts() { m->graph_casts = false; }
    } graph_casts(m);
    BrickCast prev;
    bool have_prev = false;
    auto finish_prev = [&]() -> int {
      if (!have_prev) return VDBGPU_OK;
      have_prev = false;
      int r = brick_check(m, prev);
      if (!r) r = commit_staged(m);
      if (!r && m->npend >= m->tuning.defer) r = flush_pending(m);
      return r;
    };
    for (int i = 0; i < nscans; ++i) {
      const double R = scan_range(i);
      if (!(R > 0) || counts[i] <= 0 || counts[i] > (int64_t)kMaxRays) {
        int r = finish_prev();
        if (!r) r = plain_insert(i);
        if (r) return r;
        continue;
      }
      BrickCast c = brick_cast(src, scan_ptr(i), counts[i], stride, origins + 3 * i, on_device != 0);
      c.sc.R = R;
      int r = brick_tables(m, counts[i], R);
      if (!r) r = brick_enqueue(m, c, 0);
      const int rp = finish_prev();  // the scans before this one are applied as in sequential inserts
      if (r || rp) return rp ? rp : r;
      prev = c;
      have_prev = true;
    }
    if (have_prev) {  // the last scan stays staged, as after a plain insert
      have_prev = false;
      const int r = brick_check(m, prev);
      if (r) return r;
    }
    return refresh_snapshot(m);
  }
  if ((m->tuning.traversal == 0 && !pending) || m->fast_mode || m->ntiles) {
    // Streaming every scan through the brick path (the next scan's cast
    // overlaps this one's map update) is faster than the (leaf, scan) fold;
    // a fast-mode cast depends on the map state, and the tiles of a pruned
    // map are checked after every scan, so both run scan by scan.
    for (int i = 0; i < nscans; ++i) {
      const int r = plain_insert(i);
      if (r) return r;
    }
    return refresh_snapshot(m);
  }
  if (pending) {
    // pending sets are applied with the first scan's integrate, in source-id order
    const int r = plain_insert(0);
    if (r) return r;
    i0 = 1;
  }
  while (i0 < nscans) {
    int i1 = i0;
    uint64_t rays = 0;
    while (i1 < nscans && i1 - i0 < kMaxBatchScans && rays + (uint64_t)counts[i1] <= kMaxRays) rays += counts[i1++];
    if (i1 == i0) return fail(m, VDBGPU_BAD_ARGUMENT, "a scan has more than 2^30 points");
    std::vector<ScanInfo> scans;
    for (int i = i0; i < i1; ++i) {
      ScanInfo sc{};
      for (int a = 0; a < 3; ++a) sc.o[a] = origins[3 * i + a];
      sc.R = scan_range(i);
      sc.ray0 = (uint32_t)(off[i] - off[i0]);
      sc.nrays = (uint32_t)counts[i];
      scans.push_back(sc);
    }
    int r = rays ? run_pipeline(m, src, scan_ptr(i0), stride, on_device != 0, scans, (uint32_t)rays, true)
                 : VDBGPU_OK;
    if (r == VDBGPU_OUT_OF_RANGE) {
      // outside the batched key range (or a bad scan): exact per-scan semantics
      int first = VDBGPU_OK;
      for (int i = i0; i < i1; ++i) {
        const int ri = plain_insert(i);
        if (ri && !first) first = ri;
      }
      if (first) return first;
      r = VDBGPU_OK;
    }
    if (r) return r;
    i0 = i1;
  }
  return refresh_snapshot(m);
}

int32_t vdbgpu_reset(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  int r;
  if ((r = clear_grid(m))) return r;
  for (auto& kv : m->sources) {
    Source& s = kv.second;
    if (s.cap) {
      CK(launch_fill_u64(m->stream, s.keys, s.cap, kEmpty));
      CK(hipMemsetAsync(s.masks, 0, (size_t)s.cap * 16 * sizeof(unsigned long long), m->stream));
    }
    s.pending = 0;
  }
  // resetMap keeps the artificial grid (VM:174-186), not its leaves
  CK(hipStreamSynchronize(m->stream));
  return VDBGPU_OK;
}

int32_t vdbgpu_empty(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  return m->touched ? 0 : 1;
}

int64_t vdbgpu_leaf_count(vdbgpu_map* m) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  return live_leaves(m);  // tiles are not leaves
}

int64_t vdbgpu_export_leaves(vdbgpu_map* m, int32_t* origins, float* values, uint64_t* masks,
                             int64_t cap) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  if (read_status(m)) return -1;
  const uint32_t n = m->h_st[1].leaf_count;
  if (!m->ntiles) {
    if (!origins && !values && !masks) return n;
    if ((int64_t)n > cap) return -(int64_t)n;
  } else if (!origins && !values && !masks) {
    return live_leaves(m);
  }
  std::vector<unsigned long long> keys(n), act((size_t)n * 8);
  std::vector<float> vals(values ? (size_t)n * 512 : 0);
  if (n) {
    if (hipMemcpy(keys.data(), m->leaf_keys, n * sizeof(unsigned long long), hipMemcpyDeviceToHost) ||
        hipMemcpy(act.data(), m->active, (size_t)n * 64, hipMemcpyDeviceToHost) ||
        (values && hipMemcpy(vals.data(), m->values, (size_t)n * 2048, hipMemcpyDeviceToHost))) {
      m->last_error = "export copy failed";
      return -1;
    }
  }
  std::vector<uint32_t> tb;
  if (tile_bits(m, tb, nullptr)) return -1;
  if (!tb.empty()) {  // tiles are not leaves: compact them out
    uint32_t k = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (bit_at(tb, i)) continue;
      keys[k] = keys[i];
      std::memcpy(&act[(size_t)k * 8], &act[(size_t)i * 8], 64);
      if (values) std::memcpy(&vals[(size_t)k * 512], &vals[(size_t)i * 512], 2048);
      ++k;
    }
    keys.resize(k);
    act.resize((size_t)k * 8);
    if (values) vals.resize((size_t)k * 512);
    if (!origins && !values && !masks) return k;
    if ((int64_t)k > cap) return -(int64_t)k;
  }
  return export_sorted(m, keys, vals, act, origins, values, masks, cap, values != nullptr);
}

int64_t vdbgpu_export_dirty_leaves(vdbgpu_map* m, int32_t* origins, float* values, uint64_t* masks,
                                   int64_t cap) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  if (read_status(m)) return -1;
  const uint32_t lc = m->h_st[1].leaf_count;
  if (lc == 0) return 0;
  if (hipMemsetAsync(&m->d_st[2], 0, sizeof(Status), m->stream) != hipSuccess) return -1;
  // the dirty list reuses the eidx scratch (>= leaf_count entries needed)
  if (lc > m->eidx_cap) {
    dfree(m, &m->eidx, m->eidx_cap);
    m->eidx_cap = (uint32_t)next_pow2(lc);
    if (dalloc(m, &m->eidx, m->eidx_cap) != hipSuccess) return -1;
  }
  const int grid = std::min<int>(4096, (int)((lc + 255) / 256));
  if (launch_dirty_list(m->stream, grid, m->dirty, lc, m->eidx, &m->d_st[2]) || read_status(m)) return -1;
  uint32_t nd = m->h_st[2].entries;
  std::vector<uint32_t> tb;
  if (tile_bits(m, tb, nullptr)) return -1;
  if (!tb.empty() && nd) {
    // tiles are not leaves: drop them from the list (a tile is unchanged since
    // the check after the last write, so its flag may go)
    std::vector<uint32_t> idx(nd);
    if (hipMemcpy(idx.data(), m->eidx, nd * 4, hipMemcpyDeviceToHost)) return -1;
    uint32_t k = 0;
    for (uint32_t i : idx)
      if (!bit_at(tb, i)) idx[k++] = i;
    nd = k;
    if (nd && hipMemcpy(m->eidx, idx.data(), nd * 4, hipMemcpyHostToDevice)) return -1;
  }
  if ((int64_t)nd > cap) {
    // put the flags back so that a retry with a larger buffer sees the same set
    std::vector<uint32_t> idx(nd);
    std::vector<uint8_t> one(1, 1);
    (void)hipMemcpy(idx.data(), m->eidx, nd * 4, hipMemcpyDeviceToHost);
    for (uint32_t i : idx) (void)hipMemcpy(m->dirty + i, one.data(), 1, hipMemcpyHostToDevice);
    return -(int64_t)nd;
  }
  if (nd == 0) return 0;
  unsigned long long *dk = nullptr, *da = nullptr;
  float* dv = nullptr;
  if (dalloc(m, &dk, nd) || dalloc(m, &da, (size_t)nd * 8) || dalloc(m, &dv, (size_t)nd * 512)) return -1;
  std::vector<unsigned long long> keys(nd), act((size_t)nd * 8);
  std::vector<float> vals((size_t)nd * 512);
  int64_t r = -1;
  if (!launch_gather_leaves(m->stream, std::min<int>(4096, (int)((nd + 3) / 4)), m->eidx, nd, map_view(m), dk, dv, da) &&
      !hipStreamSynchronize(m->stream) &&
      !hipMemcpy(keys.data(), dk, nd * 8, hipMemcpyDeviceToHost) &&
      !hipMemcpy(act.data(), da, (size_t)nd * 64, hipMemcpyDeviceToHost) &&
      !hipMemcpy(vals.data(), dv, (size_t)nd * 2048, hipMemcpyDeviceToHost))
    r = export_sorted(m, keys, vals, act, origins, values, masks, cap, true);
  dfree(m, &dk, nd);
  dfree(m, &da, (size_t)nd * 8);
  dfree(m, &dv, (size_t)nd * 512);
  return r;
}

int64_t vdbgpu_update_leaf_count(vdbgpu_map* m, const char* id) {
  if (!m || !id) return -1;
  Guard g(m);
  if (g.rc) return -1;
  auto it = m->sources.find(id);
  if (it == m->sources.end()) return -1;
  return it->second.pending;
}

int64_t vdbgpu_export_update(vdbgpu_map* m, const char* id, int32_t* origins, uint64_t* visited,
                             uint64_t* hit, int64_t cap) {
  if (!m || !id) return -1;
  Guard g(m);
  if (g.rc) return -1;
  auto it = m->sources.find(id);
  if (it == m->sources.end()) return -1;
  Source& s = it->second;
  if (!origins && !visited && !hit) return s.pending;
  if ((int64_t)s.pending > cap) return -(int64_t)s.pending;
  if (s.cap == 0 || s.pending == 0) return 0;
  if (hipStreamSynchronize(m->stream)) return -1;
  std::vector<unsigned long long> keys(s.cap), masks((size_t)s.cap * 16);
  if (hipMemcpy(keys.data(), s.keys, (size_t)s.cap * 8, hipMemcpyDeviceToHost) ||
      hipMemcpy(masks.data(), s.masks, (size_t)s.cap * 128, hipMemcpyDeviceToHost))
    return -1;
  std::vector<std::pair<std::array<int, 3>, uint32_t>> order;
  for (uint32_t i = 0; i < s.cap; ++i) {
    if (keys[i] == kEmpty) continue;
    std::array<int, 3> o;
    key_origin(keys[i], o.data());
    order.push_back({o, i});
  }
  std::sort(order.begin(), order.end());
  if ((int64_t)order.size() > cap) return -(int64_t)order.size();
  for (size_t j = 0; j < order.size(); ++j) {
    const uint32_t i = order[j].second;
    if (origins) std::memcpy(origins + 3 * j, order[j].first.data(), 12);
    if (visited) std::memcpy(visited + 8 * j, &masks[(size_t)i * 16], 64);
    if (hit) std::memcpy(hit + 8 * j, &masks[(size_t)i * 16 + 8], 64);
  }
  return (int64_t)order.size();
}

// Status of a finished raytrace launch: rejected queries are reported.
static int32_t raytrace_status(vdbgpu_map* m) {
  const int r = read_status(m);
  if (r) return r;
  if (m->h_st[2].error)
    return fail(m, VDBGPU_BAD_ARGUMENT,
                "query rays with non-finite values or more than 2^24 voxel steps were rejected "
                "(success = 0, end point = NaN)");
  return VDBGPU_OK;
}

int32_t vdbgpu_raytrace(vdbgpu_map* m, int64_t n, const double* o, const double* d,
                        const double* len, uint8_t* succ, double* ends) {
  if (!m || n < 0 || (n && (!o || !d || !len || !succ || !ends))) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (!m->has_isect)
    return fail(m, VDBGPU_NO_INTERSECTOR, "raytrace needs an integrate in fast mode on a non-empty map");
  if (n == 0) return VDBGPU_OK;
  if (n > m->q_cap) {
    dfree(m, &m->q_in, (size_t)m->q_cap * 7);
    dfree(m, &m->q_succ, (size_t)m->q_cap);
    dfree(m, &m->q_end, (size_t)m->q_cap * 3);
    m->q_cap = n;
    CK(dalloc(m, &m->q_in, (size_t)n * 7));
    CK(dalloc(m, &m->q_succ, (size_t)n));
    CK(dalloc(m, &m->q_end, (size_t)n * 3));
  }
  CK(hipMemcpyAsync(m->q_in, o, (size_t)n * 24, hipMemcpyHostToDevice, m->stream));
  CK(hipMemcpyAsync(m->q_in + 3 * n, d, (size_t)n * 24, hipMemcpyHostToDevice, m->stream));
  CK(hipMemcpyAsync(m->q_in + 6 * n, len, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
  const int grid = (int)std::min<int64_t>((int64_t)m->num_cu * 16, (n + 255) / 256);
  CK(hipMemsetAsync(&m->d_st[2].error, 0, sizeof(unsigned int), m->stream));
  CK(launch_raytrace(m->stream, grid, m->rt, map_view(m), m->inv, m->res, n, m->q_in, m->q_in + 3 * n,
                     m->q_in + 6 * n, m->q_succ, m->q_end, &m->d_st[2].error));
  CK(hipMemcpyAsync(succ, m->q_succ, (size_t)n, hipMemcpyDeviceToHost, m->stream));
  CK(hipMemcpyAsync(ends, m->q_end, (size_t)n * 24, hipMemcpyDeviceToHost, m->stream));
  return raytrace_status(m);
}

int32_t vdbgpu_raytrace_device(vdbgpu_map* m, int64_t n, const double* o, const double* d,
                               const double* len, uint8_t* succ, double* ends) {
  if (!m || n < 0 || (n && (!o || !d || !len || !succ || !ends))) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (!m->has_isect)
    return fail(m, VDBGPU_NO_INTERSECTOR, "raytrace needs an integrate in fast mode on a non-empty map");
  if (n == 0) return VDBGPU_OK;
  const int grid = (int)std::min<int64_t>((int64_t)m->num_cu * 16, (n + 255) / 256);
  CK(hipMemsetAsync(&m->d_st[2].error, 0, sizeof(unsigned int), m->stream));
  {
    Timer t(m, kRaytrace);
    CK(launch_raytrace(m->stream, grid, m->rt, map_view(m), m->inv, m->res, n, o, d, len, succ, ends,
                       &m->d_st[2].error));
  }
  return raytrace_status(m);
}

}  // extern "C"

// ------------------------------------------------- allocation-free raytrace --
// vdbgpu_raytrace_noalloc stages the queries, the results and the rejected-query
// counter in this static device array instead of the handle's growable query
// buffers. Every GPU the library runs on holds its own copy until the process
// ends. In host code `scratch` names a host-side stand-in, never the device
// memory: its device address comes from hipGetSymbolAddress (per device, on the
// first call there) and only that address is used.
static __device__ __align__(256) unsigned char scratch[32u << 20];

namespace {

constexpr size_t kNoAllocScratchBytes = sizeof(scratch);
constexpr int64_t kNoAllocMaxQueries = 393216;  // 81 bytes per query + 5 x 256 bytes of alignment
static_assert(kNoAllocMaxQueries * 81 + 5 * 256 <= (int64_t)kNoAllocScratchBytes, "scratch too small");
constexpr int kNoAllocMaxDevices = 64;

struct NoAllocDevice {
  std::once_flag once;
  hipError_t err = hipSuccess;
  unsigned char* ptr = nullptr;  // device address of this GPU's `scratch`
  std::mutex mu;                 // calls on one GPU share its scratch: one at a time
};
NoAllocDevice g_noalloc[kNoAllocMaxDevices];

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

int32_t vdbgpu_raytrace_noalloc(vdbgpu_map* m, int64_t n, const double* o, const double* d,
                                const double* len, uint8_t* succ, double* ends) {
  if (!m || n < 0 || (n && (!o || !d || !len || !succ || !ends))) return VDBGPU_BAD_ARGUMENT;
  Guard g(m, true);  // applying waiting scans could allocate: they are refused below instead
  if (m->staged || m->npend)
    return fail(m, VDBGPU_BAD_ARGUMENT,
                "vdbgpu_raytrace_noalloc: map updates are still waiting to be applied (applying them may "
                "allocate); call vdbgpu_synchronize first");
  if (!m->has_isect)
    return fail(m, VDBGPU_NO_INTERSECTOR, "raytrace needs an integrate in fast mode on a non-empty map");
  if (n > kNoAllocMaxQueries)
    return fail(m, VDBGPU_BAD_ARGUMENT,
                "vdbgpu_raytrace_noalloc: at most 393216 queries per call (the library's static device "
                "scratch); split the batch");
  if (n == 0) return VDBGPU_OK;
  if (m->device < 0 || m->device >= kNoAllocMaxDevices)
    return fail(m, VDBGPU_UNSUPPORTED, "vdbgpu_raytrace_noalloc: device ordinal above 63");
  NoAllocDevice& nd = g_noalloc[m->device];
  // Guard made this handle's device current: the symbol address is that device's copy
  std::call_once(nd.once, [&nd] { nd.err = hipGetSymbolAddress((void**)&nd.ptr, HIP_SYMBOL(scratch)); });
  if (nd.err != hipSuccess || !nd.ptr)
    return fail(m, VDBGPU_DEVICE_ERROR,
                std::string("vdbgpu_raytrace_noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(nd.err));
  std::lock_guard<std::mutex> lk(nd.mu);
  size_t off = 0;
  auto carve = [&](size_t bytes) {
    unsigned char* p = nd.ptr + off;
    off += align256(bytes);
    return p;
  };
  unsigned int* err = (unsigned int*)carve(sizeof(unsigned long long));
  double* q_o = (double*)carve((size_t)n * 24);
  double* q_d = (double*)carve((size_t)n * 24);
  double* q_l = (double*)carve((size_t)n * 8);
  uint8_t* q_s = (uint8_t*)carve((size_t)n);
  double* q_e = (double*)carve((size_t)n * 24);
  if (off > kNoAllocScratchBytes) return fail(m, VDBGPU_BAD_ARGUMENT, "vdbgpu_raytrace_noalloc: scratch overrun");
  // the counter keeps the last call's value: clear it with a kernel, not a memset node
  CK(launch_zero_words(m->stream, (unsigned long long*)err, 1));
  CK(hipMemcpyAsync(q_o, o, (size_t)n * 24, hipMemcpyHostToDevice, m->stream));
  CK(hipMemcpyAsync(q_d, d, (size_t)n * 24, hipMemcpyHostToDevice, m->stream));
  CK(hipMemcpyAsync(q_l, len, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
  const int grid = (int)std::min<int64_t>((int64_t)m->num_cu * 16, (n + 255) / 256);
  CK(launch_raytrace(m->stream, grid, m->rt, map_view(m), m->inv, m->res, n, q_o, q_d, q_l, q_s, q_e, err));
  unsigned int h_err = 0;
  CK(hipMemcpyAsync(succ, q_s, (size_t)n, hipMemcpyDeviceToHost, m->stream));
  CK(hipMemcpyAsync(ends, q_e, (size_t)n * 24, hipMemcpyDeviceToHost, m->stream));
  CK(hipMemcpyAsync(&h_err, err, sizeof h_err, hipMemcpyDeviceToHost, m->stream));
  CK(hipStreamSynchronize(m->stream));
  if (h_err)
    return fail(m, VDBGPU_BAD_ARGUMENT,
                "query rays with non-finite values or more than 2^24 voxel steps were rejected "
                "(success = 0, end point = NaN)");
  return VDBGPU_OK;
}

int32_t vdbgpu_device_alloc(int64_t bytes, void** out) {
  if (!out || bytes <= 0) return VDBGPU_BAD_ARGUMENT;
  *out = nullptr;
  return hipMalloc(out, (size_t)bytes) == hipSuccess ? VDBGPU_OK : VDBGPU_DEVICE_ERROR;
}

int32_t vdbgpu_device_free(void* ptr) {
  if (!ptr) return VDBGPU_OK;
  return hipFree(ptr) == hipSuccess ? VDBGPU_OK : VDBGPU_DEVICE_ERROR;
}

int32_t vdbgpu_device_upload(void* dst, const void* src, int64_t bytes) {
  if (!dst || !src || bytes < 0) return VDBGPU_BAD_ARGUMENT;
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyHostToDevice) == hipSuccess ? VDBGPU_OK
                                                                                 : VDBGPU_DEVICE_ERROR;
}

int32_t vdbgpu_device_download(void* dst, const void* src, int64_t bytes) {
  if (!dst || !src || bytes < 0) return VDBGPU_BAD_ARGUMENT;
  return hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToHost) == hipSuccess ? VDBGPU_OK
                                                                                 : VDBGPU_DEVICE_ERROR;
}

int32_t vdbgpu_device_copy(void* dst, const void* src, int64_t bytes) {
  if (!dst || !src || bytes < 0) return VDBGPU_BAD_ARGUMENT;
  // a device-to-device hipMemcpy may return before the copy has run: wait for it,
  // so that the call's duration is the copy's
  if (hipMemcpy(dst, src, (size_t)bytes, hipMemcpyDeviceToDevice) != hipSuccess) return VDBGPU_DEVICE_ERROR;
  return hipStreamSynchronize(nullptr) == hipSuccess ? VDBGPU_OK : VDBGPU_DEVICE_ERROR;
}

int32_t vdbgpu_release_scratch(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);  // applies the waiting scans first: nothing in flight uses the scratch after the syncs
  if (g.rc) return g.rc;
  CK(hipStreamSynchronize(m->cast));
  CK(hipStreamSynchronize(m->stream));
  drop_cast_graphs(m);  // their buffers go
  walk_scratch(m, true);
  return VDBGPU_OK;
}

int32_t vdbgpu_synchronize(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  CK(hipStreamSynchronize(m->cast));
  return read_status(m);
}

int32_t vdbgpu_get_stats(vdbgpu_map* m, vdbgpu_stats* out) {
  if (!m || !out) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  int r = read_status(m);
  if (r) return r;
  std::vector<uint32_t> tb;
  if ((r = tile_bits(m, tb, nullptr))) return r;
  m->stats.leaf_count = m->h_st[1].leaf_count - count_bits(tb);  // tiles are not leaves
  m->stats.voxel_updates = m->h_st[3].voxels;
  m->stats.device_bytes = m->device_bytes;
  m->stats.scratch_bytes = walk_scratch(m, false);
  *out = m->stats;
  // batch folds count on the host, every k_map_insert on the device
  out->update_leaves = m->stats.update_leaves + m->h_st[3].entries;
  return VDBGPU_OK;
}

int32_t vdbgpu_set_profiling(vdbgpu_map* m, int32_t on) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  m->profiling = on != 0;
  return VDBGPU_OK;
}

int32_t vdbgpu_set_tuning(vdbgpu_map* m, const vdbgpu_tuning* t) {
  if (!m || !t) return VDBGPU_BAD_ARGUMENT;
  if (t->block < 64 || t->block > kTraverseMaxBlock || (t->block & 63) || t->steps_per_lane < 1 ||
      t->steps_per_lane > 1 << 20 || t->waves_per_cu < 1 || t->waves_per_cu > 64 ||
      !(t->load_factor >= 0.2 && t->load_factor <= 0.95) || t->traversal < 0 || t->traversal > 1 ||
      t->defer < 0 || t->defer > kDeferMax)
    return VDBGPU_BAD_ARGUMENT;
  Guard g(m);  // applies the scans deferred under the old setting
  if (g.rc) return g.rc;
  const int sets = t->defer > 1 ? 2 * t->defer + 2 : 2;  // as at creation (kRecSets for the default 8)
  if (sets != m->rec_sets) {  // a new ring of record sets
    CK(hipStreamSynchronize(m->cast));
    CK(hipStreamSynchronize(m->stream));
    for (int i = 0; i < vdbgpu_map::kRecSets; ++i) free_set(m, i);
    m->rec_sets = sets;
    m->rec_cur = 0;
  }
  m->tuning = *t;
  return VDBGPU_OK;
}

int32_t vdbgpu_get_tuning(vdbgpu_map* m, vdbgpu_tuning* out) {
  if (!m || !out) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  *out = m->tuning;
  return VDBGPU_OK;
}

const char* vdbgpu_last_error(vdbgpu_map* m) { return m ? m->last_error.c_str() : ""; }

// ------------------------------------------------------------ map edits --

}  // extern "C"

namespace {

// Mark the voxels floor(p * 1/res) of n host points in the edit set (a NaN
// or out-of-range point clears the set and returns its status); *count =
// the set's leaves.
int mark_points(vdbgpu_map* m, const float* xyz, int64_t n, int64_t stride, uint32_t* count) {
  *count = 0;
  if (n == 0) return VDBGPU_OK;
  const size_t bytes = (size_t)(n - 1) * stride + 12;
  float* pts = nullptr;
  CK(dalloc(m, &pts, (bytes + 3) / 4));
  CK(hipMemcpyAsync(pts, xyz, bytes, hipMemcpyHostToDevice, m->stream));
  // at most n leaves; large edits start smaller and grow on overflow
  const uint32_t cap0 = (uint32_t)std::max<uint64_t>(kMinUpdateCap, next_pow2(std::min<uint64_t>(2 * (uint64_t)n, 1u << 22)));
  const int r = mark_set(m, m->edit, cap0, [&]() -> int {
    CK(launch_mark_points(m->stream, pts, stride, (uint32_t)n, m->inv, upd_view(m->edit), &m->d_st[2]));
    return VDBGPU_OK;
  }, count);
  dfree(m, &pts, (bytes + 3) / 4);  // hipFree waits for the kernels that read it
  return r;
}

// setNodeToOccupied / setNodeToFree on the marked voxels; clears the edit set.
int apply_points(vdbgpu_map* m, uint32_t count, bool occupied) {
  const float val = occupied ? m->lo[4] : m->lo[5];
  const bool need = occupied || val != 0.0f;  // does the op change a background voxel?
  int r;
  if (!(r = ensure_map(m, m->leaf_count_ub + count)) && !(r = ensure_eidx(m, count)) &&
      !(r = apply_edit(m, m->edit, count, m->eidx, occupied ? kEditOccupied : kEditFree, val, need, true))) {
    m->leaf_count_ub += count;
    ledger_add(m, count);
    return VDBGPU_OK;
  }
  (void)clear_set(m, m->edit);  // no stale marks for the next call
  return r;
}

// The grid emptied (leaf pool, leaf hash, root children); update sets and
// the artificial grid are kept.
int clear_grid(vdbgpu_map* m) {
  if (m->lcap) {
    CK(hipMemsetAsync(m->values, 0, (size_t)m->lcap * 512 * sizeof(float), m->stream));
    CK(hipMemsetAsync(m->active, 0, (size_t)m->lcap * 8 * sizeof(unsigned long long), m->stream));
    CK(hipMemsetAsync(m->dirty, 0, m->lcap, m->stream));
  }
  if (m->mcap) CK(launch_fill_u64(m->stream, m->mkeys, m->mcap, kEmpty));
  CK(launch_fill_u64(m->stream, m->noop4096, vdbgpu_map::kNoopCap, kEmpty));
  CK(hipMemsetAsync(&m->d_st[1], 0, sizeof(Status), m->stream));
  m->leaf_count_ub = 0;
  m->ledger.clear();
  m->ins_fifo.clear();
  free_tiles(m);
  m->touched = false;
  m->snap_dirty = true;
  m->art_leaves = false;
  return VDBGPU_OK;
}

}  // namespace

extern "C" {

int32_t vdbgpu_set_points(vdbgpu_map* m, const float* xyz, int64_t n, int64_t stride, int32_t occupied) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (n < 0 || stride < 12 || (stride & 3) || n > (int64_t)kMaxRays || (n && !xyz))
    return fail(m, VDBGPU_BAD_ARGUMENT, "bad point buffer arguments");
  uint32_t count = 0;
  int r;
  if ((r = mark_points(m, xyz, n, stride, &count))) return r;
  return count ? apply_points(m, count, occupied != 0) : VDBGPU_OK;
}

int32_t vdbgpu_create_from_points(vdbgpu_map* m, const float* xyz, int64_t n, int64_t stride,
                                  int32_t set_background, int32_t clear_map) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (n < 0 || stride < 12 || (stride & 3) || n > (int64_t)kMaxRays || (n && !xyz))
    return fail(m, VDBGPU_BAD_ARGUMENT, "bad point buffer arguments");
  CK(hipStreamSynchronize(m->cast));
  uint32_t count = 0;
  int r;
  if ((r = mark_points(m, xyz, n, stride, &count))) return r;  // validated before any change
  if (clear_map && (r = clear_grid(m))) return r;
  // setValueOn(ijk, max log-odds): the same voxel state as setNodeToOccupied
  if (count && (r = apply_points(m, count, true))) return r;
  if (!set_background) return prune_tiles(m);
  if ((r = read_status(m))) return r;
  const uint32_t L = m->h_st[1].leaf_count;
  if (L == 0) return prune_tiles(m);
  int* d_box = nullptr;
  const int init[6] = {INT_MAX, INT_MAX, INT_MAX, INT_MIN, INT_MIN, INT_MIN};
  if ((r = upload(m, &d_box, init, 6))) return r;
  CK(launch_active_bbox(m->stream, map_view(m), L, d_box));
  int box[6];
  CK(hipMemcpyAsync(box, d_box, sizeof box, hipMemcpyDeviceToHost, m->stream));
  CK(hipStreamSynchronize(m->stream));
  if (box[0] > box[3]) {  // no active voxel: an empty box
    dfree(m, &d_box, 6);
    return prune_tiles(m);
  }
  const int lo[3] = {box[0] >> 3, box[1] >> 3, box[2] >> 3};
  const uint32_t dims[3] = {(uint32_t)((box[3] >> 3) - lo[0] + 1), (uint32_t)((box[4] >> 3) - lo[1] + 1),
                            (uint32_t)((box[5] >> 3) - lo[2] + 1)};
  const uint64_t nb = (uint64_t)dims[0] * dims[1] * dims[2];
  if (nb > (1ull << 27)) {
    dfree(m, &d_box, 6);
    return fail(m, VDBGPU_UNSUPPORTED, "background box holds more than 2^27 leaves");
  }
  const uint32_t N = (uint32_t)nb;
  const bool need = m->lo[5] != 0.0f;  // setValueOff(min log-odds) changes a background voxel
  unsigned long long* keys = nullptr;
  uint8_t* needs = nullptr;
  uint32_t* eidx = nullptr;
  if ((r = ensure_map(m, m->leaf_count_ub + N))) return r;
  CK(dalloc(m, &keys, N));
  CK(dalloc(m, &needs, N));
  CK(dalloc(m, &eidx, N));
  CK(launch_box_keys(m->stream, lo, dims, keys, needs, need));
  CK(launch_upsert_leaves(m->stream, keys, needs, N, map_view(m), eidx, nullptr, &m->d_st[1]));
  CK(launch_box_fill(m->stream, keys, eidx, N, d_box, map_view(m), m->lo[5]));
  m->leaf_count_ub += N;
  ledger_add(m, N);
  m->snap_dirty = true;
  if ((r = read_status(m))) return r;
  m->touched = m->touched || m->h_st[1].leaf_count > 0;
  dfree(m, &keys, N);
  dfree(m, &needs, N);
  dfree(m, &eidx, N);
  dfree(m, &d_box, 6);
  return prune_tiles(m);  // pruneGrid (OCC:173)
}

int64_t vdbgpu_export_tiles(vdbgpu_map* m, int32_t* origins, float* values, uint8_t* active, int64_t cap) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  std::vector<uint32_t> tb, on;
  if (read_status(m) || tile_bits(m, tb, &on)) return -1;
  const int64_t n = count_bits(tb);
  if (!origins && !values && !active) return n;
  if (n > cap) return -n;
  if (n == 0) return 0;
  const uint32_t L = std::min<uint32_t>(m->h_st[1].leaf_count, m->tcap);
  std::vector<unsigned long long> keys(L);
  std::vector<float> tv(L);
  if (hipMemcpy(keys.data(), m->leaf_keys, (size_t)L * 8, hipMemcpyDeviceToHost) ||
      hipMemcpy(tv.data(), m->tval, (size_t)L * 4, hipMemcpyDeviceToHost)) {
    m->last_error = "export copy failed";
    return -1;
  }
  std::vector<std::pair<std::array<int, 3>, uint32_t>> order;
  for (uint32_t i = 0; i < L; ++i) {
    if (!bit_at(tb, i)) continue;
    std::array<int, 3> o;
    key_origin(keys[i], o.data());
    order.push_back({o, i});
  }
  std::sort(order.begin(), order.end());
  for (size_t j = 0; j < order.size(); ++j) {
    if (origins) std::memcpy(origins + 3 * j, order[j].first.data(), 12);
    if (values) values[j] = tv[order[j].second];
    if (active) active[j] = bit_at(on, order[j].second) ? 1 : 0;
  }
  return (int64_t)order.size();
}

int64_t vdbgpu_export_active_points(vdbgpu_map* m, float* xyz, int64_t cap) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  if (hipStreamSynchronize(m->cast) != hipSuccess || read_status(m)) return -1;
  const uint32_t n = m->h_st[1].leaf_count;
  std::vector<unsigned long long> keys(n), act((size_t)n * 8);
  if (n && (hipMemcpy(keys.data(), m->leaf_keys, n * sizeof(unsigned long long), hipMemcpyDeviceToHost) ||
            hipMemcpy(act.data(), m->active, (size_t)n * 64, hipMemcpyDeviceToHost))) {
    m->last_error = "export copy failed";
    return -1;
  }
  std::vector<uint32_t> tb;
  if (tile_bits(m, tb, nullptr)) return -1;
  // cbeginValueOn yields an active tile once, at its origin (an inactive one
  // not at all): only voxel 0 of a tile's leaf is listed
  for (uint32_t i = 0; i < n; ++i)
    if (bit_at(tb, i))
      for (int w = 0; w < 8; ++w) act[(size_t)i * 8 + w] = w == 0 ? act[(size_t)i * 8] & 1u : 0;
  int64_t total = 0;
  for (uint64_t w : act) total += __builtin_popcountll(w);
  if (!xyz) return total;
  if (total > cap) return -total;
  // cbeginValueOn order: root children by origin, then the 128^3 nodes and
  // the leaves (and tiles) inside each by offset (x-major), voxels by offset
  std::vector<std::pair<std::array<int, 9>, uint32_t>> order(n);
  for (uint32_t i = 0; i < n; ++i) {
    int o[3];
    key_origin(keys[i], o);
    order[i] = {{o[0] >> 12, o[1] >> 12, o[2] >> 12, o[0] >> 7, o[1] >> 7, o[2] >> 7, o[0] >> 3, o[1] >> 3, o[2] >> 3},
                i};
  }
  std::sort(order.begin(), order.end());
  int64_t k = 0;
  for (const auto& oi : order) {
    int o[3];
    key_origin(keys[oi.second], o);
    for (int v = 0; v < 512; ++v) {
      if (!((act[(size_t)oi.second * 8 + (v >> 6)] >> (v & 63)) & 1)) continue;
      const int c[3] = {o[0] + (v >> 6), o[1] + ((v >> 3) & 7), o[2] + (v & 7)};
      for (int a = 0; a < 3; ++a) {
        double w = (double)c[a] * m->res;  // indexToWorld (ScaleMap::applyMap)
        w += 0.5 * m->res;
        xyz[3 * k + a] = (float)w;
      }
      ++k;
    }
  }
  return total;
}
}  // extern "C"

namespace {

// Layer range [(int)(negative_height / res), (int)(positive_height / res))
// of addArtificialWall (VM:1227-1228).
int wall_layers(vdbgpu_map* m, double negative_height, double positive_height, int* k0, int* k1) {
  const double qn = negative_height / m->res, qp = positive_height / m->res;
  if (!(qn > -2147483649.0 && qn < 2147483648.0) || !(qp > -2147483649.0 && qp < 2147483648.0))
    return fail(m, VDBGPU_BAD_ARGUMENT, "height / resolution is outside the int range");
  *k0 = (int)qn;
  *k1 = (int)qp;
  if ((int64_t)*k1 - *k0 > (1 << 24)) return fail(m, VDBGPU_BAD_ARGUMENT, "more than 2^24 height layers");
  return VDBGPU_OK;
}

// Cast the walls {a, b, k0, k1} over `verts` into art_tmp (empty before);
// *count = its listed leaves. On error art_tmp is empty again.
int build_art(vdbgpu_map* m, const std::vector<double>& verts, const std::vector<int>& walls, uint32_t* count) {
  *count = 0;
  const uint32_t n_walls = (uint32_t)(walls.size() / 4);
  std::vector<uint64_t> prefix(n_walls + 1, 0);
  double ext = 0;
  for (uint32_t w = 0; w < n_walls; ++w) {
    prefix[w + 1] = prefix[w] + (uint64_t)std::max(0, walls[4 * w + 3] - walls[4 * w + 2]);
    for (int c = 0; c < 3; ++c) {
      const double d = std::fabs(verts[3 * (size_t)walls[4 * w + 1] + c] - verts[3 * (size_t)walls[4 * w] + c]) * m->inv;
      if (d > ext) ext = d;  // NaN compares false
    }
  }
  if (prefix[n_walls] == 0) return VDBGPU_OK;
  int r;
  if ((r = ensure_table(m, (uint32_t)std::min<double>(ext + 2.0, (double)kMaxTableN), m->stream))) return r;
  double* d_verts = nullptr;
  int4* d_walls = nullptr;
  uint64_t* d_prefix = nullptr;
  if ((r = upload(m, &d_verts, verts.data(), verts.size())) ||
      (r = upload(m, &d_walls, reinterpret_cast<const int4*>(walls.data()), n_walls)) ||
      (r = upload(m, &d_prefix, prefix.data(), prefix.size())))
    return r;
  r = mark_set(m, m->art_tmp, kMinUpdateCap, [&]() -> int {
    CK(launch_mark_walls(m->stream, d_verts, d_walls, d_prefix, n_walls, prefix[n_walls], m->res, m->inv, m->tbl,
                         m->tbl_n, upd_view(m->art_tmp), &m->d_st[2]));
    return VDBGPU_OK;
  }, count);
  dfree(m, &d_verts, verts.size());
  dfree(m, &d_walls, n_walls);
  dfree(m, &d_prefix, prefix.size());
  return r;
}

// The freshly built art_tmp becomes the artificial grid (the old one, already
// cleared or about to be, becomes art_tmp).
int install_art(vdbgpu_map* m, uint32_t count, std::vector<double>&& verts, std::vector<int>&& walls) {
  std::swap(m->art, m->art_tmp);
  int r;
  if ((r = clear_set(m, m->art_tmp))) return r;
  m->art_count = count;
  m->art_leaves = false;
  m->art_verts = std::move(verts);
  m->art_walls = std::move(walls);
  if (count > m->art_eidx_cap) {
    dfree(m, &m->art_eidx, m->art_eidx_cap);
    m->art_eidx_cap = (uint32_t)next_pow2(count);
    CK(dalloc(m, &m->art_eidx, m->art_eidx_cap));
  }
  return VDBGPU_OK;
}

}  // namespace

extern "C" {

int32_t vdbgpu_add_artificial_areas(vdbgpu_map* m, int32_t n_polygons, const int32_t* counts,
                                    const double* xyz, double negative_height, double positive_height) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (n_polygons < 0 || (n_polygons && !counts)) return fail(m, VDBGPU_BAD_ARGUMENT, "bad polygon arguments");
  int k0, k1, r;
  if ((r = wall_layers(m, negative_height, positive_height, &k0, &k1))) return r;
  // addArtificialPolygon (VM:1198-1207): walls i -> (i + 1) mod size
  std::vector<double> verts;
  std::vector<int> walls;
  int64_t nv = 0;
  for (int32_t p = 0; p < n_polygons; ++p) {
    const int32_t cnt = counts[p];
    if (cnt < 0 || (cnt && !xyz)) return fail(m, VDBGPU_BAD_ARGUMENT, "bad polygon vertex count");
    if (nv + cnt > (1 << 28)) return fail(m, VDBGPU_BAD_ARGUMENT, "too many polygon vertices");
    for (int32_t i = 0; i < cnt && k0 < k1; ++i) {
      walls.insert(walls.end(), {(int)(nv + i), (int)(nv + (i + 1) % cnt), k0, k1});
    }
    nv += cnt;
  }
  verts.assign(xyz ? xyz : nullptr, xyz ? xyz + 3 * nv : nullptr);
  CK(hipStreamSynchronize(m->cast));
  uint32_t count = 0;
  if ((r = build_art(m, verts, walls, &count))) return r;  // the live grid and the map are unchanged
  // restoreMapIntegrity first (VM:1181), then the new grid replaces the old
  if ((r = restore_integrity(m))) {
    (void)clear_set(m, m->art_tmp);
    return r;
  }
  if ((r = install_art(m, count, std::move(verts), std::move(walls)))) return r;
  m->art_present = true;
  return VDBGPU_OK;
}

int32_t vdbgpu_add_artificial_walls(vdbgpu_map* m, int32_t n_walls, const double* xyz, double negative_height,
                                    double positive_height) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (n_walls < 0 || (n_walls && !xyz) || n_walls > (1 << 26)) return fail(m, VDBGPU_BAD_ARGUMENT, "bad wall arguments");
  int k0, k1, r;
  if ((r = wall_layers(m, negative_height, positive_height, &k0, &k1))) return r;
  std::vector<double> verts = m->art_verts;
  std::vector<int> walls = m->art_walls;
  if (verts.size() / 3 + 2 * (size_t)n_walls > (1u << 28)) return fail(m, VDBGPU_BAD_ARGUMENT, "too many walls");
  for (int32_t w = 0; w < n_walls && k0 < k1; ++w) {
    const int a = (int)(verts.size() / 3);
    verts.insert(verts.end(), xyz + 6 * (size_t)w, xyz + 6 * (size_t)w + 6);
    walls.insert(walls.end(), {a, a + 1, k0, k1});
  }
  CK(hipStreamSynchronize(m->cast));
  uint32_t count = 0;
  // the grid is rebuilt from every wall since the last restore (marks are an OR)
  if ((r = build_art(m, verts, walls, &count))) return r;
  return install_art(m, count, std::move(verts), std::move(walls));
}

int32_t vdbgpu_restore_map_integrity(vdbgpu_map* m) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  return restore_integrity(m);
}

// CK for the exports: a HIP error returns -1.
#define CKN(expr)                                                                \
  do {                                                                           \
    hipError_t e_ = (expr);                                                      \
    if (e_ != hipSuccess) {                                                      \
      m->last_error = std::string(#expr) + ": " + hipGetErrorString(e_);         \
      return -1;                                                                 \
    }                                                                            \
  } while (0)

int64_t vdbgpu_export_artificial(vdbgpu_map* m, int32_t* origins, uint64_t* masks, int64_t cap) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  const uint32_t n = m->art_count;
  std::vector<unsigned long long> keys(n), act((size_t)n * 8);
  std::vector<float> none;
  if (n && (origins || masks)) {
    std::vector<uint32_t> list(n);
    std::vector<unsigned long long> all_keys(m->art.cap);
    CKN(hipMemcpyAsync(list.data(), m->art.list, n * sizeof(uint32_t), hipMemcpyDeviceToHost, m->stream));
    CKN(hipMemcpyAsync(all_keys.data(), m->art.keys, m->art.cap * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                      m->stream));
    std::vector<unsigned long long> all_masks((size_t)m->art.cap * 16);
    CKN(hipMemcpyAsync(all_masks.data(), m->art.masks, all_masks.size() * sizeof(unsigned long long),
                      hipMemcpyDeviceToHost, m->stream));
    CKN(hipStreamSynchronize(m->stream));
    for (uint32_t i = 0; i < n; ++i) {
      keys[i] = all_keys[list[i]];
      std::memcpy(&act[(size_t)i * 8], &all_masks[(size_t)list[i] * 16], 64);
    }
  }
  return export_sorted(m, keys, none, act, origins, nullptr, masks, cap, false);
}

// ---------------------------------------------------------- map sections --

int64_t vdbgpu_export_section(vdbgpu_map* m, const int32_t bbox[6], int32_t full_grid, int32_t* origins,
                              float* values, uint64_t* masks, int64_t cap) {
  if (!m || !bbox) return -1;
  Guard g(m);
  if (g.rc) return -1;
  CKN(hipStreamSynchronize(m->cast));
  if (read_status(m)) return -1;
  const uint32_t L = m->h_st[1].leaf_count;
  std::vector<unsigned long long> keys, act;
  std::vector<float> vals;
  if (L) {
    int* d_box = nullptr;
    uint32_t* list = nullptr;
    if (upload(m, &d_box, bbox, 6)) return -1;
    CKN(dalloc(m, &list, L));
    CKN(hipMemsetAsync(&m->d_st[2], 0, sizeof(Status), m->stream));
    CKN(launch_section_pick(m->stream, map_view(m), L, d_box, full_grid != 0, list, &m->d_st[2].entries,
                            tile_view(m)));
    if (read_status(m)) return -1;
    const uint32_t n = m->h_st[2].entries;
    const bool want = origins || values || masks;
    if (n && want && n <= cap) {
      unsigned long long *dk = nullptr, *da = nullptr;
      float* dv = nullptr;
      CKN(dalloc(m, &dk, n));
      CKN(dalloc(m, &da, (size_t)n * 8));
      CKN(dalloc(m, &dv, (size_t)n * 512));
      CKN(launch_section_gather(m->stream, map_view(m), list, n, d_box, full_grid != 0, dk, dv, da));
      keys.resize(n);
      act.resize((size_t)n * 8);
      vals.resize((size_t)n * 512);
      CKN(hipMemcpyAsync(keys.data(), dk, n * 8, hipMemcpyDeviceToHost, m->stream));
      CKN(hipMemcpyAsync(act.data(), da, (size_t)n * 64, hipMemcpyDeviceToHost, m->stream));
      CKN(hipMemcpyAsync(vals.data(), dv, (size_t)n * 2048, hipMemcpyDeviceToHost, m->stream));
      CKN(hipStreamSynchronize(m->stream));
      dfree(m, &dk, n);
      dfree(m, &da, (size_t)n * 8);
      dfree(m, &dv, (size_t)n * 512);
    } else {
      keys.resize(n);  // the count is all that is returned
    }
    dfree(m, &d_box, 6);
    dfree(m, &list, L);
  }
  return export_sorted(m, keys, vals, act, origins, values, masks, cap, true);
}

namespace {

// Upsert the map leaves of n section leaves (device arrays) and hand back
// their pool indices in *eidx_out (device, n entries; freed by the caller).
int section_upsert(vdbgpu_map* m, const unsigned long long* keys, const uint8_t* need, uint32_t n,
                   uint16_t* first, uint32_t** eidx_out) {
  int r;
  if ((r = ensure_map(m, m->leaf_count_ub + n))) return r;
  CK(dalloc(m, eidx_out, n));
  CK(launch_upsert_leaves(m->stream, keys, need, n, map_view(m), *eidx_out, first, &m->d_st[1]));
  m->leaf_count_ub += n;
  ledger_add(m, n);
  m->snap_dirty = true;
  return VDBGPU_OK;
}

}  // namespace

int32_t vdbgpu_apply_section(vdbgpu_map* m, int64_t n, const int32_t* origins, const float* values,
                             const uint64_t* masks) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (n < 0 || n >= (1 << 28) || (n && (!origins || !values || !masks)))
    return fail(m, VDBGPU_BAD_ARGUMENT, "bad section arguments");
  int r;
  if ((r = check_origins(m, n, origins))) return r;
  if (n == 0) return VDBGPU_OK;
  CK(hipStreamSynchronize(m->cast));
  const uint32_t N = (uint32_t)n;
  const uint32_t scap = (uint32_t)std::max<uint64_t>(1024, next_pow2(2 * (uint64_t)N));
  int* d_o = nullptr;
  float* d_v = nullptr;
  unsigned long long *d_a = nullptr, *keys = nullptr, *sets = nullptr, *lists = nullptr;
  uint8_t* need = nullptr;
  uint16_t* first = nullptr;
  uint32_t* eidx = nullptr;
  if ((r = upload(m, &d_o, origins, (size_t)N * 3)) || (r = upload(m, &d_v, values, (size_t)N * 512)) ||
      (r = upload(m, &d_a, reinterpret_cast<const unsigned long long*>(masks), (size_t)N * 8)))
    return r;
  CK(dalloc(m, &keys, N));
  CK(dalloc(m, &need, N));
  CK(dalloc(m, &first, N));
  CK(dalloc(m, &sets, (size_t)scap * 3));  // leaf, 128^3 and 4096^3 key sets
  CK(dalloc(m, &lists, (size_t)N * 2));    // new 128^3 and 4096^3 node keys
  CK(launch_fill_u64(m->stream, sets, (uint64_t)scap * 3, kEmpty));
  CK(hipMemsetAsync(&m->d_st[2], 0, sizeof(Status), m->stream));
  unsigned int* counters = &m->d_st[2].segs;  // [0] 128^3 nodes, [1] 4096^3 nodes
  CK(launch_section_prep(m->stream, d_o, d_v, d_a, N, keys, need, first, sets, scap - 1, sets + scap, scap - 1,
                         lists, sets + 2 * (size_t)scap, scap - 1, lists + N, counters));
  if ((r = read_status(m))) return r;
  const uint32_t c128 = m->h_st[2].segs, c4096 = m->h_st[2].items;
  if ((r = section_upsert(m, keys, need, N, first, &eidx))) return r;
  CK(launch_section_write(m->stream, d_v, d_a, eidx, first, N, map_view(m)));
  // the inactive tiles of the section's internal nodes (VM:1034 cbeginValueAll)
  CK(launch_section_tiles(m->stream, lists, c128, lists + N, c4096, sets, scap - 1, sets + scap, scap - 1,
                          map_view(m)));
  if ((r = check_tiles(m)) || (r = read_status(m))) return r;
  m->touched = m->touched || m->h_st[1].leaf_count > 0;
  dfree(m, &d_o, (size_t)N * 3);
  dfree(m, &d_v, (size_t)N * 512);
  dfree(m, &d_a, (size_t)N * 8);
  dfree(m, &keys, N);
  dfree(m, &need, N);
  dfree(m, &first, N);
  dfree(m, &sets, (size_t)scap * 3);
  dfree(m, &lists, (size_t)N * 2);
  dfree(m, &eidx, N);
  return VDBGPU_OK;
}

int32_t vdbgpu_apply_section_update(vdbgpu_map* m, const int32_t bbox[6], int64_t n, const int32_t* origins,
                                    const uint64_t* masks) {
  if (!m) return VDBGPU_BAD_ARGUMENT;
  Guard g(m);
  if (g.rc) return g.rc;
  if (!bbox || n < 0 || n >= (1 << 28) || (n && (!origins || !masks)))
    return fail(m, VDBGPU_BAD_ARGUMENT, "bad section arguments");
  int r;
  if ((r = check_origins(m, n, origins))) return r;
  CK(hipStreamSynchronize(m->cast));
  if ((r = read_status(m))) return r;
  const uint32_t L = m->h_st[1].leaf_count;
  int* d_box = nullptr;
  if ((r = upload(m, &d_box, bbox, 6))) return r;
  CK(launch_section_clear(m->stream, map_view(m), L, d_box, tile_view(m)));
  if (n) {
    const uint32_t N = (uint32_t)n;
    int* d_o = nullptr;
    unsigned long long *d_a = nullptr, *keys = nullptr;
    uint8_t* need = nullptr;
    uint16_t* first = nullptr;
    uint32_t* eidx = nullptr;
    if ((r = upload(m, &d_o, origins, (size_t)N * 3)) ||
        (r = upload(m, &d_a, reinterpret_cast<const unsigned long long*>(masks), (size_t)N * 8)))
      return r;
    CK(dalloc(m, &keys, N));
    CK(dalloc(m, &need, N));
    CK(dalloc(m, &first, N));
    CK(launch_section_prep(m->stream, d_o, nullptr, d_a, N, keys, need, first, nullptr, 0, nullptr, 0, nullptr,
                           nullptr, 0, nullptr, nullptr));
    if ((r = section_upsert(m, keys, need, N, nullptr, &eidx))) return r;
    CK(launch_section_or(m->stream, d_a, eidx, N, map_view(m)));
    if ((r = read_status(m))) return r;
    m->touched = m->touched || m->h_st[1].leaf_count > 0;
    dfree(m, &d_o, (size_t)N * 3);
    dfree(m, &d_a, (size_t)N * 8);
    dfree(m, &keys, N);
    dfree(m, &need, N);
    dfree(m, &first, N);
    dfree(m, &eidx, N);
  }
  if ((r = check_tiles(m))) return r;
  CK(hipStreamSynchronize(m->stream));
  dfree(m, &d_box, 6);
  return VDBGPU_OK;
}

// ------------------------------------------------------------ morphology --

}  // extern "C"

namespace {

constexpr uint64_t kMorphMaxLeaves = 1ull << 26;

// Device leaf table of one morphology call (morph_kernels.hip): entries
// 0 .. count-1 by index (origin, mask ping-pong pair, keep flag, neighbour
// indices) and the key hash over them.
struct MorphTable {
  vdbgpu_map* m;
  uint32_t count = 0, ecap = 0, hcap = 0;
  int* org = nullptr;
  unsigned long long *ma = nullptr, *mb = nullptr, *keys = nullptr;
  uint32_t *hidx = nullptr, *nbr = nullptr;
  uint8_t* keep = nullptr;
  unsigned int* added = nullptr;
  explicit MorphTable(vdbgpu_map* mm) : m(mm) {}
  ~MorphTable() {
    (void)hipStreamSynchronize(m->stream);
    free_entries();
    dfree(m, &keys, hcap);
    dfree(m, &hidx, hcap);
    dfree(m, &added, 1);
  }
  void free_entries() {
    dfree(m, &org, (size_t)ecap * 3);
    dfree(m, &ma, (size_t)ecap * 8);
    dfree(m, &mb, (size_t)ecap * 8);
    dfree(m, &nbr, (size_t)ecap * 27);
    dfree(m, &keep, ecap);
  }
  MorphView view() const { return MorphView{org, keys, hidx, hcap - 1, nbr}; }

  // Room for `need` entries; the first `count` (origin, mask, keep) move along.
  int reserve(uint32_t need) {
    if (need <= ecap) return VDBGPU_OK;
    int* o2 = nullptr;
    unsigned long long *a2 = nullptr, *b2 = nullptr;
    uint32_t* n2 = nullptr;
    uint8_t* k2 = nullptr;
    CK(dalloc(m, &o2, (size_t)need * 3));
    CK(dalloc(m, &a2, (size_t)need * 8));
    CK(dalloc(m, &b2, (size_t)need * 8));
    CK(dalloc(m, &n2, (size_t)need * 27));
    CK(dalloc(m, &k2, need));
    if (count) {
      CK(hipMemcpyAsync(o2, org, (size_t)count * 12, hipMemcpyDeviceToDevice, m->stream));
      CK(hipMemcpyAsync(a2, ma, (size_t)count * 64, hipMemcpyDeviceToDevice, m->stream));
      CK(hipMemcpyAsync(k2, keep, count, hipMemcpyDeviceToDevice, m->stream));
    }
    CK(hipStreamSynchronize(m->stream));
    free_entries();
    org = o2;
    ma = a2;
    mb = b2;
    nbr = n2;
    keep = k2;
    ecap = need;
    return VDBGPU_OK;
  }

  // Hash of the `count` entries, sized for ecap of them (at most half full).
  int rehash() {
    const uint32_t want = (uint32_t)std::max<uint64_t>(1024, next_pow2(2 * (uint64_t)ecap));
    if (want != hcap) {
      CK(hipStreamSynchronize(m->stream));
      dfree(m, &keys, hcap);
      dfree(m, &hidx, hcap);
      CK(dalloc(m, &keys, want));
      CK(dalloc(m, &hidx, want));
      hcap = want;
    }
    CK(launch_fill_u64(m->stream, keys, hcap, kEmpty));
    CK(launch_morph_seed(m->stream, view(), count));
    return VDBGPU_OK;
  }

  // `iterations` dilation steps: before each run of up to 8 steps the table
  // takes in the 26 neighbours of every leaf with an active voxel.
  int dilate(int32_t iterations) {
    for (int32_t left = iterations; left > 0;) {
      const uint64_t bound = (uint64_t)count * 27;
      if (bound > kMorphMaxLeaves)
        return fail(m, VDBGPU_UNSUPPORTED, "morphology would exceed 2^26 leaves");
      int r;
      if ((r = reserve((uint32_t)bound)) || (r = rehash())) return r;
      CK(hipMemsetAsync(added, 0, sizeof(unsigned int), m->stream));
      CK(launch_morph_grow(m->stream, view(), ma, keep, count, ecap, added));
      unsigned int h_added = 0;
      CK(hipMemcpyAsync(&h_added, added, sizeof h_added, hipMemcpyDeviceToHost, m->stream));
      CK(hipStreamSynchronize(m->stream));
      count += h_added;
      CK(launch_morph_nbrs(m->stream, view(), count));
      const int32_t steps = std::min<int32_t>(8, left);
      for (int32_t i = 0; i < steps; ++i) {
        CK(launch_morph_step(m->stream, ma, mb, nbr, count, false));
        std::swap(ma, mb);
      }
      left -= steps;
    }
    CK(launch_morph_keep(m->stream, ma, keep, count));
    return VDBGPU_OK;
  }

  int erode(int32_t iterations, bool nbrs_current) {
    if (!nbrs_current) {
      int r;
      if ((r = rehash())) return r;
      CK(launch_morph_nbrs(m->stream, view(), count));
    }
    for (int32_t i = 0; i < iterations; ++i) {
      CK(launch_morph_step(m->stream, ma, mb, nbr, count, true));
      std::swap(ma, mb);
    }
    return VDBGPU_OK;
  }
};

}  // namespace

extern "C" {

int64_t vdbgpu_morphology(vdbgpu_map* m, int32_t op, int32_t iterations, int64_t n, const int32_t* origins,
                          const uint64_t* masks, int32_t* out_origins, uint64_t* out_masks, int64_t cap) {
  if (!m) return -1;
  Guard g(m);
  if (g.rc) return -1;
  if (op < 0 || op > 3 || n < 0 || (uint64_t)n > kMorphMaxLeaves || (n && (!origins || !masks))) {
    fail(m, VDBGPU_BAD_ARGUMENT, "bad morphology arguments");
    return -1;
  }
  if (check_origins(m, n, origins)) return -1;
  // dilation moves the leaf set by at most one leaf per 8 steps; it must stay
  // inside the leaf key range
  const int64_t margin = op == 1 || iterations <= 0 ? 0 : 8 * (((int64_t)iterations + 7) / 8);
  for (int64_t i = 0; i < 3 * n; ++i)
    if (origins[i] - margin < -kCoordLimit || origins[i] + 8 + margin > kCoordLimit) {
      fail(m, VDBGPU_OUT_OF_RANGE, "dilation would leave the voxel range [-2^23, 2^23)");
      return -1;
    }
  std::vector<int32_t> ro(origins, origins + 3 * n);
  std::vector<uint64_t> rm(masks, masks + 8 * n);
  if (n && iterations > 0) {
    MorphTable t(m);
    const uint32_t N = (uint32_t)n;
    if (t.reserve(N) || dalloc(m, &t.added, 1) != hipSuccess) return -1;
    CKN(hipMemcpyAsync(t.org, origins, (size_t)N * 12, hipMemcpyHostToDevice, m->stream));
    CKN(hipMemcpyAsync(t.ma, masks, (size_t)N * 64, hipMemcpyHostToDevice, m->stream));
    CKN(hipMemsetAsync(t.keep, 0, N, m->stream));
    t.count = N;
    const bool dilate_first = op == 0 || op == 2;
    for (int p = 0; p < (op >= 2 ? 2 : 1); ++p) {
      const bool dilate = (p == 0) == dilate_first;
      if (dilate ? t.dilate(iterations) : t.erode(iterations, p == 1)) return -1;
    }
    const uint32_t C = t.count;
    std::vector<int32_t> o((size_t)C * 3);
    std::vector<uint64_t> a((size_t)C * 8);
    std::vector<uint8_t> k(C);
    CKN(hipMemcpyAsync(o.data(), t.org, (size_t)C * 12, hipMemcpyDeviceToHost, m->stream));
    CKN(hipMemcpyAsync(a.data(), t.ma, (size_t)C * 64, hipMemcpyDeviceToHost, m->stream));
    CKN(hipMemcpyAsync(k.data(), t.keep, C, hipMemcpyDeviceToHost, m->stream));
    CKN(hipStreamSynchronize(m->stream));
    std::memcpy(rm.data(), a.data(), (size_t)N * 64);
    // the leaves dilation created, sorted by origin
    std::vector<std::pair<std::array<int32_t, 3>, uint32_t>> fresh;
    for (uint32_t i = N; i < C; ++i)
      if (k[i]) fresh.push_back({{o[3 * (size_t)i], o[3 * (size_t)i + 1], o[3 * (size_t)i + 2]}, i});
    std::sort(fresh.begin(), fresh.end());
    for (const auto& f : fresh) {
      ro.insert(ro.end(), f.first.begin(), f.first.end());
      rm.insert(rm.end(), a.begin() + (size_t)f.second * 8, a.begin() + (size_t)f.second * 8 + 8);
    }
  }
  const int64_t total = (int64_t)(ro.size() / 3);
  if (!out_origins && !out_masks) return total;
  if (total > cap) return -total;
  if (out_origins) std::memcpy(out_origins, ro.data(), ro.size() * sizeof(int32_t));
  if (out_masks) std::memcpy(out_masks, rm.data(), rm.size() * sizeof(uint64_t));
  return total;
}

// createIndexBoundingBox (VM:811-874): corners of [min, max] as float, each
// transformed by the row-major 4x4 matrix in double (((m0 p0 + m1 p1) + m2 p2)
// + m3, rounded to float: pcl::transformPointCloud), float min / max,
// Coord::floor(value * 1/res).
int32_t vdbgpu_index_bbox(double resolution, const double min_boundary[3], const double max_boundary[3],
                          const double tf[16], int32_t out[6]) {
  if (!(resolution > 0) || !min_boundary || !max_boundary || !tf || !out) return VDBGPU_BAD_ARGUMENT;
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int c = 0; c < 8; ++c) {
    const double p[3] = {(double)(float)((c & 4) ? max_boundary[0] : min_boundary[0]),
                         (double)(float)((c & 2) ? max_boundary[1] : min_boundary[1]),
                         (double)(float)((c & 1) ? max_boundary[2] : min_boundary[2])};
    for (int r = 0; r < 3; ++r) {
      const double acc = tf[4 * r] * p[0] + tf[4 * r + 1] * p[1] + tf[4 * r + 2] * p[2] + tf[4 * r + 3];
      const float q = (float)acc;
      lo[r] = std::min(lo[r], q);
      hi[r] = std::max(hi[r], q);
    }
  }
  const double inv = 1.0 / resolution;
  for (int r = 0; r < 3; ++r) {
    const double a = std::floor((double)lo[r] * inv), b = std::floor((double)hi[r] * inv);
    if (!(a >= -2147483648.0 && a < 2147483648.0 && b >= -2147483648.0 && b < 2147483648.0))
      return VDBGPU_OUT_OF_RANGE;
    out[r] = (int32_t)a;
    out[3 + r] = (int32_t)b;
  }
  return VDBGPU_OK;
}

}  // extern "C"
