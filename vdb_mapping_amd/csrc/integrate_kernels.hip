// gfx950 kernels of the per-scan integration path.
//
//   k_prep        one lane per point: NaN skip, max-range truncation and rounded
//                 world->index (VDBMapping.hpp:499-537, :612-631), visit count
//                 1+|dx|+|dy|+|dz| per ray and its block-local prefix
//   k_scan_blocks exclusive scan of the 256-ray block sums
//   k_traverse    load-balanced DDA (VDBMapping.hpp:550-566): the global visit
//                 range is cut into equal chunks, one per lane; a lane enters
//                 its first ray mid-way at the exact DDA state found from the
//                 crossing table, then walks; visits are OR-ed into the
//                 source's update set at leaf-word granularity
//   k_compact     list of occupied update-set slots + visited-voxel count
//   k_map_insert  one lane per update leaf: find or create the map leaf
//   k_apply       one wave per update leaf, eight voxels per lane: clamped
//                 log-odds update of the visited voxels
//                 (OccupancyVDBMapping.hpp:92-117), two float4 per lane
//
// Build with -ffp-contract=off: every double operation must round on its own,
// exactly like the reference's x86-64 build.
#include <float.h>

#include "dda_math.h"
#include "device_common.h"
#include "kernels.h"
#include "rt_common.h"
#include "upd_common.h"

namespace vdbgpu {

// ---------------------------------------------------------------- helpers --

// Workgroup-local aggregation table in LDS: leaf key -> 8 visited words.
constexpr int kLdsEntries = 512;  // 36 KiB per workgroup
constexpr int kLdsProbe = 16;

__device__ __forceinline__ int lds_slot(unsigned long long* lkey, uint64_t key) {
  uint32_t h = (uint32_t)(fmix64(key) >> 40) & (kLdsEntries - 1);
  for (int p = 0; p < kLdsProbe; ++p) {
    const unsigned long long k =
        __hip_atomic_load(&lkey[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return (int)h;
    if (k == kEmpty) {
      const unsigned long long prev = atomicCAS(&lkey[h], kEmpty, (unsigned long long)key);
      if (prev == kEmpty || prev == key) return (int)h;
    }
    h = (h + 1) & (kLdsEntries - 1);
  }
  return -1;
}

// Per-lane accumulator of visited bits for one (leaf, mask word); a finished
// run is OR-ed into the workgroup's LDS entry of the leaf, or straight into
// the global update set when the LDS table is full.
struct Run {
  int lx, ly, lz;
  uint32_t lb;   // scan of the current leaf entry (batched keys)
  int ls;        // LDS entry of the current leaf, -1 = global
  uint32_t gs;   // global slot when ls < 0
  int word;
  uint64_t bits;
  __device__ void init() {
    lx = INT_MIN;
    ly = lz = 0;
    lb = 0;
    ls = -1;
    gs = kNone;
    word = 0;
    bits = 0;
  }
  __device__ __forceinline__ void flush(const UpdView& u, unsigned long long* lvis) {
    if (bits == 0) return;
    if (ls >= 0)
      atomicOr(&lvis[ls * 8 + word], (unsigned long long)bits);
    else
      upd_or(u, gs, word, bits);
    bits = 0;
  }
  __device__ __forceinline__ void add(const UpdView& u, unsigned int* ovf, unsigned long long* lkey,
                                      unsigned long long* lvis, int x, int y, int z, uint32_t scan,
                                      bool batched) {
    const int w = x & 7;
    const uint64_t b = 1ull << (((y & 7) << 3) | (z & 7));
    const int nx = x >> 3, ny = y >> 3, nz = z >> 3;
    if (nx == lx && ny == ly && nz == lz && scan == lb) {
      if (w == word) {
        bits |= b;
        return;
      }
      flush(u, lvis);
    } else {
      flush(u, lvis);
      lx = nx;
      ly = ny;
      lz = nz;
      lb = scan;
      const uint64_t key = update_key(x, y, z, scan, batched);
      ls = lds_slot(lkey, key);
      if (ls < 0) gs = upd_slot(u, key, ovf);
    }
    word = w;
    bits = b;
  }
};

// ---------------------------------------------------------------- kernels --

// Row n of the crossing table: next_0 = 0.5 * fl(1/n), next_j = fl(next_{j-1} + fl(1/n)),
// which is what math::DDA::init/step produce for a ray from a voxel centre
// with integer direction component +-n (VDBMapping.hpp:555-564).
__global__ void k_build_table(double* __restrict__ tbl, uint32_t n_lo, uint32_t n_hi) {
  const uint32_t n = n_lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (n > n_hi) return;
  const double d = 1.0 / (double)n;
  double t = 0.5 * d;
  double* row = tbl + (((uint64_t)n * (n - 1)) >> 1);
  for (uint32_t j = 0; j < n; ++j) {
    row[j] = t;
    t = t + d;
  }
}

template <bool kVec4>
__global__ __launch_bounds__(kPrepBlock) void k_prep(const float* __restrict__ pts,
                                                     int64_t stride_f, ScanInfo* __restrict__ scans,
                                                     const BatchCtl* __restrict__ ctl, double res,
                                                     double inv, uint32_t tbl_n,
                                                     int4* __restrict__ rayinfo,
                                                     uint32_t* __restrict__ pre,
                                                     uint32_t* __restrict__ bsum, Status* st) {
  __shared__ uint32_t wsum[kPrepBlock / 64];
  const uint32_t n = ctl->n_rays;
  if (blockIdx.x * kPrepBlock >= n) return;  // whole workgroup beyond the rays
  const uint32_t nscans = ctl->n_scans;
  const int lim = ctl->batched ? kBatchCoordLimit : kCoordLimit;
  const uint32_t i = blockIdx.x * kPrepBlock + threadIdx.x;
  uint32_t cnt = 0;
  bool valid = false;
  if (i < n) {
    // scan of ray i: the last scan whose first ray is <= i
    uint32_t b = 0;
    if (nscans > 1) {
      uint32_t lo = 0, hi = nscans - 1;
      while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (scans[mid].ray0 <= i)
          lo = mid;
        else
          hi = mid - 1;
      }
      b = lo;
    }
    const double ox = scans[b].o[0], oy = scans[b].o[1], oz = scans[b].o[2], R = scans[b].R;
    double px, py, pz;
    if (kVec4 && i + 1 < n) {  // the last point may end 4 bytes before a 16-byte boundary
      const float4 p = reinterpret_cast<const float4*>(pts)[i];
      px = p.x;
      py = p.y;
      pz = p.z;
    } else {
      const float* p = pts + (int64_t)i * stride_f;
      px = p[0];
      py = p[1];
      pz = p[2];
    }
    int4 ri = make_int4(0, 0, 0, (int)(b << kScanShift));
    // a scan with max range <= 0 casts nothing (accumulateUpdate's gate, VDBMapping.hpp:331)
    valid = R > 0.0 && !(isnan(px) || isnan(py) || isnan(pz) || isnan(ox) || isnan(oy) || isnan(oz));
    const int sx = world_to_index1(ox, res, inv), sy = world_to_index1(oy, res, inv),
              sz = world_to_index1(oz, res, inv);
    if (valid) {
      bool trunc = false;
      const double dx = px - ox, dy = py - oy, dz = pz - oz;
      const double len = sqrt(dx * dx + dy * dy + dz * dz);
      if (R > 0.0 && len > R) {
        px = ox + (dx / len) * R;
        py = oy + (dy / len) * R;
        pz = oz + (dz / len) * R;
        trunc = true;
      }
      const int ex = world_to_index1(px, res, inv), ey = world_to_index1(py, res, inv),
                ez = world_to_index1(pz, res, inv);
      if (!(sx >= -lim && sx < lim && sy >= -lim && sy < lim && sz >= -lim && sz < lim &&
            ex >= -lim && ex < lim && ey >= -lim && ey < lim && ez >= -lim && ez < lim)) {
        atomicOr(&st->error, kErrRange);
      } else {
        const uint32_t nx = (uint32_t)abs(ex - sx), ny = (uint32_t)abs(ey - sy),
                       nz = (uint32_t)abs(ez - sz);
        if (nx > tbl_n || ny > tbl_n || nz > tbl_n) {
          atomicOr(&st->error, kErrTable);
        } else {
          cnt = (nx | ny | nz) ? 1u + nx + ny + nz : 0u;
          ri = make_int4(ex, ey, ez,
                         (int)(cnt | (b << kScanShift) | kValidBit | (trunc ? 0u : kHitBit)));
        }
      }
    }
    rayinfo[i] = ri;
    if (i == scans[b].ray0) {
      scans[b].s[0] = sx;
      scans[b].s[1] = sy;
      scans[b].s[2] = sz;
    }
  }
  // valid-ray count, one atomic per wave
  const uint64_t vm = __ballot(valid);
  if (lane_id() == 0 && vm) atomicAdd(&st->rays, (unsigned long long)__popcll(vm));
  // block-inclusive prefix of the visit counts
  uint32_t v = cnt;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t t = __shfl_up(v, o, 64);
    if (lane_id() >= o) v += t;
  }
  const int wv = threadIdx.x >> 6;
  if (lane_id() == 63) wsum[wv] = v;
  __syncthreads();
  uint32_t add = 0;
  for (int w = 0; w < wv; ++w) add += wsum[w];
  v += add;
  if (i < n) pre[i] = v;
  if (threadIdx.x == kPrepBlock - 1) bsum[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void k_scan_blocks(const uint32_t* __restrict__ bsum,
                                                     uint64_t* __restrict__ boff,
                                                     const BatchCtl* __restrict__ ctl,
                                                     Status* st) {
  __shared__ unsigned long long wsum[4];
  const uint32_t nb = ctl->nb;
  unsigned long long carry = 0;
  for (uint32_t base = 0; base < nb; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const unsigned long long x = i < nb ? bsum[i] : 0ull;
    unsigned long long v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const unsigned long long t = __shfl_up(v, o, 64);
      if (lane_id() >= o) v += t;
    }
    const int wv = threadIdx.x >> 6;
    if (lane_id() == 63) wsum[wv] = v;
    __syncthreads();
    unsigned long long add = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wv) add += wsum[w];
      tot += wsum[w];
    }
    if (i < nb) boff[i] = carry + add + v - x;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) st->total = carry;
}

// Global visit offset of each ray and, for every chunk of `chunk` visits,
// the ray holding the chunk's first visit.
__global__ __launch_bounds__(256) void k_ray_offsets(const int4* __restrict__ rayinfo,
                                                     const uint32_t* __restrict__ pre,
                                                     const uint64_t* __restrict__ boff,
                                                     const BatchCtl* __restrict__ ctl,
                                                     uint32_t chunk, uint64_t* __restrict__ roff,
                                                     uint32_t* __restrict__ first_ray,
                                                     uint64_t max_chunks, Status* st) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ctl->n_rays) return;
  const uint32_t cnt = (uint32_t)rayinfo[i].w & kCntMask;
  const uint64_t off = boff[i / kPrepBlock] + pre[i] - cnt;
  roff[i] = off;
  if (cnt == 0) return;
  const uint64_t c0 = (off + chunk - 1) / chunk, c1 = (off + cnt + chunk - 1) / chunk;
  if (c1 > max_chunks) {
    atomicOr(&st->error, kErrTable);
    return;
  }
  for (uint64_t c = c0; c < c1; ++c) first_ray[c] = i;
}

__global__ __launch_bounds__(256) void k_traverse(const int4* __restrict__ rayinfo,
                                                  const uint64_t* __restrict__ roff,
                                                  const uint32_t* __restrict__ first_ray,
                                                  const ScanInfo* __restrict__ scans,
                                                  const BatchCtl* __restrict__ ctl,
                                                  const double* __restrict__ tbl, UpdView u,
                                                  Status* st, uint32_t chunk) {
  __shared__ unsigned long long lkey[kLdsEntries];
  __shared__ unsigned long long lvis[kLdsEntries * 8];
  __shared__ unsigned int s_stop;
  // every exit before the last barrier must be taken by the whole workgroup
  if (threadIdx.x == 0)
    s_stop = __hip_atomic_load(&st->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  if (s_stop) return;
  const uint64_t S = st->total;
  const uint32_t n = ctl->n_rays;
  const bool batched = ctl->batched != 0;
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned int* ovf = &st->overflow;
  for (int e = threadIdx.x; e < kLdsEntries; e += blockDim.x) {
    lkey[e] = kEmpty;
#pragma unroll
    for (int w = 0; w < 8; ++w) lvis[e * 8 + w] = 0;
  }

  // Endpoints of rays inside max range: setValueOn(end, true) (VDBMapping.hpp:533-536).
  for (uint64_t i = gid; i < n; i += G) {
    const int4 r = rayinfo[i];
    if ((uint32_t)r.w & kHitBit) {
      const uint32_t sc = ((uint32_t)r.w >> kScanShift) & kScanMask;
      const uint32_t slot = upd_slot(u, update_key(r.x, r.y, r.z, sc, batched), ovf);
      const int w = r.x & 7;
      const uint64_t b = 1ull << (((r.y & 7) << 3) | (r.z & 7));
      upd_or(u, slot, w, b);
      upd_or(u, slot, 8 + w, b);
    }
  }
  __syncthreads();

  // Workgroup iterations: blockDim.x consecutive chunks of `chunk` visits.
  for (uint64_t it = blockIdx.x;; it += gridDim.x) {
    const uint64_t cfirst = it * blockDim.x;
    if (cfirst * chunk >= S) break;  // same for every thread of the workgroup
    if (threadIdx.x == 0)
      s_stop = __hip_atomic_load(ovf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    if (s_stop) break;  // full update set: the host grows it and re-runs

    const uint64_t c = cfirst + threadIdx.x;
    if (c * chunk < S) {
      uint64_t g = c * chunk;
      const uint64_t gend = min(g + chunk, S);
      uint32_t i = first_ray[c];
      uint32_t k = (uint32_t)(g - roff[i]);
      Run run;
      run.init();
      uint32_t cur_scan = kNone;
      int sx = 0, sy = 0, sz = 0;
      while (g < gend && i < n) {
        const int4 r = rayinfo[i];
        const uint32_t cnt = (uint32_t)r.w & kCntMask;
        if (k >= cnt) {
          ++i;
          k = 0;
          continue;
        }
        const uint32_t m = (uint32_t)min((uint64_t)(cnt - k), gend - g);
        const uint32_t sc = ((uint32_t)r.w >> kScanShift) & kScanMask;
        if (sc != cur_scan) {
          cur_scan = sc;
          sx = scans[sc].s[0];
          sy = scans[sc].s[1];
          sz = scans[sc].s[2];
        }
        const uint32_t nn[3] = {(uint32_t)abs(r.x - sx), (uint32_t)abs(r.y - sy),
                                (uint32_t)abs(r.z - sz)};
        const int stx = r.x > sx ? 1 : (r.x < sx ? -1 : 0);
        const int sty = r.y > sy ? 1 : (r.y < sy ? -1 : 0);
        const int stz = r.z > sz ? 1 : (r.z < sz ? -1 : 0);
        uint32_t cc[3] = {0, 0, 0};
        if (k > 0 && !dda_enter(tbl, nn, k, cc)) {
          atomicOr(&st->error, kErrEntry);
          break;
        }
        IntDda d;
        d.init(nn, cc);
        int vx = sx + stx * (int)cc[0], vy = sy + sty * (int)cc[1], vz = sz + stz * (int)cc[2];
        for (uint32_t v = 0;;) {
          run.add(u, ovf, lkey, lvis, vx, vy, vz, sc, batched);
          if (++v == m) break;
          const int a = d.next_axis(tbl);
          d.advance(a);
          vx += a == 0 ? stx : 0;
          vy += a == 1 ? sty : 0;
          vz += a == 2 ? stz : 0;
        }
        g += m;
        ++i;
        k = 0;
      }
      run.flush(u, lvis);
    }
    __syncthreads();
    // one global probe per distinct leaf, one atomic per new visited word
    for (int e = threadIdx.x; e < kLdsEntries; e += blockDim.x) {
      const uint64_t key = lkey[e];
      if (key == kEmpty) continue;
      const uint32_t slot = upd_slot(u, key, ovf);
#pragma unroll
      for (int w = 0; w < 8; ++w) {
        const uint64_t v = lvis[e * 8 + w];
        if (v) {
          upd_or(u, slot, w, v);
          lvis[e * 8 + w] = 0;
        }
      }
      lkey[e] = kEmpty;
    }
    __syncthreads();
  }
}

// Fast-mode raycast (castRayIntoGridFast, VDBMapping.hpp:577-602), one lane
// per ray: Ray(s + 0.5, e - s, 0, 1) clipped to the snapshot's padded active
// bbox (result ignored, VDBMapping.hpp:590), VolumeHDDA::hits over the
// topology snapshot (4096 -> 128 -> 8: every run of existing leaves, as a time
// span), and for every span a voxel DDA over [t0, t1] that marks the voxels
// active in the live map as visited. No e == s guard: a zero-length ray
// probes its origin voxel. Endpoints inside max range are hits as in the
// normal path (VDBMapping.hpp:533-536). cast == 0 (the map is empty,
// VDBMapping.hpp:495, :522): endpoints only.
__global__ __launch_bounds__(256) void k_fast_cast(const int4* __restrict__ rayinfo,
                                                   const ScanInfo* __restrict__ scans,
                                                   const BatchCtl* __restrict__ ctl, RtView rt,
                                                   MapView mp, UpdView u, uint32_t cast,
                                                   Status* st) {
  const uint32_t n = ctl->n_rays;
  unsigned int* ovf = &st->overflow;
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += G) {
    const int4 r = rayinfo[i];
    const uint32_t wbits = (uint32_t)r.w;
    if (!(wbits & kValidBit)) continue;
    if (wbits & kHitBit) mark_voxel(u, ovf, r.x, r.y, r.z, true);
    if (!cast) continue;
    const uint32_t sc = (wbits >> kScanShift) & kScanMask;
    const int s[3] = {scans[sc].s[0], scans[sc].s[1], scans[sc].s[2]};
    const int e[3] = {r.x, r.y, r.z};
    double eye[3], dir[3], idir[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      eye[a] = (double)s[a] + 0.5;
      dir[a] = (double)e[a] - (double)s[a];
      idir[a] = 1.0 / dir[a];
    }
    // every DDA level takes at most |dir|_1 / dim + 3 steps over t in [0, 1]
    int budget = 4 * ((int)(fabs(dir[0]) + fabs(dir[1]) + fabs(dir[2])) + 8);
    double rt0 = 0.0, rt1 = 1.0;
    rt_clip(rt, eye, idir, rt0, rt1);
    ActiveProbe probe;
    auto walk_span = [&](double t0, double t1) {
      RtDda fd;
      fd.init(eye, dir, idir, t0, t1, 0);
      do {
        if (probe.on(mp, fd.vox[0], fd.vox[1], fd.vox[2])) mark_voxel(u, ovf, fd.vox[0], fd.vox[1], fd.vox[2], false);
      } while (fd.step() && --budget > 0);
    };
    double s0 = -1, s1 = -1;  // the open TimeSpan
    RtDda d12;
    d12.init(eye, dir, idir, rt0, rt1, 12);
    do {
      if (set_has(rt.n4096, rt.n4096_mask, node_key(d12.vox[0], d12.vox[1], d12.vox[2], 12))) {
        RtDda d7;
        d7.init(eye, dir, idir, d12.t0, d12.next(), 7);
        do {
          if (set_has(rt.n128, rt.n128_mask, node_key(d7.vox[0], d7.vox[1], d7.vox[2], 7))) {
            RtDda d3;
            d3.init(eye, dir, idir, d7.t0, d7.next(), 3);
            do {
              if (set_has(rt.leaf, rt.leaf_mask, leaf_key(d3.vox[0], d3.vox[1], d3.vox[2]))) {
                if (s0 < 0) s0 = d3.t0;
              } else if (s0 >= 0) {
                s1 = d3.t0;
                if (span_valid(s0, s1)) walk_span(s0, s1);
                s0 = s1 = -1;
              }
            } while (d3.step() && --budget > 0);
            if (s0 >= 0) s1 = d3.t1;
          } else if (s0 >= 0) {
            s1 = d7.t0;
            if (span_valid(s0, s1)) walk_span(s0, s1);
            s0 = s1 = -1;
          }
        } while (d7.step() && --budget > 0);
        if (s0 >= 0) s1 = d7.t1;
      } else if (s0 >= 0) {
        s1 = d12.t0;
        if (span_valid(s0, s1)) walk_span(s0, s1);
        s0 = s1 = -1;
      }
    } while (d12.step() && --budget > 0);
    if (s0 >= 0) s1 = d12.t1;
    if (span_valid(s0, s1)) walk_span(s0, s1);
    if (budget <= 0) atomicOr(&st->error, kErrWalk);
  }
}

// List the occupied update-set slots (keys only: coalesced 8-byte reads).
__global__ __launch_bounds__(256) void k_compact(UpdView u, uint32_t cap,
                                                 uint32_t* __restrict__ list, unsigned int* counter,
                                                 const Status* st) {
  __shared__ uint32_t s_wave[4], s_base;
  if (st->error != 0 || st->overflow != 0) return;  // same for every thread
  const uint32_t G = gridDim.x * blockDim.x;
  const uint32_t rounds = (cap + G - 1) / G;
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t slot = rd * G + blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = slot < cap && u.keys[slot] != kEmpty;
    const uint32_t pos = block_append(valid, counter, s_wave, &s_base);
    if (valid) list[pos] = slot;
  }
}

// One lane per update leaf: find its map leaf, or create it when the update
// changes a background voxel (OpenVDB creates a child node only when the
// modify op changes the tile's value or active state).
__global__ __launch_bounds__(256) void k_map_insert(UpdView u, const uint32_t* __restrict__ list,
                                                    uint32_t count, MapView mp,
                                                    uint32_t* __restrict__ eidx,
                                                    uint32_t create_hit, uint32_t create_miss,
                                                    unsigned long long* noop_set,
                                                    uint32_t noop_mask, Status* st, Status* stc) {
  __shared__ uint32_t s_wave[4], s_base;
  uint32_t listed = 0;
  const uint32_t G = gridDim.x * blockDim.x;
  const uint32_t e0 = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t rounds = (count + G - 1) / G;
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t e = e0 + rd * G;
    bool created = false;
    uint32_t h = kNone, idx = kNone;
    uint64_t key = 0;
    if (e < count && (list || u.keys[e] != kEmpty)) {
      ++listed;
      const uint32_t slot = list ? list[e] : e;
      key = u.keys[slot];
      // every listed leaf has a visited voxel: with both ops creating, it needs
      // its leaf whatever its masks hold (the usual case: p_hit, p_miss != 0.5)
      bool need = true;
      if (!(create_hit && create_miss)) {
        const unsigned long long* w = &u.masks[(size_t)slot * 16];
        uint64_t anyhit = 0, anymiss = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          anyhit |= w[j] & w[8 + j];
          anymiss |= w[j] & ~w[8 + j];
        }
        need = (anyhit && create_hit) || (anymiss && create_miss);
      }
      uint32_t p = (uint32_t)fmix64(key) & mp.cap_mask;
      for (uint32_t probe = 0; probe <= mp.cap_mask; ++probe) {
        const unsigned long long k = mp.keys[p];
        if (k == key) {
          idx = mp.idx[p];
          break;
        }
        if (k == kEmpty) {
          if (!need) break;
          const unsigned long long prev = atomicCAS(&mp.keys[p], kEmpty, (unsigned long long)key);
          if (prev == kEmpty) {
            created = true;
            h = p;
            break;
          }
        }
        p = (p + 1) & mp.cap_mask;
      }
      if (!need && idx == kNone) {
        // the root still gains the 4096^3 child even when no leaf is created
        int o[3];
        key_origin(key, o);
        set_insert(noop_set, noop_mask, node_key(o[0], o[1], o[2], 12));
      }
    }
    const uint32_t pos = block_append(created, &st->leaf_count, s_wave, &s_base);
    if (created) {
      if (pos < mp.lcap) {
        idx = pos;
        mp.idx[h] = idx;
        mp.leaf_keys[idx] = key;
      } else {  // the host sizes the pool first; never write past it
        mp.idx[h] = kNone;
        atomicOr(&st->map_full, 1u);
      }
    }
    if (e < count) {
      eidx[e] = idx;
      if (idx != kNone) mp.dirty[idx] = 1;
    }
  }
  // update leaves of this call: one atomic per workgroup
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) listed += __shfl_xor(listed, o, 64);
  if (lane_id() == 0) s_wave[threadIdx.x >> 6] = listed;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_wave[w];
    if (t) atomicAdd(&stc->entries, t);
  }
}

// Clamped log-odds update of one voxel (updateOccupiedNode / updateFreeNode,
// OccupancyVDBMapping.hpp:92-117): hit -> x += l_hit, above t_max: active,
// clamp at c_max; miss -> x += l_miss, below t_min: inactive, clamp at c_min.
// Written with selects, no branches (the same values and states).
__device__ __forceinline__ void update_voxel(float& x, uint32_t& ab, int k, bool visited, bool hit,
                                             float l_hit, float l_miss, float t_min, float t_max,
                                             float c_max, float c_min) {
  const float a = x + (hit ? l_hit : l_miss);
  const bool cross = hit ? (a > t_max) : (a < t_min);  // the threshold branch is taken
  const float cl = hit ? (a > c_max ? c_max : a) : (a < c_min ? c_min : a);
  x = visited ? (cross ? cl : a) : x;
  const uint32_t bit = 1u << k;
  ab = (visited && cross) ? (hit ? (ab | bit) : (ab & ~bit)) : ab;
}

// updateMap (VDBMapping.hpp:764-783): every visited voxel of every listed
// update leaf gets one clamped log-odds update. One wave per update leaf:
// lane l owns voxels 8l .. 8l+7 (two float4, so the wave reads and writes the
// leaf's 2 KiB of values as one contiguous run), takes its eight visited /
// hit / active bits from byte l & 7 of mask word l >> 3, and the eight lanes
// of a word OR-reduce its new active word. The values of a 128-byte line (four
// lanes) are read and written only when one of its 32 voxels is visited: a
// line is never written in part, and the 8% of an update leaf's lines with no
// visit (outdoor stream) are skipped (synchronous latency after 100 scans
// 2.76-2.77 -> 2.71-2.72 ms at p50; 64-byte blocks were slower). Each wave carries kApplyLeaves leaves per
// iteration: four (103 VGPRs, four waves per SIMD) gave p50 2.64-2.65 ms
// against 2.70-2.71 with two (64 VGPRs, eight waves), 2.72 with six and 2.87
// with eight. Optionally clears the update-set slots afterwards; counts the
// updated voxels.
constexpr int kApplyLeaves = 4;  // leaves in flight per wave

__global__ __launch_bounds__(256) void k_apply(UpdView u, const uint32_t* __restrict__ list,
                                               const uint32_t* __restrict__ eidx, uint32_t count,
                                               MapView mp, float l_hit, float l_miss, float t_min,
                                               float t_max, float c_max, float c_min, Status* st,
                                               uint32_t clear) {
  const int lane = lane_id();
  const int j = lane >> 3;        // mask word of the lane's voxels
  const int sh = (lane & 7) * 8;  // their first bit in it
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  unsigned long long vox = 0;
  for (uint32_t e0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kApplyLeaves; e0 < count;
       e0 += waves * kApplyLeaves) {
    uint32_t slot[kApplyLeaves], idx[kApplyLeaves];
    uint64_t vis[kApplyLeaves], hit[kApplyLeaves], aw[kApplyLeaves];
    bool blk[kApplyLeaves];
    float4 va[kApplyLeaves], vb4[kApplyLeaves];
#pragma unroll
    for (int q = 0; q < kApplyLeaves; ++q) {
      const uint32_t e = e0 + q;
      slot[q] = e < count ? list[e] : kNone;
      idx[q] = e < count ? eidx[e] : kNone;
    }
#pragma unroll
    for (int q = 0; q < kApplyLeaves; ++q) {
      vis[q] = hit[q] = aw[q] = 0;
      va[q] = vb4[q] = make_float4(0.f, 0.f, 0.f, 0.f);
      if (slot[q] != kNone) {
        vis[q] = u.masks[(size_t)slot[q] * 16 + j];
        hit[q] = u.masks[(size_t)slot[q] * 16 + 8 + j];
      }
      uint32_t lv = (uint32_t)(vis[q] >> sh) & 0xFFu;  // a visit in the lane's 128-byte line
      lv |= __shfl_xor(lv, 1, 64);
      lv |= __shfl_xor(lv, 2, 64);
      blk[q] = lv != 0;
      if (idx[q] != kNone) {
        const float4* pv = reinterpret_cast<const float4*>(mp.values) + (size_t)idx[q] * 128 + 2 * lane;
        if (blk[q]) {
          va[q] = pv[0];
          vb4[q] = pv[1];
        }
        aw[q] = mp.active[(size_t)idx[q] * 8 + j];
      }
    }
#pragma unroll
    for (int q = 0; q < kApplyLeaves; ++q) {
      if (clear && slot[q] != kNone) {  // after the loads of the iteration
        if ((lane & 7) == 0 && vis[q]) {
          u.masks[(size_t)slot[q] * 16 + j] = 0;
          u.masks[(size_t)slot[q] * 16 + 8 + j] = 0;
        }
        if (lane == 0) u.keys[slot[q]] = kEmpty;
      }
      const uint32_t v8 = (uint32_t)(vis[q] >> sh) & 0xFFu, h8 = (uint32_t)(hit[q] >> sh) & 0xFFu;
      uint32_t ab = (uint32_t)(aw[q] >> sh) & 0xFFu;
      vox += (unsigned long long)__popc(v8);
      update_voxel(va[q].x, ab, 0, v8 & 1u, h8 & 1u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(va[q].y, ab, 1, v8 & 2u, h8 & 2u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(va[q].z, ab, 2, v8 & 4u, h8 & 4u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(va[q].w, ab, 3, v8 & 8u, h8 & 8u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(vb4[q].x, ab, 4, v8 & 16u, h8 & 16u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(vb4[q].y, ab, 5, v8 & 32u, h8 & 32u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(vb4[q].z, ab, 6, v8 & 64u, h8 & 64u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      update_voxel(vb4[q].w, ab, 7, v8 & 128u, h8 & 128u, l_hit, l_miss, t_min, t_max, c_max, c_min);
      const bool live = vis[q] != 0 && idx[q] != kNone;  // the same for the word's eight lanes
      if (live && blk[q]) {
        float4* pv = reinterpret_cast<float4*>(mp.values) + (size_t)idx[q] * 128 + 2 * lane;
        pv[0] = va[q];
        pv[1] = vb4[q];
      }
      uint64_t na = (uint64_t)ab << sh;
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) na |= __shfl_xor(na, o, 64);
      if (live && (lane & 7) == 0) mp.active[(size_t)idx[q] * 8 + j] = na;
    }
  }
  __shared__ unsigned long long s_vox[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) vox += __shfl_xor(vox, o, 64);
  if (lane_id() == 0) s_vox[threadIdx.x >> 6] = vox;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tv = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) tv += s_vox[k];
    if (tv) atomicAdd(&st->voxels, tv);
  }
}

// ---- deferred scans: several brick-path scans applied in one pass ----------
// The records of up to kDeferMax consecutive scans (scan t of the group,
// t = 0 .. nscan-1 in insert order) are grouped by map leaf; each group's leaf
// is then read and written once while every scan's updates are applied to it
// in scan order, in registers (the per-voxel sequence of clamped adds of
// updateMap, VDBMapping.hpp:764-783, scan after scan). gslot (one entry per
// map leaf, kNone outside a grouping) names the record that opened the group.

// Lane per record of scan t (launched for t = 0, 1, ... in order): the first
// record of a map leaf opens its group (the group's record positions cleared,
// its own entered); a later one enters its position in the group that an
// earlier scan's record opened (a map leaf has one record per scan).
__global__ __launch_bounds__(256) void k_group_add(const uint32_t* __restrict__ eidx,
                                                   const uint32_t* __restrict__ list, uint32_t count, uint32_t t,
                                                   uint32_t* gslot, uint32_t* rgrp, uint32_t rgrp_stride,
                                                   uint32_t* glist, uint32_t* grec, unsigned int* ngroups) {
  __shared__ uint32_t s_wave[4];
  __shared__ uint32_t s_base;
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const uint32_t e = i < count ? eidx[i] : kNone;
  uint32_t code = kNone;
  if (e != kNone) code = atomicCAS(&gslot[e], kNone, (t << kGroupScanShift) | i);
  const bool opener = e != kNone && code == kNone;
  const uint32_t g = block_append(opener, ngroups, s_wave, &s_base);
  if (opener) {
    glist[g] = e;
    rgrp[(size_t)t * rgrp_stride + i] = g;
#pragma unroll
    for (int k = 0; k < kDeferMax; ++k) grec[(size_t)g * kDeferMax + k] = k == (int)t ? list[i] : kNone;
  } else if (e != kNone && (code >> kGroupScanShift) < t) {
    // opened by an earlier scan's launch (a scan's records name distinct leaves)
    const uint32_t go = rgrp[(size_t)(code >> kGroupScanShift) * rgrp_stride + (code & kGroupRecMask)];
    grec[(size_t)go * kDeferMax + t] = list[i];
  }
}

// Group apply: one wave per group (map leaf), lane l owning voxels 8l ..
// 8l+7 as in k_apply. The leaf's values and active words are loaded once
// (beside the masks), every scan of the group with a record of the leaf gives
// the lane's visited / hit byte (byte l & 7 of mask word l >> 3), the updates
// are applied in scan order, visits only, and the leaf is stored once. Each
// wave carries kGroupLeaves groups per iteration. Frees the group's gslot
// entry.
constexpr int kGroupLeaves = 2;
static_assert(kDeferMax == 8, "a group's record positions are read as two uint4; one byte per scan");

// Every scan's update of voxel i of a lane (bit 8i + s of vt / ht: visited /
// hit in scan s, the lane's masks transposed once per group), in scan order,
// visits only: one iteration per visit, the clamped update of update_voxel.
__device__ __forceinline__ void apply_visits(float& x, uint32_t& ab, int i, uint64_t vt, uint64_t ht,
                                             float l_hit, float l_miss, float t_min, float t_max,
                                             float c_max, float c_min) {
  uint32_t m = (uint32_t)(vt >> (8 * i)) & 0xFFu;
  if (!__any(m != 0)) return;
  const uint32_t hs = (uint32_t)(ht >> (8 * i)) & 0xFFu;
  float v = x;
  uint32_t on = (ab >> i) & 1u;
  if (!__any(hs != 0)) {
    // misses only (most voxels of most waves): the same clamped add once per
    // visit, so which scans made them does not matter
    while (m) {
      m &= m - 1;
      const float a = v + l_miss;
      const bool cross = a < t_min;
      v = cross ? (a < c_min ? c_min : a) : a;
      on = cross ? 0u : on;
    }
    x = v;
    ab = (ab & ~(1u << i)) | (on << i);
    return;
  }
  while (m) {
    const uint32_t b = (uint32_t)__builtin_ctz(m);
    m &= m - 1;
    const bool h = (hs >> b) & 1u;
    const float a = v + (h ? l_hit : l_miss);
    // hit: above t_max -> active, clamped at c_max; miss: below t_min -> inactive, clamped at c_min
    const bool cross = h ? (a > t_max) : (a < t_min);
    const float cl = h ? (a > c_max ? c_max : a) : (a < c_min ? c_min : a);
    v = cross ? cl : a;
    on = cross ? (h ? 1u : 0u) : on;
  }
  x = v;
  ab = (ab & ~(1u << i)) | (on << i);
}

__global__ __launch_bounds__(256) void k_apply_groups(GroupScans gs, const uint32_t* __restrict__ glist,
                                                      const uint32_t* __restrict__ grec,
                                                      const unsigned int* __restrict__ ngroups, uint32_t* gslot,
                                                      MapView mp, float l_hit, float l_miss, float t_min,
                                                      float t_max, float c_max, float c_min, Status* st) {
  const uint32_t n = *ngroups;
  const int lane = lane_id();
  const int j = lane >> 3;        // mask word of the lane's voxels
  const int sh = (lane & 7) * 8;  // their first bit in it
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  unsigned long long vox = 0;
  for (uint32_t g0 = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * kGroupLeaves; g0 < n;
       g0 += waves * kGroupLeaves) {
    uint32_t e[kGroupLeaves], slot[kGroupLeaves][kDeferMax];
#pragma unroll
    for (int q = 0; q < kGroupLeaves; ++q) {  // the group: leaf and record positions
      const uint32_t g = g0 + q;
      e[q] = kNone;
#pragma unroll
      for (int k = 0; k < kDeferMax; ++k) slot[q][k] = kNone;
      if (g < n) {  // the same for every lane: kept in scalar registers
        e[q] = __builtin_amdgcn_readfirstlane(glist[g]);
        const uint4* r = reinterpret_cast<const uint4*>(grec + (size_t)g * kDeferMax);
        const uint4 a = r[0], b = r[1];
        slot[q][0] = __builtin_amdgcn_readfirstlane(a.x), slot[q][1] = __builtin_amdgcn_readfirstlane(a.y);
        slot[q][2] = __builtin_amdgcn_readfirstlane(a.z), slot[q][3] = __builtin_amdgcn_readfirstlane(a.w);
        slot[q][4] = __builtin_amdgcn_readfirstlane(b.x), slot[q][5] = __builtin_amdgcn_readfirstlane(b.y);
        slot[q][6] = __builtin_amdgcn_readfirstlane(b.z), slot[q][7] = __builtin_amdgcn_readfirstlane(b.w);
      }
    }
    uint64_t vb[kGroupLeaves], hb[kGroupLeaves], aw[kGroupLeaves];
    float4 va[kGroupLeaves], vc[kGroupLeaves];
#pragma unroll
    for (int q = 0; q < kGroupLeaves; ++q) {  // every scan's masks, and the leaf beside them
      vb[q] = hb[q] = 0;
#pragma unroll
      for (int k = 0; k < kDeferMax; ++k) {
        // the wave holds one group: a scalar branch. The lane's byte is in one
        // 32-bit half of the mask word: only that half is loaded.
        if (k < gs.nscan && slot[q][k] != kNone) {
          const uint32_t* m = reinterpret_cast<const uint32_t*>(gs.masks[k] + (size_t)slot[q][k] * 16);
          const uint32_t v = m[2 * j + (sh >> 5)], h = m[16 + 2 * j + (sh >> 5)];
          vb[q] |= (uint64_t)((v >> (sh & 31)) & 0xFFu) << (8 * k);
          hb[q] |= (uint64_t)((h >> (sh & 31)) & 0xFFu) << (8 * k);
        }
      }
      va[q] = vc[q] = make_float4(0.f, 0.f, 0.f, 0.f);
      aw[q] = 0;
      if (e[q] != kNone) {
        const float4* pv = reinterpret_cast<const float4*>(mp.values) + (size_t)e[q] * 128 + 2 * lane;
        va[q] = pv[0];
        vc[q] = pv[1];
        aw[q] = mp.active[(size_t)e[q] * 8 + j];
      }
    }
#pragma unroll
    for (int q = 0; q < kGroupLeaves; ++q) {
      if (e[q] == kNone) continue;  // the whole wave
      if (lane == 0) gslot[e[q]] = kNone;
      vox += (unsigned long long)__popcll(vb[q]);
      uint32_t ab = (uint32_t)(aw[q] >> sh) & 0xFFu;
      // byte i: voxel i's visits / hits, bit s per scan
      const uint64_t vt = transpose8(vb[q]), ht = transpose8(hb[q]);
      apply_visits(va[q].x, ab, 0, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(va[q].y, ab, 1, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(va[q].z, ab, 2, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(va[q].w, ab, 3, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(vc[q].x, ab, 4, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(vc[q].y, ab, 5, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(vc[q].z, ab, 6, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      apply_visits(vc[q].w, ab, 7, vt, ht, l_hit, l_miss, t_min, t_max, c_max, c_min);
      // every word with a visit in some scan is written back whole (one owner per
      // row): the eight lanes of the word agree on it through the reduction
      uint64_t na = (uint64_t)ab << sh;
      uint32_t wany = vb[q] != 0 ? 1u : 0u;
#pragma unroll
      for (int o = 1; o < 8; o <<= 1) {
        na |= __shfl_xor(na, o, 64);
        wany |= __shfl_xor(wany, o, 64);
      }
      if (wany) {
        float4* pv = reinterpret_cast<float4*>(mp.values) + (size_t)e[q] * 128 + 2 * lane;
        pv[0] = va[q];
        pv[1] = vc[q];
        if ((lane & 7) == 0) mp.active[(size_t)e[q] * 8 + j] = na;
      }
    }
  }
  __shared__ unsigned long long s_vox[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) vox += __shfl_xor(vox, o, 64);
  if (lane_id() == 0) s_vox[threadIdx.x >> 6] = vox;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long tv = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) tv += s_vox[k];
    if (tv) atomicAdd(&st->voxels, tv);
  }
}

// Move every entry of an update set into a larger one.
__global__ __launch_bounds__(256) void k_rehash_update(UpdView from, uint32_t from_cap, UpdView to,
                                                       Status* st) {
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < from_cap; s += G) {
    const uint64_t key = from.keys[s];
    if (key == kEmpty) continue;
    const uint32_t d = upd_slot(to, key, &st->overflow);
    if (d == kNone) continue;
    for (int j = 0; j < 16; ++j) to.masks[(size_t)d * 16 + j] |= from.masks[(size_t)s * 16 + j];
  }
}

// Rebuild the map leaf hash from the leaf pool's key list.
__global__ __launch_bounds__(256) void k_rehash_map(MapView mp, uint32_t leaf_count, const Status* stc) {
  if (stc) leaf_count = stc->leaf_count;  // the count at this point of the stream
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < leaf_count; i += G) {
    const uint64_t key = mp.leaf_keys[i];
    uint32_t p = (uint32_t)fmix64(key) & mp.cap_mask;
    for (uint32_t probe = 0; probe <= mp.cap_mask; ++probe) {
      if (atomicCAS(&mp.keys[p], kEmpty, (unsigned long long)key) == kEmpty) {
        mp.idx[p] = i;
        break;
      }
      p = (p + 1) & mp.cap_mask;
    }
  }
}

// Per-scan set-up of a single-scan cast in one launch: the scan record and
// the size block (by value, so no host-to-device copy is queued behind the
// previous scan's map update), a zeroed scan status, and optionally the
// brick table cleared (keys EMPTY, counts and cursors 0).
__global__ void k_stage_scan(ScanInfo sc, BatchCtl ctl, ScanInfo* __restrict__ d_scans,
                             BatchCtl* __restrict__ d_ctl, Status* st, unsigned long long* bkeys,
                             unsigned long long* bcnt, uint64_t nb) {
  const uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) {
    d_scans[0] = sc;
    *d_ctl = ctl;
  }
  if (i0 < sizeof(Status) / 4) reinterpret_cast<uint32_t*>(st)[i0] = 0u;
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = i0; i < nb; i += G) {
    bkeys[i] = kEmpty;
    bcnt[i] = 0;
  }
}

__global__ void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += G) p[i] = v;
}

// List (and clear) the dirty leaves, then gather their data contiguously.
__global__ __launch_bounds__(256) void k_dirty_list(uint8_t* dirty, uint32_t leaf_count,
                                                    uint32_t* list, Status* st) {
  const uint32_t G = gridDim.x * blockDim.x;
  const uint32_t rounds = (leaf_count + G - 1) / G;
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x + rd * G;
    const bool d = i < leaf_count && dirty[i];
    const uint64_t m = __ballot(d);
    if (m) {
      const int leader = __ffsll((long long)m) - 1;
      uint32_t base = 0;
      if (lane_id() == leader) base = atomicAdd(&st->entries, (unsigned int)__popcll(m));
      base = __shfl(base, leader);
      if (d) {
        list[base + __popcll(m & ((1ull << lane_id()) - 1))] = i;
        dirty[i] = 0;
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_gather_leaves(const uint32_t* __restrict__ list,
                                                       uint32_t count, MapView mp,
                                                       unsigned long long* __restrict__ keys_out,
                                                       float* __restrict__ vals_out,
                                                       unsigned long long* __restrict__ act_out) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const int lane = lane_id();
  for (uint32_t e = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); e < count; e += waves) {
    const uint32_t i = list[e];
    const float* src = mp.values + (size_t)i * 512;
    float* dst = vals_out + (size_t)e * 512;
#pragma unroll
    for (int j = 0; j < 8; ++j) dst[j * 64 + lane] = src[j * 64 + lane];
    if (lane < 8) act_out[(size_t)e * 8 + lane] = mp.active[(size_t)i * 8 + lane];
    if (lane == 0) keys_out[e] = mp.leaf_keys[i];
  }
}

// ------------------------------------------------------------- launchers --

hipError_t launch_build_table(hipStream_t s, double* tbl, uint32_t n_lo, uint32_t n_hi) {
  const uint32_t cnt = n_hi - n_lo + 1;
  hipLaunchKernelGGL(k_build_table, dim3((cnt + 255) / 256), dim3(256), 0, s, tbl, n_lo, n_hi);
  return hipGetLastError();
}

hipError_t launch_prep(hipStream_t s, int grid, const float* pts, int64_t stride_bytes, bool vec4,
                       ScanInfo* scans, const BatchCtl* ctl, double res, double inv,
                       uint32_t tbl_n, int4* rayinfo, uint32_t* pre, uint32_t* bsum, Status* st) {
  if (vec4)
    hipLaunchKernelGGL(k_prep<true>, dim3(grid), dim3(kPrepBlock), 0, s, pts, stride_bytes / 4,
                       scans, ctl, res, inv, tbl_n, rayinfo, pre, bsum, st);
  else
    hipLaunchKernelGGL(k_prep<false>, dim3(grid), dim3(kPrepBlock), 0, s, pts, stride_bytes / 4,
                       scans, ctl, res, inv, tbl_n, rayinfo, pre, bsum, st);
  return hipGetLastError();
}

hipError_t launch_scan_blocks(hipStream_t s, const uint32_t* bsum, uint64_t* boff,
                              const BatchCtl* ctl, Status* st) {
  hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(256), 0, s, bsum, boff, ctl, st);
  return hipGetLastError();
}

hipError_t launch_ray_offsets(hipStream_t s, int grid, const int4* rayinfo, const uint32_t* pre,
                              const uint64_t* boff, const BatchCtl* ctl, uint32_t chunk,
                              uint64_t* roff, uint32_t* first_ray, uint64_t max_chunks,
                              Status* st) {
  hipLaunchKernelGGL(k_ray_offsets, dim3(grid), dim3(256), 0, s, rayinfo, pre, boff, ctl, chunk, roff,
                     first_ray, max_chunks, st);
  return hipGetLastError();
}

hipError_t launch_traverse(hipStream_t s, int grid, int block, const int4* rayinfo,
                           const uint64_t* roff, const uint32_t* first_ray, const ScanInfo* scans,
                           const BatchCtl* ctl, const double* tbl, UpdView u, Status* st,
                           uint32_t chunk) {
  hipLaunchKernelGGL(k_traverse, dim3(grid), dim3(block), 0, s, rayinfo, roff, first_ray, scans, ctl,
                     tbl, u, st, chunk);
  return hipGetLastError();
}

hipError_t launch_fast_cast(hipStream_t s, int grid, const int4* rayinfo, const ScanInfo* scans,
                            const BatchCtl* ctl, RtView rt, MapView mp, UpdView u, bool cast,
                            Status* st) {
  hipLaunchKernelGGL(k_fast_cast, dim3(grid), dim3(256), 0, s, rayinfo, scans, ctl, rt, mp, u,
                     (uint32_t)cast, st);
  return hipGetLastError();
}

hipError_t launch_compact(hipStream_t s, int grid, UpdView u, uint32_t cap, uint32_t* list,
                          unsigned int* counter, const Status* st) {
  hipLaunchKernelGGL(k_compact, dim3(grid), dim3(256), 0, s, u, cap, list, counter, st);
  return hipGetLastError();
}

hipError_t launch_map_insert(hipStream_t s, int grid, UpdView u, const uint32_t* list,
                             uint32_t count, MapView mp, uint32_t* eidx, bool create_hit,
                             bool create_miss, unsigned long long* noop_set, uint32_t noop_mask,
                             Status* st, Status* stc) {
  hipLaunchKernelGGL(k_map_insert, dim3(grid), dim3(256), 0, s, u, list, count, mp, eidx,
                     (uint32_t)create_hit, (uint32_t)create_miss, noop_set, noop_mask, st, stc);
  return hipGetLastError();
}

hipError_t launch_apply(hipStream_t s, int grid, UpdView u, const uint32_t* list,
                        const uint32_t* eidx, uint32_t count, MapView mp, const float lo[6],
                        Status* st, bool clear) {
  hipLaunchKernelGGL(k_apply, dim3(grid), dim3(256), 0, s, u, list, eidx, count, mp, lo[0], lo[1],
                     lo[2], lo[3], lo[4], lo[5], st, (uint32_t)clear);
  return hipGetLastError();
}

hipError_t launch_group_add(hipStream_t s, const uint32_t* eidx, const uint32_t* list, uint32_t count, uint32_t t,
                            uint32_t* gslot, uint32_t* rgrp, uint32_t rgrp_stride, uint32_t* glist, uint32_t* grec,
                            unsigned int* ngroups) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_group_add, dim3((count + 255) / 256), dim3(256), 0, s, eidx, list, count, t, gslot, rgrp,
                     rgrp_stride, glist, grec, ngroups);
  return hipGetLastError();
}

hipError_t launch_apply_groups(hipStream_t s, int grid, const GroupScans& gs, const uint32_t* glist,
                               const uint32_t* grec, const unsigned int* ngroups, uint32_t* gslot, MapView mp,
                               const float lo[6], Status* st) {
  hipLaunchKernelGGL(k_apply_groups, dim3(grid), dim3(256), 0, s, gs, glist, grec, ngroups, gslot, mp, lo[0],
                     lo[1], lo[2], lo[3], lo[4], lo[5], st);
  return hipGetLastError();
}

hipError_t launch_rehash_update(hipStream_t s, int grid, UpdView from, uint32_t from_cap,
                                UpdView to, Status* st) {
  hipLaunchKernelGGL(k_rehash_update, dim3(grid), dim3(256), 0, s, from, from_cap, to, st);
  return hipGetLastError();
}

hipError_t launch_rehash_map(hipStream_t s, int grid, MapView mp, uint32_t leaf_count, const Status* stc) {
  hipLaunchKernelGGL(k_rehash_map, dim3(grid), dim3(256), 0, s, mp, leaf_count, stc);
  return hipGetLastError();
}

hipError_t launch_stage_scan(hipStream_t s, const ScanInfo& sc, const BatchCtl& ctl, ScanInfo* d_scans,
                             BatchCtl* d_ctl, Status* st, unsigned long long* bkeys, unsigned long long* bcnt,
                             uint64_t nb) {
  const uint64_t blocks = (nb + 255) / 256;
  const int grid = (int)(blocks < 1 ? 1 : (blocks < 4096 ? blocks : 4096));
  hipLaunchKernelGGL(k_stage_scan, dim3(grid), dim3(256), 0, s, sc, ctl, d_scans, d_ctl, st, bkeys, bcnt, nb);
  return hipGetLastError();
}

hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, uint64_t n, uint64_t v) {
  if (n == 0) return hipSuccess;
  const uint64_t blocks = (n + 255) / 256;
  const int grid = (int)(blocks < 4096 ? blocks : 4096);
  hipLaunchKernelGGL(k_fill_u64, dim3(grid), dim3(256), 0, s, p, n, (unsigned long long)v);
  return hipGetLastError();
}

hipError_t launch_dirty_list(hipStream_t s, int grid, uint8_t* dirty, uint32_t leaf_count,
                             uint32_t* list, Status* st) {
  hipLaunchKernelGGL(k_dirty_list, dim3(grid), dim3(256), 0, s, dirty, leaf_count, list, st);
  return hipGetLastError();
}

hipError_t launch_gather_leaves(hipStream_t s, int grid, const uint32_t* list, uint32_t count,
                                MapView mp, unsigned long long* keys_out, float* vals_out,
                                unsigned long long* act_out) {
  hipLaunchKernelGGL(k_gather_leaves, dim3(grid), dim3(256), 0, s, list, count, mp, keys_out,
                     vals_out, act_out);
  return hipGetLastError();
}

}  // namespace vdbgpu
