// gfx950 kernels of batched integration: many scans of one source folded into
// the map with a result identical to inserting them one after another
// (VDBMapping.hpp:399-406 per scan, updateMap :764-783 + OccupancyVDBMapping.hpp:92-117
// per voxel and scan).
//
// The batch's update set is keyed by (leaf, scan). Its entries are bucketed
// by leaf into a group table whose 128-bit scan mask orders them by scan
// index, which is the (leaf key, scan index) order: a wave per leaf then loads
// the map leaf once, applies every scan's hit/miss masks in ascending scan
// order in registers and writes the leaf back once.
//
// Every kernel reads its sizes from device memory and returns early when an
// earlier kernel flagged an error, an update-set overflow or a full map, so
// the host can grow the tables and re-run before anything is applied, and the
// whole sequence can be captured in a graph and replayed.
#include "device_common.h"
#include "kernels.h"

namespace vdbgpu {

__device__ __forceinline__ int bk_lane() { return (int)__lane_id(); }

// Slot of `key` in an open-addressing set, inserting it when absent.
__device__ __forceinline__ uint32_t set_slot(unsigned long long* keys, uint32_t mask, uint64_t key) {
  uint32_t h = (uint32_t)fmix64(key) & mask;
  for (uint32_t p = 0; p <= mask; ++p) {
    const unsigned long long k = keys[h];
    if (k == key) return h;
    if (k == kEmpty) {
      const unsigned long long prev = atomicCAS(&keys[h], kEmpty, (unsigned long long)key);
      if (prev == kEmpty || prev == key) return h;
    }
    h = (h + 1) & mask;
  }
  return kNone;
}

__device__ __forceinline__ uint32_t find_slot(const unsigned long long* __restrict__ keys,
                                              uint32_t mask, uint64_t key) {
  uint32_t h = (uint32_t)fmix64(key) & mask;
  for (uint32_t p = 0; p <= mask; ++p) {
    const unsigned long long k = keys[h];
    if (k == key) return h;
    if (k == kEmpty) return kNone;
    h = (h + 1) & mask;
  }
  return kNone;
}

__device__ __forceinline__ bool batch_stopped(const Status* st) {
  return st->error != 0 || st->overflow != 0 || st->map_full != 0;
}

// Lane per (leaf, scan) entry: add its scan bit to the leaf's group, and mark
// the group when the entry would create a missing leaf.
__global__ __launch_bounds__(256) void k_group_build(UpdView u, const uint32_t* __restrict__ list,
                                                     GroupView gv, uint32_t create_hit,
                                                     uint32_t create_miss, const Status* st) {
  if (batch_stopped(st)) return;
  const uint32_t count = st->entries;
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += G) {
    const uint32_t slot = list[e];
    const uint64_t key = u.keys[slot];
    const unsigned long long* w = &u.masks[(size_t)slot * 16];
    uint64_t anyhit = 0, anymiss = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      anyhit |= w[j] & w[8 + j];
      anymiss |= w[j] & ~w[8 + j];
    }
    const uint32_t gs = set_slot(gv.keys, gv.cap_mask, key >> 7);
    if (gs == kNone) continue;  // cannot happen: groups <= entries <= capacity
    const uint32_t scan = (uint32_t)(key & kScanMask);
    atomicOr(&gv.scans[(size_t)gs * 2 + (scan >> 6)], 1ull << (scan & 63));
    if ((anyhit && create_hit) || (anymiss && create_miss)) atomicOr(&gv.need[gs], 1u);
  }
}

// One thread: does the map have room for one new leaf per group?
__global__ void k_map_check(Status* st, const Status* stmap, uint32_t lcap, uint32_t mcap) {
  if (batch_stopped(st)) return;
  const uint64_t need = (uint64_t)stmap->leaf_count + st->groups;
  if (need > lcap || 2 * need > mcap) st->map_full = kMapFull;
}

// Lane per group: find or create the map leaf (same creation rule as
// k_map_insert), one counter atomic per workgroup.
__global__ __launch_bounds__(256) void k_group_insert(GroupView gv, const uint32_t* __restrict__ glist,
                                                      MapView mp, uint32_t* __restrict__ gidx,
                                                      unsigned long long* noop_set, uint32_t noop_mask,
                                                      const Status* st, Status* stmap) {
  __shared__ uint32_t s_wave[4], s_base;
  if (batch_stopped(st)) return;  // same value for every thread of the grid
  const uint32_t count = st->groups;
  const uint32_t G = gridDim.x * blockDim.x;
  const uint32_t rounds = (count + G - 1) / G;
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t g = rd * G + blockIdx.x * blockDim.x + threadIdx.x;
    bool created = false;
    uint32_t h = kNone, idx = kNone;
    uint64_t key = 0;
    if (g < count) {
      const uint32_t gs = glist[g];
      key = batched_to_leaf_key(gv.keys[gs] << 7);  // group keys drop the 7 scan bits
      const bool need = gv.need[gs] != 0;
      uint32_t p = (uint32_t)fmix64(key) & mp.cap_mask;
      for (uint32_t probe = 0; probe <= mp.cap_mask; ++probe) {
        const unsigned long long k = mp.keys[p];
        if (k == key) {
          idx = mp.idx[p];
          break;
        }
        if (k == kEmpty) {
          if (!need) break;
          if (atomicCAS(&mp.keys[p], kEmpty, (unsigned long long)key) == kEmpty) {
            created = true;
            h = p;
            break;
          }
        }
        p = (p + 1) & mp.cap_mask;
      }
      if (!need && idx == kNone) {
        int o[3];
        key_origin(key, o);
        set_insert(noop_set, noop_mask, node_key(o[0], o[1], o[2], 12));
      }
    }
    // workgroup-aggregated allocation of the new leaf indices
    const uint64_t m = __ballot(created);
    const int w = threadIdx.x >> 6;
    if (bk_lane() == 0) s_wave[w] = (uint32_t)__popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
      uint32_t tot = 0;
      for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
        const uint32_t c = s_wave[i];
        s_wave[i] = tot;
        tot += c;
      }
      s_base = tot ? atomicAdd(&stmap->leaf_count, tot) : 0u;
    }
    __syncthreads();
    if (created) {
      idx = s_base + s_wave[w] + (uint32_t)__popcll(m & ((1ull << bk_lane()) - 1));
      mp.idx[h] = idx;
      mp.leaf_keys[idx] = key;
    }
    __syncthreads();
    if (g < count) {
      gidx[g] = idx;
      if (idx != kNone) mp.dirty[idx] = 1;
    }
  }
}

// Wave per leaf: every scan's update of the leaf in ascending scan order.
// Lane l holds voxels l, 64+l, ..., 448+l in registers across the scans.
// Zeroes the consumed update masks and the group slot (keys are cleared by
// k_clear_keys afterwards: other waves still probe the update set).
__global__ __launch_bounds__(256) void k_fold(UpdView u, GroupView gv, const uint32_t* __restrict__ glist,
                                              const uint32_t* __restrict__ gidx, MapView mp, float l_hit,
                                              float l_miss, float t_min, float t_max, float c_max,
                                              float c_min, const Status* st, Status* stcount) {
  if (batch_stopped(st)) return;
  const uint32_t count = st->groups;
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const int lane = bk_lane();
  unsigned long long vox = 0;
  for (uint32_t g = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); g < count; g += waves) {
    const uint32_t gs = glist[g];
    const uint64_t gkey = gv.keys[gs];
    const uint64_t sm[2] = {gv.scans[(size_t)gs * 2], gv.scans[(size_t)gs * 2 + 1]};
    const uint32_t idx = gidx[g];
    float* vals = idx != kNone ? mp.values + (size_t)idx * 512 : nullptr;
    unsigned long long* act = idx != kNone ? mp.active + (size_t)idx * 8 : nullptr;
    float x[8];
    bool a[8];
    uint64_t touched[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      touched[j] = 0;
      x[j] = vals ? vals[j * 64 + lane] : 0.0f;
      a[j] = act ? ((act[j] >> lane) & 1) : false;
    }
    for (int half = 0; half < 2; ++half) {
      uint64_t bits = sm[half];
      while (bits) {
        const uint32_t scan = (uint32_t)(half * 64 + __ffsll((long long)bits) - 1);
        bits &= bits - 1;
        const uint32_t slot = find_slot(u.keys, u.cap_mask, (gkey << 7) | scan);
        if (slot == kNone) continue;
        unsigned long long* w = &u.masks[(size_t)slot * 16];
        uint64_t vis[8], hit[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
          vis[j] = w[j];
          hit[j] = w[8 + j];
          vox += __popcll(vis[j]);
        }
        if (vals) {
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            if (!((vis[j] >> lane) & 1)) continue;
            float v = x[j];
            if ((hit[j] >> lane) & 1) {
              v += l_hit;
              if (v > t_max) {
                a[j] = true;
                if (v > c_max) v = c_max;
              }
            } else {
              v += l_miss;
              if (v < t_min) {
                a[j] = false;
                if (v < c_min) v = c_min;
              }
            }
            x[j] = v;
          }
#pragma unroll
          for (int j = 0; j < 8; ++j) touched[j] |= vis[j];
        }
        if (lane < 16) w[lane] = 0;
      }
    }
    if (vals) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (touched[j] == 0) continue;
        if ((touched[j] >> lane) & 1) vals[j * 64 + lane] = x[j];
        const uint64_t na = __ballot(a[j]);
        if (lane == 0) act[j] = na;
      }
    }
    if (lane == 0) {
      gv.keys[gs] = kEmpty;
      gv.scans[(size_t)gs * 2] = 0;
      gv.scans[(size_t)gs * 2 + 1] = 0;
      gv.need[gs] = 0;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) vox += __shfl_xor(vox, o, 64);
  if (lane == 0 && vox) atomicAdd(&stcount->voxels, vox);
}

// Empty the update-set keys listed by k_compact (after the fold).
__global__ __launch_bounds__(256) void k_clear_keys(UpdView u, const uint32_t* __restrict__ list,
                                                    const Status* st) {
  if (batch_stopped(st)) return;
  const uint32_t count = st->entries;
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += G) u.keys[list[e]] = kEmpty;
}

// Zero n 64-bit words (graph-safe replacement of hipMemsetAsync).
__global__ void k_zero_words(unsigned long long* p, uint32_t n) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

hipError_t launch_group_build(hipStream_t s, int grid, UpdView u, const uint32_t* list, GroupView gv,
                              bool create_hit, bool create_miss, const Status* st) {
  hipLaunchKernelGGL(k_group_build, dim3(grid), dim3(256), 0, s, u, list, gv, (uint32_t)create_hit,
                     (uint32_t)create_miss, st);
  return hipGetLastError();
}

hipError_t launch_map_check(hipStream_t s, Status* st, const Status* stmap, uint32_t lcap,
                            uint32_t mcap) {
  hipLaunchKernelGGL(k_map_check, dim3(1), dim3(1), 0, s, st, stmap, lcap, mcap);
  return hipGetLastError();
}

hipError_t launch_group_insert(hipStream_t s, int grid, GroupView gv, const uint32_t* glist, MapView mp,
                               uint32_t* gidx, unsigned long long* noop_set, uint32_t noop_mask,
                               const Status* st, Status* stmap) {
  hipLaunchKernelGGL(k_group_insert, dim3(grid), dim3(256), 0, s, gv, glist, mp, gidx, noop_set,
                     noop_mask, st, stmap);
  return hipGetLastError();
}

hipError_t launch_fold(hipStream_t s, int grid, UpdView u, GroupView gv, const uint32_t* glist,
                       const uint32_t* gidx, MapView mp, const float lo[6], const Status* st,
                       Status* stcount) {
  hipLaunchKernelGGL(k_fold, dim3(grid), dim3(256), 0, s, u, gv, glist, gidx, mp, lo[0], lo[1], lo[2],
                     lo[3], lo[4], lo[5], st, stcount);
  return hipGetLastError();
}

hipError_t launch_clear_keys(hipStream_t s, int grid, UpdView u, const uint32_t* list,
                             const Status* st) {
  hipLaunchKernelGGL(k_clear_keys, dim3(grid), dim3(256), 0, s, u, list, st);
  return hipGetLastError();
}

hipError_t launch_zero_words(hipStream_t s, unsigned long long* p, uint32_t n) {
  hipLaunchKernelGGL(k_zero_words, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace vdbgpu
