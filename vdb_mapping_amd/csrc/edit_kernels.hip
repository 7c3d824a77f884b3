// gfx950 kernels of the direct map edits, artificial areas and map sections.
//
//   k_mark_points     addPointsToGrid / removePointsFromGrid (VDBMapping.hpp:413-447):
//                     one lane per point, voxel floor(p * 1/res), marked in an
//                     edit set (update-set layout: leaf key -> voxel mask)
//   k_mark_walls      addArtificialWall (VDBMapping.hpp:1217-1236): one lane per
//                     (polygon edge, height layer); rounded edge ends
//                     (VDBMapping.hpp:612-631) shifted by the layer, then the
//                     castRayIntoGrid walk (VDBMapping.hpp:550-566) marks the
//                     artificial set
//   k_set_voxels      one edit op on the marked voxels of listed edit-set leaves
//                     (set occupied / set free / restore state / activate)
//   k_section_pick    getMapSection (VDBMapping.hpp:922-1014): per map leaf, the
//                     section leaf it yields (full or sparse), appended to a list
//   k_upsert_leaves   find or create the map leaf of each section leaf
//   k_section_write   applyMapSectionGrid (VDBMapping.hpp:1022-1050) voxel writes
//   k_section_tiles   the section's internal-node tiles visited by the same loop
//   k_section_clear   applyMapSectionUpdateGrid (VDBMapping.hpp:1058-1085):
//                     deactivate map voxels inside the box, then OR the section
//
// Build with -ffp-contract=off like the other kernels.
#include <float.h>

#include "dda_math.h"
#include "device_common.h"
#include "kernels.h"
#include "upd_common.h"

namespace vdbgpu {

// One lane per point. A NaN or out-of-range voxel sets kErrRange (the host
// then discards the whole edit set); the others are marked visited.
__global__ __launch_bounds__(256) void k_mark_points(const float* __restrict__ pts, int64_t stride_f,
                                                     uint32_t n, double inv, UpdView u, Status* st) {
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += G) {
    const float* p = pts + (int64_t)i * stride_f;
    int c[3];
    bool ok = true;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double f = floor((double)p[a] * inv);
      ok = ok && f >= -(double)kCoordLimit && f < (double)kCoordLimit;
      c[a] = ok ? (int)f : 0;
    }
    if (!ok) {
      atomicOr(&st->error, kErrRange);
      continue;
    }
    mark_voxel(u, &st->overflow, c[0], c[1], c[2], false);
  }
}

// One lane per (wall, layer): wall w = {a, b, k0, k1} joins vertices a and b
// of `verts` (x, y, z doubles); its layers k in [k0, k1) are lanes
// prefix[w] .. prefix[w + 1] - 1 and shift both rounded ends by (0, 0, k). A
// zero-length wall marks nothing (castRayIntoGrid's e == s guard); otherwise
// the walk visits 1 + |dx| + |dy| + |dz| voxels in exact DDA order (integer
// crossing comparisons, the crossing table for equal-time ties).
__global__ __launch_bounds__(64) void k_mark_walls(const double* __restrict__ verts,
                                                   const int4* __restrict__ walls,
                                                   const uint64_t* __restrict__ prefix, uint32_t n_walls,
                                                   double res, double inv, const double* __restrict__ tbl,
                                                   uint32_t tbl_n, UpdView u, Status* st) {
  const uint64_t total = prefix[n_walls];
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += G) {
    uint32_t lo = 0, hi = n_walls - 1;  // last wall with prefix <= t
    while (lo < hi) {
      const uint32_t mid = (lo + hi + 1) >> 1;
      if (prefix[mid] <= t)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int4 wr = walls[lo];
    const int k = wr.z + (int)(t - prefix[lo]);
    const double* a = verts + 3 * (int64_t)wr.x;
    const double* b = verts + 3 * (int64_t)wr.y;
    int s[3], e[3];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      s[j] = world_to_index1(a[j], res, inv);
      e[j] = world_to_index1(b[j], res, inv);
    }
    const long long sz = (long long)s[2] + k, ez = (long long)e[2] + k;
#pragma unroll
    for (int j = 0; j < 2; ++j)
      ok = ok && s[j] >= -kCoordLimit && s[j] < kCoordLimit && e[j] >= -kCoordLimit && e[j] < kCoordLimit;
    ok = ok && s[2] != INT_MIN && e[2] != INT_MIN && sz >= -kCoordLimit && sz < kCoordLimit &&
         ez >= -kCoordLimit && ez < kCoordLimit;
    if (!ok) {
      atomicOr(&st->error, kErrRange);
      continue;
    }
    s[2] = (int)sz;
    e[2] = (int)ez;
    if (s[0] == e[0] && s[1] == e[1] && s[2] == e[2]) continue;
    const uint32_t nn[3] = {(uint32_t)abs(e[0] - s[0]), (uint32_t)abs(e[1] - s[1]),
                            (uint32_t)abs(e[2] - s[2])};
    if (nn[0] > tbl_n || nn[1] > tbl_n || nn[2] > tbl_n) {
      atomicOr(&st->error, kErrTable);
      continue;
    }
    const int stx = e[0] > s[0] ? 1 : (e[0] < s[0] ? -1 : 0);
    const int sty = e[1] > s[1] ? 1 : (e[1] < s[1] ? -1 : 0);
    const int stz = e[2] > s[2] ? 1 : (e[2] < s[2] ? -1 : 0);
    const uint32_t cc[3] = {0, 0, 0};
    IntDda d;
    d.init(nn, cc);
    int vx = s[0], vy = s[1], vz = s[2];
    const uint32_t m = 1 + nn[0] + nn[1] + nn[2];
    for (uint32_t v = 0;;) {
      mark_voxel(u, &st->overflow, vx, vy, vz, false);
      if (++v == m) break;
      const int ax = d.next_axis(tbl);
      d.advance(ax);
      vx += ax == 0 ? stx : 0;
      vy += ax == 1 ? sty : 0;
      vz += ax == 2 ? stz : 0;
    }
  }
}

// One lane per (listed leaf, mask word): the op on every marked voxel of the
// word. kEditOccupied: (val, on) (setNodeToOccupied, OccupancyVDBMapping.hpp:124-129);
// kEditFree: (val, off) (setNodeToFree, :118-123); kEditState: on = value > t_max
// (setNodeState, :130-134); kEditActivate: on (setActiveState). Optionally
// clears the edit-set slots.
__global__ __launch_bounds__(256) void k_set_voxels(UpdView u, const uint32_t* __restrict__ list,
                                                    const uint32_t* __restrict__ eidx, uint32_t count,
                                                    MapView mp, int op, float val, float t_max,
                                                    uint32_t clear) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)count * 8; t += G) {
    const uint32_t e = (uint32_t)(t >> 3);
    const int j = (int)(t & 7);
    const uint32_t slot = list[e];
    unsigned long long* w = &u.masks[(size_t)slot * 16];
    const uint64_t vis = w[j];
    if (clear) {
      if (vis) {
        w[j] = 0;
        w[8 + j] = 0;
      }
      if (j == 0) u.keys[slot] = kEmpty;
    }
    const uint32_t idx = eidx[e];
    if (vis == 0 || idx == kNone) continue;
    float* pv = mp.values + (size_t)idx * 512 + j * 64;
    unsigned long long* pa = mp.active + (size_t)idx * 8 + j;
    uint64_t aw = *pa;
    if (op == kEditOccupied || op == kEditFree) {
      for (uint64_t b = vis; b; b &= b - 1) pv[__builtin_ctzll(b)] = val;
      aw = op == kEditOccupied ? (aw | vis) : (aw & ~vis);
    } else if (op == kEditState) {
      for (uint64_t b = vis; b; b &= b - 1) {
        const int k = __builtin_ctzll(b);
        const uint64_t bit = 1ull << k;
        aw = pv[k] > t_max ? (aw | bit) : (aw & ~bit);
      }
    } else {
      aw |= vis;
    }
    *pa = aw;
    mp.dirty[idx] = 1;
  }
}

// Voxels of the leaf at origin o inside the inclusive box b, as 8 words (one
// per x slab, bit (y & 7) * 8 + (z & 7)).
__device__ __forceinline__ uint64_t box_word(const int o[3], const int* b, int j) {
  const int x = o[0] + j;
  if (x < b[0] || x > b[3]) return 0;
  uint64_t ym = 0, zm = 0;
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    if (o[1] + k >= b[1] && o[1] + k <= b[4]) ym |= 1ull << (8 * k);
    if (o[2] + k >= b[2] && o[2] + k <= b[5]) zm |= 1ull << k;
  }
  return ym * zm;  // product of the y-rows and z-columns selected
}

// One wave per map leaf: the section leaf it yields, if any. Full grid: the
// section's leaf appears at the first voxel inside the box that is on or has
// a value other than 0 (a setValueOff of 0 on the background tile changes
// nothing); from there every inside voxel keeps value and state, earlier ones
// stay (0, off). Sparse: inside active voxels become (1.0, on). The leaf's
// pool index is appended to `out` (order arbitrary; the host sorts).
__global__ __launch_bounds__(256) void k_section_pick(MapView mp, uint32_t leaf_count, const int* bbox,
                                                      int full, uint32_t* __restrict__ out,
                                                      unsigned int* counter, TileView tv) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const int lane = lane_id();
  __shared__ int sb[6];
  if (threadIdx.x < 6) sb[threadIdx.x] = bbox[threadIdx.x];
  __syncthreads();
  for (uint32_t L = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); L < leaf_count; L += waves) {
    if (tile_at(tv, L)) continue;  // the leaf iteration of getMapSection (VM:937) skips tiles
    int o[3];
    key_origin(mp.leaf_keys[L], o);
    if (o[0] + 7 < sb[0] || o[0] > sb[3] || o[1] + 7 < sb[1] || o[1] > sb[4] || o[2] + 7 < sb[2] ||
        o[2] > sb[5])
      continue;
    bool any = false;
    if (lane < 8) {
      const uint64_t in = box_word(o, sb, lane);
      const uint64_t act = mp.active[(size_t)L * 8 + lane];
      any = (in & act) != 0;
      if (full && !any && in) {
        const float* v = mp.values + (size_t)L * 512 + lane * 64;
        for (uint64_t b = in; b; b &= b - 1)
          if (v[__builtin_ctzll(b)] != 0.0f) {
            any = true;
            break;
          }
      }
    }
    if (__ballot(any) != 0 && lane == 0) out[atomicAdd(counter, 1u)] = L;
  }
}

// One wave per picked leaf: its section values and masks (see k_section_pick).
__global__ __launch_bounds__(256) void k_section_gather(MapView mp, const uint32_t* __restrict__ list,
                                                        uint32_t count, const int* bbox, int full,
                                                        unsigned long long* keys_out, float* vals_out,
                                                        unsigned long long* act_out) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const int lane = lane_id();
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < count; i += waves) {
    const uint32_t L = list[i];
    const uint64_t key = mp.leaf_keys[L];
    int o[3];
    key_origin(key, o);
    // first creating voxel (full grid): lowest offset inside the box that is on or non-zero
    int first = 512;
    uint64_t in = 0, act = 0;
    if (lane < 8) {
      in = box_word(o, bbox, lane);
      act = mp.active[(size_t)L * 8 + lane];
      act_out[(size_t)i * 8 + lane] = in & act;
      if (full) {
        uint64_t cre = in & act;
        const float* v = mp.values + (size_t)L * 512 + lane * 64;
        for (uint64_t b = in & ~act; b; b &= b - 1)
          if (v[__builtin_ctzll(b)] != 0.0f) cre |= b & (~b + 1);
        if (cre) first = lane * 64 + __builtin_ctzll(cre);
      }
    }
#pragma unroll
    for (int s = 1; s < 8; s <<= 1) first = min(first, __shfl_xor(first, s, 64));
    first = __shfl(first, 0, 64);  // lanes 0-7 hold the minimum; broadcast it
    if (lane == 0) keys_out[i] = key;
    for (int n = lane; n < 512; n += 64) {
      const uint64_t w = __shfl(in, n >> 6, 64), aw = __shfl(act, n >> 6, 64);
      const bool inside = (w >> (n & 63)) & 1, on = (aw >> (n & 63)) & 1;
      float x = 0.0f;
      if (full)
        x = inside && n >= first ? mp.values[(size_t)L * 512 + n] : 0.0f;
      else
        x = inside && on ? 1.0f : 0.0f;
      vals_out[(size_t)i * 512 + n] = x;
    }
  }
}

// One lane per record: find the map leaf of keys[i], or create it when
// need[i] is set (leaf count in st->leaf_count; never past the pool, else
// st->map_full). Records that create nothing leave kNone in eidx. With
// `first`, a record whose leaf already existed gets first[i] = 0 (all of its
// voxels are written).
__global__ __launch_bounds__(256) void k_upsert_leaves(const unsigned long long* __restrict__ keys,
                                                       const uint8_t* __restrict__ need, uint32_t n,
                                                       MapView mp, uint32_t* __restrict__ eidx,
                                                       uint16_t* __restrict__ first, Status* st) {
  __shared__ uint32_t s_wave[4], s_base;
  const uint32_t G = gridDim.x * blockDim.x;
  const uint32_t rounds = (n + G - 1) / G;
  for (uint32_t rd = 0; rd < rounds; ++rd) {
    const uint32_t e = rd * G + blockIdx.x * blockDim.x + threadIdx.x;
    bool created = false;
    uint32_t h = kNone, idx = kNone;
    uint64_t key = 0;
    if (e < n) {
      key = keys[e];
      const bool nd = need[e] != 0;
      uint32_t p = (uint32_t)fmix64(key) & mp.cap_mask;
      for (uint32_t probe = 0; probe <= mp.cap_mask; ++probe) {
        const unsigned long long k = mp.keys[p];
        if (k == key) {
          idx = mp.idx[p];
          break;
        }
        if (k == kEmpty) {
          if (!nd) break;
          const unsigned long long prev = atomicCAS(&mp.keys[p], kEmpty, (unsigned long long)key);
          if (prev == kEmpty) {
            created = true;
            h = p;
            break;
          }
        }
        p = (p + 1) & mp.cap_mask;
      }
    }
    const uint32_t pos = block_append(created, &st->leaf_count, s_wave, &s_base);
    if (created) {
      if (pos < mp.lcap) {
        idx = pos;
        mp.idx[h] = idx;
        mp.leaf_keys[idx] = key;
      } else {
        mp.idx[h] = kNone;
        atomicOr(&st->map_full, 1u);
      }
    }
    if (e < n) {
      eidx[e] = idx;
      if (idx != kNone) mp.dirty[idx] = 1;
      if (first && idx != kNone && !created) first[e] = 0;
    }
  }
}

// One wave per section leaf: write its voxels into the map leaf eidx[i].
// A leaf that existed before gets all 512 voxels; a created one (first[i] >
// 0) keeps the background for the voxels before its creating voxel.
__global__ __launch_bounds__(256) void k_section_write(const float* __restrict__ vals,
                                                       const unsigned long long* __restrict__ act,
                                                       const uint32_t* __restrict__ eidx,
                                                       const uint16_t* __restrict__ first, uint32_t n,
                                                       MapView mp) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const int lane = lane_id();
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    const uint32_t idx = eidx[i];
    if (idx == kNone) continue;
    const int f = first[i];
    for (int k = lane; k < 512; k += 64)
      if (k >= f) mp.values[(size_t)idx * 512 + k] = vals[(size_t)i * 512 + k];
    if (lane < 8) {
      const uint64_t keep = f >= (lane + 1) * 64 ? ~0ull : (f <= lane * 64 ? 0ull : ((1ull << (f - lane * 64)) - 1));
      unsigned long long* pa = mp.active + (size_t)idx * 8 + lane;
      *pa = (*pa & keep) | (act[(size_t)i * 8 + lane] & ~keep);
    }
  }
}

// The inactive tiles of the section's internal nodes, each visited at its
// origin: 4096 8^3 slots per 128^3 node (slots holding a section leaf are
// children, skipped), 32768 128^3 slots per 4096^3 node (slots holding a
// 128^3 node skipped). A visit sets the voxel at the slot origin of an
// existing map leaf to (0, off).
__global__ __launch_bounds__(256) void k_section_tiles(const unsigned long long* __restrict__ n128,
                                                       uint32_t c128, const unsigned long long* __restrict__ n4096,
                                                       uint32_t c4096, const unsigned long long* __restrict__ sleaf,
                                                       uint32_t sleaf_mask, const unsigned long long* __restrict__ s128,
                                                       uint32_t s128_mask, MapView mp) {
  const uint64_t total = (uint64_t)c128 * 4096 + (uint64_t)c4096 * 32768;
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += G) {
    int o[3], x, y, z;
    if (t < (uint64_t)c128 * 4096) {
      key_origin(n128[t >> 12], o);  // origin / 128 * 8
      const int j = (int)(t & 4095);
      x = o[0] * 16 + 8 * (j >> 8);
      y = o[1] * 16 + 8 * ((j >> 4) & 15);
      z = o[2] * 16 + 8 * (j & 15);
      if (set_has(sleaf, sleaf_mask, leaf_key(x, y, z))) continue;
    } else {
      const uint64_t u = t - (uint64_t)c128 * 4096;
      key_origin(n4096[u >> 15], o);  // origin / 4096 * 8
      const int j = (int)(u & 32767);
      x = o[0] * 512 + 128 * (j >> 10);
      y = o[1] * 512 + 128 * ((j >> 5) & 31);
      z = o[2] * 512 + 128 * (j & 31);
      if (set_has(s128, s128_mask, node_key(x, y, z, 7))) continue;
    }
    const uint64_t key = leaf_key(x, y, z);
    uint32_t p = (uint32_t)fmix64(key) & mp.cap_mask;
    for (uint32_t probe = 0; probe <= mp.cap_mask; ++probe) {
      const unsigned long long k = mp.keys[p];
      if (k == key) {
        const uint32_t idx = mp.idx[p];
        if (idx != kNone) {
          mp.values[(size_t)idx * 512] = 0.0f;  // voxel offset 0: the slot origin
          atomicAnd(&mp.active[(size_t)idx * 8], ~1ull);
          mp.dirty[idx] = 1;
        }
        break;
      }
      if (k == kEmpty) break;
      p = (p + 1) & mp.cap_mask;
    }
  }
}

// applyMapSectionUpdateGrid, first loop: every active map voxel inside the box
// set inactive. One lane per (leaf, word).
__global__ __launch_bounds__(256) void k_section_clear(MapView mp, uint32_t leaf_count, const int* bbox,
                                                       TileView tv) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)leaf_count * 8; t += G) {
    const uint32_t L = (uint32_t)(t >> 3);
    const int j = (int)(t & 7);
    int o[3];
    key_origin(mp.leaf_keys[L], o);
    const uint64_t in = box_word(o, bbox, j);
    if (tile_at(tv, L)) {
      // cbeginValueOn (VM:1072) yields an active tile once, at its origin: only that
      // voxel is switched off, which turns the tile into a leaf
      if (j == 0 && (in & 1u) && tile_active(tv, L)) {
        mp.active[(size_t)L * 8] &= ~1ull;
        mp.dirty[L] = 1;
        atomicAnd(&tv.bits[L >> 5], ~(1u << (L & 31)));
      }
      continue;
    }
    if (!in) continue;
    unsigned long long* pa = mp.active + (size_t)L * 8 + j;
    const uint64_t aw = *pa;
    if (aw & in) {
      *pa = aw & ~in;
      mp.dirty[L] = 1;
    }
  }
}

// Second loop: OR the section's active masks into the map leaves eidx[i].
__global__ __launch_bounds__(256) void k_section_or(const unsigned long long* __restrict__ act,
                                                    const uint32_t* __restrict__ eidx, uint32_t n,
                                                    MapView mp) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * 8; t += G) {
    const uint32_t i = (uint32_t)(t >> 3);
    const uint32_t idx = eidx[i];
    if (idx == kNone) continue;
    mp.active[(size_t)idx * 8 + (t & 7)] |= act[t];
  }
}

// One wave per section leaf (origins as x, y, z triples): its leaf key, and
// the offset of its creating voxel, the first voxel that is on or (with
// values) non-zero (512: none, applying it creates no missing leaf). With
// topology sets: inserts the leaf key, and the 128^3 and 4096^3 node keys,
// each new node key appended to its list (counters[0], counters[1]).
__global__ __launch_bounds__(256) void k_section_prep(const int* __restrict__ origins,
                                                      const float* __restrict__ vals,
                                                      const unsigned long long* __restrict__ act,
                                                      uint32_t n, unsigned long long* __restrict__ keys,
                                                      uint8_t* __restrict__ need, uint16_t* __restrict__ first,
                                                      unsigned long long* sleaf, uint32_t sleaf_mask,
                                                      unsigned long long* s128, uint32_t s128_mask,
                                                      unsigned long long* n128, unsigned long long* s4096,
                                                      uint32_t s4096_mask, unsigned long long* n4096,
                                                      unsigned int* counters) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  const int lane = lane_id();
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    // lane owns voxels 8 lane .. 8 lane + 7
    uint32_t cre = (uint32_t)(act[(size_t)i * 8 + (lane >> 3)] >> ((lane & 7) * 8)) & 0xFFu;
    if (vals) {
      const float4* pv = reinterpret_cast<const float4*>(vals + (size_t)i * 512 + lane * 8);
      const float4 a = pv[0], b = pv[1];
      cre |= (a.x != 0.0f ? 1u : 0u) | (a.y != 0.0f ? 2u : 0u) | (a.z != 0.0f ? 4u : 0u) |
             (a.w != 0.0f ? 8u : 0u) | (b.x != 0.0f ? 16u : 0u) | (b.y != 0.0f ? 32u : 0u) |
             (b.z != 0.0f ? 64u : 0u) | (b.w != 0.0f ? 128u : 0u);
    }
    int f = cre ? lane * 8 + __builtin_ctz(cre) : 512;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) f = min(f, __shfl_xor(f, o, 64));
    if (lane == 0) {
      const int x = origins[3 * i], y = origins[3 * i + 1], z = origins[3 * i + 2];
      const uint64_t key = leaf_key(x, y, z);
      keys[i] = key;
      need[i] = f < 512 ? 1 : 0;
      first[i] = (uint16_t)f;
      if (sleaf) {
        set_insert(sleaf, sleaf_mask, key);
        const uint64_t k7 = node_key(x, y, z, 7), k12 = node_key(x, y, z, 12);
        if (set_insert(s128, s128_mask, k7)) n128[atomicAdd(&counters[0], 1u)] = k7;
        if (set_insert(s4096, s4096_mask, k12)) n4096[atomicAdd(&counters[1], 1u)] = k12;
      }
    }
  }
}

// evalActiveVoxelBoundingBox: min / max voxel coordinate of the active
// voxels, folded into box[0..2] (atomicMin) and box[3..5] (atomicMax). One
// lane per (leaf, word).
__global__ __launch_bounds__(256) void k_active_bbox(MapView mp, uint32_t leaf_count, int* box) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)leaf_count * 8; t += G) {
    const uint32_t L = (uint32_t)(t >> 3);
    const int j = (int)(t & 7);
    const uint64_t a = mp.active[(size_t)L * 8 + j];
    if (!a) continue;
    int o[3];
    key_origin(mp.leaf_keys[L], o);
    uint32_t rows = 0, cols = 0;  // y rows and z columns present in the word
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      if ((a >> (8 * k)) & 0xFFu) rows |= 1u << k;
      cols |= (uint32_t)(a >> (8 * k)) & 0xFFu;
    }
    lo[0] = min(lo[0], o[0] + j);
    hi[0] = max(hi[0], o[0] + j);
    lo[1] = min(lo[1], o[1] + __builtin_ctz(rows));
    hi[1] = max(hi[1], o[1] + 31 - __builtin_clz(rows));
    lo[2] = min(lo[2], o[2] + __builtin_ctz(cols));
    hi[2] = max(hi[2], o[2] + 31 - __builtin_clz(cols));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], o, 64));
      hi[a] = max(hi[a], __shfl_xor(hi[a], o, 64));
    }
  }
  if (lane_id() == 0 && lo[0] != INT_MAX) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      atomicMin(&box[a], lo[a]);
      atomicMax(&box[3 + a], hi[a]);
    }
  }
}

// Keys of every leaf overlapping the box (leaf-grid origin lo, extents dims,
// x-major), each with the same creation flag.
__global__ __launch_bounds__(256) void k_box_keys(int lx, int ly, int lz, uint32_t dx, uint32_t dy, uint32_t dz,
                                                  unsigned long long* __restrict__ keys, uint8_t* __restrict__ need,
                                                  uint8_t need_v) {
  const uint64_t n = (uint64_t)dx * dy * dz;
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += G) {
    const uint32_t z = (uint32_t)(i % dz), y = (uint32_t)((i / dz) % dy), x = (uint32_t)(i / ((uint64_t)dz * dy));
    keys[i] = leaf_key((lx + (int)x) * 8, (ly + (int)y) * 8, (lz + (int)z) * 8);
    need[i] = need_v;
  }
}

// The background of createMapFromPointCloud: inside the box, every inactive
// voxel of the leaves eidx[i] gets value v (its state stays off). One lane
// per (leaf, word).
__global__ __launch_bounds__(256) void k_box_fill(const unsigned long long* __restrict__ keys,
                                                  const uint32_t* __restrict__ eidx, uint32_t n, const int* box,
                                                  MapView mp, float v) {
  const uint64_t G = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (uint64_t)n * 8; t += G) {
    const uint32_t i = (uint32_t)(t >> 3);
    const int j = (int)(t & 7);
    const uint32_t idx = eidx[i];
    if (idx == kNone) continue;
    int o[3];
    key_origin(keys[i], o);
    const uint64_t m = box_word(o, box, j) & ~(uint64_t)mp.active[(size_t)idx * 8 + j];
    if (!m) continue;
    float* pv = mp.values + (size_t)idx * 512 + j * 64;
    for (uint64_t b = m; b; b &= b - 1) pv[__builtin_ctzll(b)] = v;
    mp.dirty[idx] = 1;
  }
}

// ------------------------------------------------------------- launchers --

hipError_t launch_mark_points(hipStream_t s, const float* pts, int64_t stride_bytes, uint32_t n, double inv,
                              UpdView u, Status* st) {
  const int grid = (int)std::min<uint32_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_mark_points, dim3(grid), dim3(256), 0, s, pts, stride_bytes / 4, n, inv, u, st);
  return hipGetLastError();
}

hipError_t launch_mark_walls(hipStream_t s, const double* verts, const int4* walls, const uint64_t* prefix,
                             uint32_t n_walls, uint64_t total, double res, double inv, const double* tbl,
                             uint32_t tbl_n, UpdView u, Status* st) {
  if (total == 0 || n_walls == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(16384, (total + 63) / 64);
  hipLaunchKernelGGL(k_mark_walls, dim3(grid), dim3(64), 0, s, verts, walls, prefix, n_walls, res, inv, tbl,
                     tbl_n, u, st);
  return hipGetLastError();
}

hipError_t launch_set_voxels(hipStream_t s, UpdView u, const uint32_t* list, const uint32_t* eidx,
                             uint32_t count, MapView mp, int op, float val, float t_max, bool clear) {
  const int grid = (int)std::min<uint64_t>(8192, ((uint64_t)count * 8 + 255) / 256);
  hipLaunchKernelGGL(k_set_voxels, dim3(grid), dim3(256), 0, s, u, list, eidx, count, mp, op, val, t_max,
                     clear ? 1u : 0u);
  return hipGetLastError();
}

hipError_t launch_section_pick(hipStream_t s, MapView mp, uint32_t leaf_count, const int* bbox, bool full,
                               uint32_t* out, unsigned int* counter, TileView tv) {
  const int grid = (int)std::min<uint32_t>(8192, (leaf_count + 3) / 4);
  hipLaunchKernelGGL(k_section_pick, dim3(grid), dim3(256), 0, s, mp, leaf_count, bbox, full ? 1 : 0, out,
                     counter, tv);
  return hipGetLastError();
}

hipError_t launch_section_gather(hipStream_t s, MapView mp, const uint32_t* list, uint32_t count,
                                 const int* bbox, bool full, unsigned long long* keys_out, float* vals_out,
                                 unsigned long long* act_out) {
  const int grid = (int)std::min<uint32_t>(8192, (count + 3) / 4);
  hipLaunchKernelGGL(k_section_gather, dim3(grid), dim3(256), 0, s, mp, list, count, bbox, full ? 1 : 0,
                     keys_out, vals_out, act_out);
  return hipGetLastError();
}

hipError_t launch_upsert_leaves(hipStream_t s, const unsigned long long* keys, const uint8_t* need, uint32_t n,
                                MapView mp, uint32_t* eidx, uint16_t* first, Status* st) {
  const int grid = (int)std::min<uint32_t>(4096, (n + 255) / 256);
  hipLaunchKernelGGL(k_upsert_leaves, dim3(grid), dim3(256), 0, s, keys, need, n, mp, eidx, first, st);
  return hipGetLastError();
}

hipError_t launch_section_write(hipStream_t s, const float* vals, const unsigned long long* act,
                                const uint32_t* eidx, const uint16_t* first, uint32_t n, MapView mp) {
  const int grid = (int)std::min<uint32_t>(8192, (n + 3) / 4);
  hipLaunchKernelGGL(k_section_write, dim3(grid), dim3(256), 0, s, vals, act, eidx, first, n, mp);
  return hipGetLastError();
}

hipError_t launch_section_tiles(hipStream_t s, const unsigned long long* n128, uint32_t c128,
                                const unsigned long long* n4096, uint32_t c4096,
                                const unsigned long long* sleaf, uint32_t sleaf_mask,
                                const unsigned long long* s128, uint32_t s128_mask, MapView mp) {
  const uint64_t total = (uint64_t)c128 * 4096 + (uint64_t)c4096 * 32768;
  if (total == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(8192, (total + 255) / 256);
  hipLaunchKernelGGL(k_section_tiles, dim3(grid), dim3(256), 0, s, n128, c128, n4096, c4096, sleaf, sleaf_mask,
                     s128, s128_mask, mp);
  return hipGetLastError();
}

hipError_t launch_section_clear(hipStream_t s, MapView mp, uint32_t leaf_count, const int* bbox, TileView tv) {
  if (leaf_count == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(8192, ((uint64_t)leaf_count * 8 + 255) / 256);
  hipLaunchKernelGGL(k_section_clear, dim3(grid), dim3(256), 0, s, mp, leaf_count, bbox, tv);
  return hipGetLastError();
}

hipError_t launch_section_or(hipStream_t s, const unsigned long long* act, const uint32_t* eidx, uint32_t n,
                             MapView mp) {
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(8192, ((uint64_t)n * 8 + 255) / 256);
  hipLaunchKernelGGL(k_section_or, dim3(grid), dim3(256), 0, s, act, eidx, n, mp);
  return hipGetLastError();
}

hipError_t launch_section_prep(hipStream_t s, const int* origins, const float* vals, const unsigned long long* act,
                               uint32_t n, unsigned long long* keys, uint8_t* need, uint16_t* first,
                               unsigned long long* sleaf, uint32_t sleaf_mask, unsigned long long* s128,
                               uint32_t s128_mask, unsigned long long* n128, unsigned long long* s4096,
                               uint32_t s4096_mask, unsigned long long* n4096, unsigned int* counters) {
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<uint32_t>(8192, (n + 3) / 4);
  hipLaunchKernelGGL(k_section_prep, dim3(grid), dim3(256), 0, s, origins, vals, act, n, keys, need, first,
                     sleaf, sleaf_mask, s128, s128_mask, n128, s4096, s4096_mask, n4096, counters);
  return hipGetLastError();
}

hipError_t launch_active_bbox(hipStream_t s, MapView mp, uint32_t leaf_count, int* box) {
  if (leaf_count == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(4096, ((uint64_t)leaf_count * 8 + 255) / 256);
  hipLaunchKernelGGL(k_active_bbox, dim3(grid), dim3(256), 0, s, mp, leaf_count, box);
  return hipGetLastError();
}

hipError_t launch_box_keys(hipStream_t s, const int lo[3], const uint32_t dims[3], unsigned long long* keys,
                           uint8_t* need, bool need_v) {
  const uint64_t n = (uint64_t)dims[0] * dims[1] * dims[2];
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(8192, (n + 255) / 256);
  hipLaunchKernelGGL(k_box_keys, dim3(grid), dim3(256), 0, s, lo[0], lo[1], lo[2], dims[0], dims[1], dims[2], keys,
                     need, (uint8_t)(need_v ? 1 : 0));
  return hipGetLastError();
}

hipError_t launch_box_fill(hipStream_t s, const unsigned long long* keys, const uint32_t* eidx, uint32_t n,
                           const int* box, MapView mp, float v) {
  if (n == 0) return hipSuccess;
  const int grid = (int)std::min<uint64_t>(8192, ((uint64_t)n * 8 + 255) / 256);
  hipLaunchKernelGGL(k_box_fill, dim3(grid), dim3(256), 0, s, keys, eidx, n, box, mp, v);
  return hipGetLastError();
}

}  // namespace vdbgpu
