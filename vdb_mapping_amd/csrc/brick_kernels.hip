// gfx950 kernels of the brick path of one scan (the raycast of
// insertPointCloud, VDBMapping.hpp:466-566). Space is cut into bricks of
// 32^3 voxels (4^3 leaves). Each ray is cut into per-brick segments in
// closed form (dda_math.h), the segments are bucketed by brick, and one
// workgroup owns a brick's visited/hit bitmap (8 KiB of LDS) while it walks
// that brick's segments. The result is the scan's update set as a list of
// leaf records (key + 8 visited + 8 hit mask words), every update leaf
// exactly once, written with plain stores at fixed positions (no shared
// counters); k_map_insert and k_apply take the dense list of filled
// positions built from the per-item leaf masks. No global atomics touch voxel bits except for the few
// bricks so crowded (more than kItemSegs segments, around the sensor) that
// several workgroups share them.
//
//   k_seg_count     one lane per ray: enumerate its bricks, count segments
//                   per brick (aggregated per workgroup in an LDS table,
//                   which is stored for the fill)
//   k_brick_sums    per 256 brick slots: sums of segments and work items
//   k_scan_u64      exclusive scan of those sums (one 256-thread workgroup)
//   k_brick_offsets segment offset and work items of each brick
//   k_seg_fill      reserve the run of each entry of the workgroup's stored
//                   table, enumerate again, write each segment record at its
//                   brick's next position
//   k_brick_walk    one workgroup per work item: walk the item's segments
//                   into the LDS bitmap, write the brick's leaf records at
//                   the item's 64 fixed positions, or OR into the shared
//                   bitmap of a split brick
//   k_split_emit    leaf records of split bricks (positions after the items)
//   k_mask_sums, k_mask_list  dense list of the filled record positions from
//                   the per-item leaf masks (scan, no shared counters)
//   k_rec_gather    the listed records copied densely into the scan's record
//                   set, named in device memory (read by the map stream,
//                   possibly several scans later)
//
// Segment record (uint4), self-contained so the walk loads nothing else:
//   x = n0 | n1 << 14, y = n2 | c0 << 14, z = c1 | c2 << 14  (|delta| per axis
//       and the crossing counts at the segment's first visit, <= 8192 each)
//   w = brick-local first voxel lx | ly << 5 | lz << 10 | visits m << 16 |
//       negative-step bits x, y, z << 24 | end-voxel hit flag << 31
// The walk of a segment starts at that voxel with the integer DDA state of
// those counts and visits m (<= 94) voxels, all inside the brick (checked).
#include "dda_math.h"
#include "device_common.h"
#include "kernels.h"

namespace vdbgpu {

namespace {

constexpr int kLdsBricks = kSegTab;  // per-workgroup brick aggregation table
constexpr int kLdsBrickProbe = 32;
constexpr uint32_t kMaxBrickProbe = 512;
constexpr int kBrickWords = 1024;  // 64 leaves x (8 visited + 8 hit) words
// The walk's LDS bitmap: [0, 512) visited words of layout X, [512, 1024) hit
// words (leaf-major, as emitted), [1024, 1536) layout Y, [1536, 2048) layout Z.
// A segment marks its voxels in the layout whose word axis it steps along
// least (see walk_segment); the layouts are merged into leaf-major visited
// words before the records are written.
constexpr int kWalkWords = 2048;

__device__ __forceinline__ int bw_lane() { return (int)__lane_id(); }

struct RayGeom {
  int s[3], e[3];
  uint32_t cnt;
  bool hit;
};

// Geometry of ray i, false if it marks nothing (NaN, zero range, rejected).
__device__ __forceinline__ bool ray_geom(const int4* __restrict__ rayinfo,
                                         const ScanInfo* __restrict__ scans, uint32_t i,
                                         RayGeom& g) {
  const int4 r = rayinfo[i];
  const uint32_t w = (uint32_t)r.w;
  g.cnt = w & kCntMask;
  g.hit = (w & kHitBit) != 0;
  if (g.cnt == 0 && !g.hit) return false;
  const uint32_t sc = (w >> kScanShift) & kScanMask;
  g.s[0] = scans[sc].s[0];
  g.s[1] = scans[sc].s[1];
  g.s[2] = scans[sc].s[2];
  g.e[0] = r.x;
  g.e[1] = r.y;
  g.e[2] = r.z;
  return true;
}

__device__ __forceinline__ int lds_brick(unsigned long long* lkey, uint64_t key, bool insert) {
  uint32_t h = (uint32_t)(fmix64(key) >> 40) & (kLdsBricks - 1);
  for (int p = 0; p < kLdsBrickProbe; ++p) {
    const unsigned long long k =
        __hip_atomic_load(&lkey[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k == key) return (int)h;
    if (k == kEmpty) {
      if (!insert) return -1;
      const unsigned long long prev = atomicCAS(&lkey[h], kEmpty, (unsigned long long)key);
      if (prev == kEmpty || prev == key) return (int)h;
    }
    h = (h + 1) & (kLdsBricks - 1);
  }
  return -1;
}

// Slot of a brick in the scan's brick table (inserted if `insert`), kNone if
// absent or after kMaxBrickProbe probes (then the table is flagged full).
__device__ __forceinline__ uint32_t brick_slot(const BrickView& bv, uint64_t key, bool insert,
                                               unsigned int* ovf) {
  uint32_t h = (uint32_t)fmix64(key) & bv.cap_mask;
  const uint32_t limit = bv.cap_mask < kMaxBrickProbe ? bv.cap_mask + 1 : kMaxBrickProbe;
  for (uint32_t p = 0; p < limit; ++p) {
    const unsigned long long k =
        __hip_atomic_load(&bv.keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return h;
    if (k == kEmpty) {
      if (!insert) return kNone;
      const unsigned long long prev = atomicCAS(&bv.keys[h], kEmpty, (unsigned long long)key);
      if (prev == kEmpty || prev == key) return h;
    }
    h = (h + 1) & bv.cap_mask;
  }
  if (insert) atomicOr(ovf, kOvfBricks);
  return kNone;
}

// Workgroup-uniform stop flag: error or any overflow already raised.
__device__ __forceinline__ bool stopped(const Status* st, unsigned int* s_stop) {
  if (threadIdx.x == 0)
    *s_stop = __hip_atomic_load(&st->error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) |
              __hip_atomic_load(&st->overflow, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __syncthreads();
  return *s_stop != 0;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ uint64_t block_sum_u64(uint64_t v, uint64_t* s_w) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (bw_lane() == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  uint64_t t = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += s_w[w];
  __syncthreads();
  return t;
}

// Leaf records of the brick bitmap `bits` (brick origin o) at the fixed
// positions base .. base + 63 (leaf l of the brick at base + l), for every
// leaf with a visited voxel; *leaf_mask gets the set of filled positions
// (bit l). Called by the whole workgroup.
__device__ void emit_leaves(const unsigned long long* bits, int ox, int oy, int oz, UpdView rec,
                            uint32_t base, uint32_t* s_pos, unsigned long long* leaf_mask) {
  const int tid = threadIdx.x;
  if (tid < 64) {  // lane l = leaf l of the brick
    uint64_t any = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) any |= bits[tid * 8 + j];
    const uint64_t m = __ballot(any != 0);
    if (tid == 0) *leaf_mask = m;
    s_pos[tid] = any ? base + tid : kNone;
    if (any) rec.keys[base + tid] = leaf_key(ox + (tid >> 4) * 8, oy + ((tid >> 2) & 3) * 8, oz + (tid & 3) * 8);
  }
  __syncthreads();
  // 16 records x 16 words per pass: contiguous 128-byte rows
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int l = p * 16 + (tid >> 4), w = tid & 15;
    const uint32_t pos = s_pos[l];
    if (pos != kNone) rec.masks[(size_t)pos * 16 + w] = bits[(w < 8 ? 0 : 512) + l * 8 + (w & 7)];
  }
}

}  // namespace

__global__ __launch_bounds__(256) void k_seg_count(const int4* __restrict__ rayinfo,
                                                   const ScanInfo* __restrict__ scans,
                                                   const BatchCtl* __restrict__ ctl,
                                                   const double* __restrict__ tbl, BrickView bv,
                                                   unsigned long long* __restrict__ wkey,
                                                   unsigned long long* __restrict__ wsc, Status* st) {
  __shared__ unsigned long long lkey[kLdsBricks];
  __shared__ uint32_t lcnt[kLdsBricks];
  __shared__ unsigned int s_stop;
  if (stopped(st, &s_stop)) return;
  for (int e = threadIdx.x; e < kLdsBricks; e += blockDim.x) {
    lkey[e] = kEmpty;
    lcnt[e] = 0;
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t nseg = 0;
  RayGeom g;
  if (i < ctl->n_rays && ray_geom(rayinfo, scans, i, g)) {
    SegEnum en;
    en.init(tbl, g.s, g.e, g.cnt);
    Segment sg;
    while (en.next(tbl, sg)) {
      const uint64_t key = brick_key(sg.bx, sg.by, sg.bz);
      const int ls = lds_brick(lkey, key, true);
      if (ls >= 0) {
        atomicAdd(&lcnt[ls], 1u);
      } else {
        const uint32_t gs = brick_slot(bv, key, true, &st->overflow);
        if (gs != kNone) atomicAdd(&bv.cnt[gs], 1u);
      }
      ++nseg;
    }
  }
  __syncthreads();
  // the table, as it stands, for the fill: a segment is counted here iff its
  // brick is in the table (a key once in the table stays; one that found its
  // probe run full never gets in), so the fill's lookups in the same table
  // split the workgroup's segments exactly as the counts did
  const size_t tb = (size_t)blockIdx.x * kLdsBricks;
  for (int e = threadIdx.x; e < kLdsBricks; e += blockDim.x) {
    const uint64_t key = lkey[e];
    wkey[tb + e] = key;
    if (key == kEmpty) continue;
    const uint32_t gs = brick_slot(bv, key, true, &st->overflow);
    if (gs != kNone) atomicAdd(&bv.cnt[gs], lcnt[e]);
    wsc[tb + e] = ((uint64_t)gs << 32) | lcnt[e];
  }
  nseg = wave_sum_u32(nseg);
  if (bw_lane() == 0 && nseg) atomicAdd(&st->segs, nseg);
}

// Per 256 slots: (work items << 32) | segments.
__global__ __launch_bounds__(256) void k_brick_sums(BrickView bv, uint64_t* __restrict__ bsum,
                                                    const Status* st) {
  __shared__ uint64_t s_w[4];
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = bv.cnt[slot];
  const uint64_t v = ((uint64_t)((c + kItemSegs - 1) / kItemSegs) << 32) | c;
  const uint64_t t = block_sum_u64(v, s_w);
  if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}

// Exclusive scan of nb packed sums (one workgroup of 1024); totals to st.
__global__ __launch_bounds__(256) void k_scan_u64(const uint64_t* __restrict__ bsum,
                                                  uint64_t* __restrict__ boff, uint32_t nb,
                                                  unsigned int* total, int total_shift) {
  __shared__ uint64_t wsum[4];
  uint64_t carry = 0;
  for (uint32_t base = 0; base < nb; base += 256) {
    const uint32_t i = base + threadIdx.x;
    const uint64_t x = i < nb ? bsum[i] : 0ull;
    uint64_t v = x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint64_t t = __shfl_up(v, o, 64);
      if (bw_lane() >= o) v += t;
    }
    const int wv = threadIdx.x >> 6;
    if (bw_lane() == 63) wsum[wv] = v;
    __syncthreads();
    uint64_t add = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wv) add += wsum[w];
      tot += wsum[w];
    }
    if (i < nb) boff[i] = carry + add + v - x;
    carry += tot;
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = (unsigned int)(carry >> total_shift);
}

__global__ __launch_bounds__(256) void k_brick_offsets(BrickView bv, const uint64_t* __restrict__ boff,
                                                       BrickItem* __restrict__ items,
                                                       uint32_t seg_cap, uint32_t item_cap,
                                                       uint32_t* __restrict__ split_slot,
                                                       uint32_t split_cap, Status* st) {
  __shared__ uint64_t s_w[4];
  const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t c = bv.cnt[slot];
  const uint32_t ni = (c + kItemSegs - 1) / kItemSegs;
  const uint64_t x = ((uint64_t)ni << 32) | c;
  uint64_t v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(v, o, 64);
    if (bw_lane() >= o) v += t;
  }
  const int wv = threadIdx.x >> 6;
  if (bw_lane() == 63) s_w[wv] = v;
  __syncthreads();
  uint64_t add = 0;
  for (int w = 0; w < wv; ++w) add += s_w[w];
  const uint64_t off = boff[blockIdx.x] + add + v - x;
  const uint32_t seg0 = (uint32_t)off, item0 = (uint32_t)(off >> 32);
  bv.cur[slot] = seg0;
  if (ni == 0) return;
  if ((uint64_t)seg0 + c > seg_cap) {
    atomicOr(&st->overflow, kOvfSegs);
    return;
  }
  if ((uint64_t)item0 + ni > item_cap) {
    atomicOr(&st->overflow, kOvfItems);
    return;
  }
  uint32_t shared = kNone;
  if (ni > 1) {
    shared = atomicAdd(&st->splits, 1u);
    if (shared >= split_cap) {
      atomicOr(&st->overflow, kOvfSplit);
      return;
    }
    split_slot[shared] = slot;
  }
  for (uint32_t j = 0; j < ni; ++j) {
    const uint32_t a = seg0 + j * kItemSegs;
    const uint32_t b = min(a + kItemSegs, seg0 + c);
    items[item0 + j] = BrickItem{slot, a, b, shared};
  }
}

__global__ __launch_bounds__(256) void k_seg_fill(const int4* __restrict__ rayinfo,
                                                  const ScanInfo* __restrict__ scans,
                                                  const BatchCtl* __restrict__ ctl,
                                                  const double* __restrict__ tbl, BrickView bv,
                                                  const unsigned long long* __restrict__ wkey,
                                                  const unsigned long long* __restrict__ wsc,
                                                  uint4* __restrict__ segs, uint32_t seg_cap,
                                                  Status* st) {
  __shared__ unsigned long long lkey[kLdsBricks];
  __shared__ uint32_t lcnt[kLdsBricks];
  __shared__ uint32_t lbase[kLdsBricks];
  __shared__ unsigned int s_stop;
  // stops when the count stopped: a count workgroup that stored no table saw
  // the flags this one sees (nothing clears them between the two launches)
  if (stopped(st, &s_stop)) return;
  // 1: the count's table of this workgroup (the same rays), one reservation
  // per (workgroup, brick)
  const size_t tb = (size_t)blockIdx.x * kLdsBricks;
  for (int e = threadIdx.x; e < kLdsBricks; e += blockDim.x) {
    const uint64_t key = wkey[tb + e];
    lkey[e] = key;
    lcnt[e] = 0;
    if (key == kEmpty) continue;
    const uint64_t sc = wsc[tb + e];
    const uint32_t gs = (uint32_t)(sc >> 32);
    if (gs == kNone) {
      atomicOr(&st->error, kErrBrick);
      lbase[e] = seg_cap;
    } else {
      lbase[e] = atomicAdd(&bv.cur[gs], (uint32_t)sc);
    }
  }
  __syncthreads();
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  RayGeom g;
  const bool valid = i < ctl->n_rays && ray_geom(rayinfo, scans, i, g);
  // 2: write the records
  if (valid) {
    SegEnum en;
    en.init(tbl, g.s, g.e, g.cnt);
    Segment sg;
    while (en.next(tbl, sg)) {
      const uint64_t key = brick_key(sg.bx, sg.by, sg.bz);
      const int ls = lds_brick(lkey, key, false);
      uint32_t pos;
      if (ls >= 0) {
        pos = lbase[ls] + atomicAdd(&lcnt[ls], 1u);
      } else {
        const uint32_t gs = brick_slot(bv, key, false, &st->overflow);
        pos = gs == kNone ? seg_cap : atomicAdd(&bv.cur[gs], 1u);
      }
      if (pos >= seg_cap) {
        atomicOr(&st->overflow, kOvfSegs);
        continue;
      }
      uint32_t n[3], neg = 0, loc = 0;
      bool inside = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const int d = g.e[a] - g.s[a];
        n[a] = (uint32_t)abs(d);
        neg |= (d < 0 ? 1u : 0u) << a;
        const int b = a == 0 ? sg.bx : (a == 1 ? sg.by : sg.bz);
        const int l = g.s[a] + (d < 0 ? -(int)sg.c[a] : (int)sg.c[a]) - b * kBrickDim;
        inside = inside && (unsigned)l < (unsigned)kBrickDim;
        loc |= (uint32_t)(l & (kBrickDim - 1)) << (5 * a);
      }
      if (!inside) atomicOr(&st->error, kErrBrick);  // the cut and its brick disagree
      segs[pos] = make_uint4(n[0] | n[1] << 14, n[2] | sg.c[0] << 14, sg.c[1] | sg.c[2] << 14,
                             loc | sg.m << 16 | neg << 24 | ((sg.last && g.hit) ? 0x80000000u : 0u));
    }
  }
}

// Walk one segment record into the brick bitmap `bits`.
//
// Word layouts: a segment's voxels go into the layout whose word axis A is
// the axis it steps along least, so most steps stay inside one 64-bit word
// and the word is OR-ed into LDS only when the walk leaves it. With (a, b, c)
// the voxel's brick-local coordinates along (A, B, C) -- (x, y, z) for layout
// X, (y, x, z) for Y, (z, x, y) for Z -- the word is a * 16 + (b >> 3) * 4 +
// (c >> 3) and the bit (b & 7) * 8 + (c & 7). The walk state is one packed
// code u = a << 12 | b << 6 | c (6-bit fields; bit 5 of a field is a guard: a
// field that leaves [0, 31] sets it, and the OR of every code is checked at
// the end), stepped by one add per visit. The path is monotone along every
// axis, so it stays inside the brick when no field ever left it.
//
// The step order is the exact integer DDA of dda_math.h (signs of the three
// pairwise differences); the rational-tie rule runs only when a lane of the
// wave meets a tie, with the crossing counts recovered from the walk state.
__device__ __forceinline__ uint32_t walk_word(uint32_t u) {
  return ((u >> 8) & 0x1F0u) | ((u >> 7) & 0xCu) | ((u >> 3) & 3u);
}
__device__ __forceinline__ uint32_t walk_bit(uint32_t u) { return ((u >> 3) & 0x38u) | (u & 7u); }
constexpr uint32_t kWalkFields = 0x1F7DFu;  // the valid bits of a walk code

__device__ __forceinline__ void walk_segment(const uint4 sr, const double* __restrict__ tbl,
                                             unsigned long long* bits, uint32_t& used, Status* st) {
  const uint32_t n0 = sr.x & 0x3FFFu, n1 = (sr.x >> 14) & 0x3FFFu, n2 = sr.y & 0x3FFFu;
  const uint32_t c00 = (sr.y >> 14) & 0x3FFFu, c10 = sr.z & 0x3FFFu, c20 = (sr.z >> 14) & 0x3FFFu;
  const uint32_t lx0 = sr.w & 31u, ly0 = (sr.w >> 5) & 31u, lz0 = (sr.w >> 10) & 31u;
  const uint32_t m = (sr.w >> 16) & 0x7Fu;
  const int sgx = n0 == 0 ? 0 : (((sr.w >> 24) & 1u) ? -1 : 1);
  const int sgy = n1 == 0 ? 0 : (((sr.w >> 25) & 1u) ? -1 : 1);
  const int sgz = n2 == 0 ? 0 : (((sr.w >> 26) & 1u) ? -1 : 1);
  // word axis: the one with the fewest steps
  const int L = (n0 <= n1 && n0 <= n2) ? 0 : (n1 <= n2 ? 1 : 2);
  used |= 1u << L;
  const int shx = L == 0 ? 12 : 6, shy = L == 1 ? 12 : (L == 0 ? 6 : 0), shz = L == 2 ? 12 : 0;
  uint32_t u = lx0 << shx | ly0 << shy | lz0 << shz;
  const int ix = sgx * (1 << shx), iy = sgy * (1 << shy), iz = sgz * (1 << shz);
  const uint32_t base = L == 0 ? 0u : (L == 1 ? 1024u : 1536u);
  const int tn0 = 2 * (int)n0, tn1 = 2 * (int)n1, tn2 = 2 * (int)n2;
  int dyx = (int)((2 * c10 + 1) * n0) - (int)((2 * c00 + 1) * n1);
  int dzx = (int)((2 * c20 + 1) * n0) - (int)((2 * c00 + 1) * n2);
  int dzy = (int)((2 * c20 + 1) * n1) - (int)((2 * c10 + 1) * n2);
  uint32_t esc = u, cw = walk_word(u);
  uint64_t cb = 0;
  for (uint32_t k = 0;;) {
    const uint32_t w = walk_word(u);
    if (w != cw) {
      atomicOr(&bits[base + cw], (unsigned long long)cb);
      cw = w;
      cb = 0;
    }
    cb |= 1ull << walk_bit(u);
    if (++k >= m) break;  // m == 0: the endpoint of a zero-length ray
    bool y = dyx < 0;
    const int dzm = y ? dzy : dzx;
    bool z = dzm < 0;
    if (__any(dyx == 0 || dzm == 0)) {
      // exact rule for every lane: crossing counts from the walk state
      const uint32_t fa = (u >> 12) & 31u, fb = (u >> 6) & 31u, fc = u & 31u;
      const uint32_t lx = L == 0 ? fa : fb, ly = L == 1 ? fa : (L == 0 ? fb : fc), lz = L == 2 ? fa : fc;
      IntDda t;
      t.n0 = n0, t.n1 = n1, t.n2 = n2;
      t.c0 = c00 + (uint32_t)abs((int)lx - (int)lx0);
      t.c1 = c10 + (uint32_t)abs((int)ly - (int)ly0);
      t.c2 = c20 + (uint32_t)abs((int)lz - (int)lz0);
      t.dyx = dyx, t.dzx = dzx, t.dzy = dzy;
      const int ax = t.next_axis(tbl);
      z = ax == 2;
      y = ax == 1;
    }
    y = y && !z;
    const bool x = !z && !y;
    dyx += z ? 0 : (y ? tn0 : -tn1);
    dzx += y ? 0 : (z ? tn0 : -tn2);
    dzy += x ? 0 : (z ? tn1 : -tn2);
    u += (uint32_t)(z ? iz : (y ? iy : ix));
    esc |= u;
  }
  atomicOr(&bits[base + cw], (unsigned long long)cb);
  if (esc & ~kWalkFields) {
    atomicOr(&st->error, kErrBrick);  // the scan is rejected
  } else if (sr.w & 0x80000000u) {
    // setValueOn(end, true) (VDBMapping.hpp:533-536): the last visit is the end voxel
    const uint32_t fa = (u >> 12) & 31u, fb = (u >> 6) & 31u, fc = u & 31u;
    const uint32_t lx = L == 0 ? fa : fb, ly = L == 1 ? fa : (L == 0 ? fb : fc), lz = L == 2 ? fa : fc;
    const uint32_t hw = ((lx >> 3) << 7) | ((ly >> 3) << 5) | ((lz >> 3) << 3) | (lx & 7u);
    atomicOr(&bits[512 + hw], 1ull << (((ly & 7u) << 3) | (lz & 7u)));
  }
}

// Merge the layouts of the walk bitmap into leaf-major visited words at
// [0, 512) (word j of leaf l at l * 8 + j, bit (y & 7) * 8 + (z & 7)); `used`
// names the layouts any segment of the item marked. Called by the whole
// workgroup (two words per thread); ends with a barrier.
__device__ void merge_layouts(unsigned long long* bits, uint32_t used) {
  uint64_t out[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int q = (int)threadIdx.x + r * 256;
    const int l = q >> 3, j = q & 7;
    const int lx = l >> 4, ly = (l >> 2) & 3, lz = l & 3;
    uint64_t v = (used & 1u) ? bits[(lx * 8 + j) * 16 + ly * 4 + lz] : 0ull;
    if (used & 2u) {  // layout Y: byte j of the word at y = 8 ly + yy is byte yy here
      const unsigned long long* p = bits + 1024 + (ly * 8) * 16 + lx * 4 + lz;
#pragma unroll
      for (int yy = 0; yy < 8; ++yy) v |= ((p[yy * 16] >> (8 * j)) & 0xFFull) << (8 * yy);
    }
    if (used & 4u) {  // layout Z: bit yy of byte j of the word at z = 8 lz + zz is bit 8 yy + zz here
      const unsigned long long* p = bits + 1536 + (lz * 8) * 16 + lx * 4 + ly;
      uint64_t mz = 0;
#pragma unroll
      for (int zz = 0; zz < 8; ++zz) mz |= ((p[zz * 16] >> (8 * j)) & 0xFFull) << (8 * zz);
      v |= transpose8(mz);
    }
    out[r] = v;
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 2; ++r) bits[threadIdx.x + r * 256] = out[r];
  __syncthreads();
}

constexpr int kWalkThreads = 256;
constexpr int kSegsPerThread = kItemSegs / kWalkThreads;  // an item's segments, per thread
static_assert(kSegsPerThread <= 4, "segment lengths are packed one byte per segment");

__global__ __launch_bounds__(kWalkThreads, 8) void k_brick_walk(const BrickItem* __restrict__ items,
                                                             const uint4* __restrict__ segs,
                                                             const int4* __restrict__ rayinfo,
                                                             const ScanInfo* __restrict__ scans,
                                                             const double* __restrict__ tbl, BrickView bv,
                                                             unsigned long long* __restrict__ shared_bits,
                                                             UpdView rec, unsigned long long* __restrict__ leaf_mask,
                                                             uint32_t rec_cap, Status* st) {
  __shared__ unsigned long long bits[kWalkWords];
  __shared__ uint32_t s_pos[64];
  __shared__ uint32_t s_hist[128];
  __shared__ uint16_t s_order[kItemSegs];
  __shared__ unsigned int s_stop;
  __shared__ uint32_t s_used;
  if (stopped(st, &s_stop)) return;
  const uint32_t n_items = st->items;
  if ((uint64_t)n_items * 64 > rec_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&st->overflow, kOvfRecords);
    return;
  }
  const uint32_t t = threadIdx.x;
  // static assignment: item it -> workgroup it % grid (no shared queue counter)
  for (uint32_t it = blockIdx.x; it < n_items; it += gridDim.x) {
    const BrickItem item = items[it];
    int b[3];
    brick_coords(bv.keys[item.slot], b);
    const int ox = b[0] * kBrickDim, oy = b[1] * kBrickDim, oz = b[2] * kBrickDim;
    for (int w = t; w < kWalkWords; w += kWalkThreads) bits[w] = 0;
    if (t < 128) s_hist[t] = 0;
    if (t == 0) s_used = 0;
    // Counting sort of the item's segments by visit count, longest first, so
    // the lanes of a wave walk segments of about the same length. Segment k
    // belongs to thread k % 256; its length is loaded once, for both passes.
    const uint32_t nseg = item.seg1 - item.seg0;  // <= kItemSegs
    uint32_t lens = 0;  // 7-bit lengths of the thread's segments, one byte each
#pragma unroll
    for (int r = 0; r < kSegsPerThread; ++r) {
      const uint32_t k = t + r * kWalkThreads;
      if (k < nseg) lens |= ((segs[item.seg0 + k].w >> 16) & 0x7Fu) << (8 * r);
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegsPerThread; ++r)
      if (t + r * kWalkThreads < nseg) atomicAdd(&s_hist[127 - ((lens >> (8 * r)) & 0xFFu)], 1u);
    __syncthreads();
    if (t < 64) {  // exclusive scan of the 128 bins, two per lane
      const uint32_t h0 = s_hist[2 * t], h1 = s_hist[2 * t + 1];
      uint32_t v = h0 + h1;
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const uint32_t u = __shfl_up(v, o, 64);
        if ((int)t >= o) v += u;
      }
      const uint32_t ex = v - h0 - h1;
      s_hist[2 * t] = ex;
      s_hist[2 * t + 1] = ex + h0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kSegsPerThread; ++r) {
      const uint32_t k = t + r * kWalkThreads;
      if (k < nseg) s_order[atomicAdd(&s_hist[127 - ((lens >> (8 * r)) & 0xFFu)], 1u)] = (uint16_t)k;
    }
    __syncthreads();
    // Round r of the walk covers the sorted segments [256 r, 256 r + 256); wave
    // w takes the 64 of block (w + r) mod 4, so over the rounds every wave
    // walks long and short ones alike. The next round's record is loaded
    // before the current one is walked.
    uint32_t used = 0;
    const uint32_t lane = t & 63u, wave = t >> 6;
    uint32_t k = (((wave + 0) & 3u) << 6) + lane;
    uint4 sr = k < nseg ? segs[item.seg0 + s_order[k]] : make_uint4(0, 0, 0, 0);
#pragma unroll 1
    for (uint32_t r = 0; k < nseg;) {
      ++r;
      const uint32_t kn = r * kWalkThreads + (((wave + r) & 3u) << 6) + lane;
      const uint4 sn = kn < nseg ? segs[item.seg0 + s_order[kn]] : make_uint4(0, 0, 0, 0);
      walk_segment(sr, tbl, bits, used, st);
      k = kn;
      sr = sn;
    }
    if (used) atomicOr(&s_used, used);
    __syncthreads();
    merge_layouts(bits, s_used);
    if (item.shared == kNone) {
      emit_leaves(bits, ox, oy, oz, rec, it * 64, s_pos, &leaf_mask[it]);
    } else {
      unsigned long long* dst = shared_bits + (size_t)item.shared * kBrickWords;
      for (int w = t; w < kBrickWords; w += kWalkThreads) {
        const uint64_t v = bits[w];
        if (v) atomicOr(&dst[w], (unsigned long long)v);
      }
      if (t == 0) leaf_mask[it] = 0;  // its records come from k_split_emit
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_split_emit(const uint32_t* __restrict__ split_slot,
                                                    unsigned long long* __restrict__ shared_bits,
                                                    BrickView bv, UpdView rec,
                                                    unsigned long long* __restrict__ leaf_mask,
                                                    uint32_t rec_cap, Status* st) {
  __shared__ unsigned long long bits[kBrickWords];
  __shared__ uint32_t s_pos[64];
  __shared__ unsigned int s_stop;
  if (stopped(st, &s_stop)) return;
  const uint32_t ns = st->splits, n_items = st->items;
  if ((uint64_t)(n_items + ns) * 64 > rec_cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&st->overflow, kOvfRecords);
    return;
  }
  for (uint32_t k = blockIdx.x; k < ns; k += gridDim.x) {
    unsigned long long* src = shared_bits + (size_t)k * kBrickWords;
    for (int w = threadIdx.x; w < kBrickWords; w += blockDim.x) {
      bits[w] = src[w];
      src[w] = 0;
    }
    __syncthreads();
    int b[3];
    brick_coords(bv.keys[split_slot[k]], b);
    emit_leaves(bits, b[0] * kBrickDim, b[1] * kBrickDim, b[2] * kBrickDim, rec, (n_items + k) * 64, s_pos,
                &leaf_mask[n_items + k]);
    __syncthreads();
  }
}

// Record list from the per-group (64 positions) leaf masks: per 256 groups
// the number of filled positions, then after the scan of those sums every
// group writes its positions in order. No shared counters.
__global__ __launch_bounds__(256) void k_mask_sums(const unsigned long long* __restrict__ leaf_mask,
                                                   uint64_t* __restrict__ bsum, const Status* st) {
  __shared__ uint64_t s_w[4];
  __shared__ unsigned int s_stop;
  if (stopped(st, &s_stop)) return;
  const uint32_t ng = st->items + st->splits;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t c = g < ng ? (uint64_t)__popcll(leaf_mask[g]) : 0;
  const uint64_t t = block_sum_u64(c, s_w);
  if (threadIdx.x == 0) bsum[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_mask_list(const unsigned long long* __restrict__ leaf_mask,
                                                   const uint64_t* __restrict__ boff,
                                                   uint32_t* __restrict__ list, const Status* st) {
  __shared__ uint64_t s_w[4];
  __shared__ unsigned int s_stop;
  if (stopped(st, &s_stop)) return;
  const uint32_t ng = st->items + st->splits;
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint64_t m = g < ng ? leaf_mask[g] : 0;
  const uint64_t x = (uint64_t)__popcll(m);
  uint64_t v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint64_t t = __shfl_up(v, o, 64);
    if (bw_lane() >= o) v += t;
  }
  const int wv = threadIdx.x >> 6;
  if (bw_lane() == 63) s_w[wv] = v;
  __syncthreads();
  uint64_t add = 0;
  for (int w = 0; w < wv; ++w) add += s_w[w];
  uint32_t pos = (uint32_t)(boff[blockIdx.x] + add + v - x);
  for (uint64_t r = m; r; r &= r - 1) list[pos++] = g * 64 + (uint32_t)__builtin_ctzll(r);
}

// Dense copy of the listed records (key and 16 mask words each, sixteen lanes
// per record) into the record set the map stream reads; the positional set is
// then free for the next scan's walk. Copies nothing when the count exceeds
// dcap (the host grows the set and runs it again).
__global__ __launch_bounds__(256) void k_rec_gather(const uint32_t* __restrict__ list, UpdView sparse,
                                                    const RecDest* __restrict__ dest, const Status* st) {
  if (st->error | st->overflow) return;
  const uint32_t n = st->records;
  const RecDest d = *dest;
  if (n > d.cap) return;
  const UpdView dense{d.keys, d.masks, 0};
  // eight lanes per record, 16 bytes each (the 128 mask bytes); each lane
  // group keeps kGatherRecs records in flight (list, then record loads, then
  // stores, issued together)
  constexpr int kGatherRecs = 4;
  const int l = (int)(threadIdx.x & 7);
  const uint64_t G = ((uint64_t)gridDim.x * blockDim.x) >> 3;
  const uint4* __restrict__ src = reinterpret_cast<const uint4*>(sparse.masks);
  uint4* __restrict__ dst = reinterpret_cast<uint4*>(dense.masks);
  for (uint64_t i0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 3; i0 < n; i0 += kGatherRecs * G) {
    uint32_t pos[kGatherRecs];
#pragma unroll
    for (int r = 0; r < kGatherRecs; ++r) pos[r] = i0 + r * G < n ? list[i0 + r * G] : kNone;
    uint4 v[kGatherRecs];
    unsigned long long key[kGatherRecs];
#pragma unroll
    for (int r = 0; r < kGatherRecs; ++r) {
      if (pos[r] == kNone) continue;
      v[r] = src[(size_t)pos[r] * 8 + l];
      if (l == 0) key[r] = sparse.keys[pos[r]];
    }
#pragma unroll
    for (int r = 0; r < kGatherRecs; ++r) {
      if (pos[r] == kNone) continue;
      const uint64_t i = i0 + r * G;
      dst[i * 8 + l] = v[r];
      if (l == 0) dense.keys[i] = key[r];
    }
  }
}

// The record set of the next cast (one lane).
__global__ void k_put_dest(RecDest* dst, RecDest v) {
  if (threadIdx.x == 0) *dst = v;
}

// The map's leaf count into host-visible memory (one lane; no copy engine).
__global__ void k_snap_count(const Status* st, uint32_t* host_dst) {
  if (threadIdx.x == 0) __hip_atomic_store(host_dst, st->leaf_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ __launch_bounds__(256) void k_iota_u32(uint32_t* p, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = i;
}

// ---------------------------------------------------------------- launchers --

hipError_t launch_rec_gather(hipStream_t s, int grid, const uint32_t* list, UpdView sparse, const RecDest* dest,
                             const Status* st) {
  hipLaunchKernelGGL(k_rec_gather, dim3(grid), dim3(256), 0, s, list, sparse, dest, st);
  return hipGetLastError();
}

hipError_t launch_put_dest(hipStream_t s, RecDest* dst, RecDest v) {
  hipLaunchKernelGGL(k_put_dest, dim3(1), dim3(64), 0, s, dst, v);
  return hipGetLastError();
}

hipError_t launch_snap_count(hipStream_t s, const Status* st, uint32_t* host_dst) {
  hipLaunchKernelGGL(k_snap_count, dim3(1), dim3(64), 0, s, st, host_dst);
  return hipGetLastError();
}

hipError_t launch_iota_u32(hipStream_t s, uint32_t* p, uint32_t n) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(k_iota_u32, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

hipError_t launch_seg_count(hipStream_t s, int grid, const int4* rayinfo, const ScanInfo* scans,
                            const BatchCtl* ctl, const double* tbl, BrickView bv, unsigned long long* wkey,
                            unsigned long long* wsc, Status* st) {
  hipLaunchKernelGGL(k_seg_count, dim3(grid), dim3(256), 0, s, rayinfo, scans, ctl, tbl, bv, wkey, wsc, st);
  return hipGetLastError();
}

hipError_t launch_brick_offsets(hipStream_t s, BrickView bv, uint64_t* bsum, uint64_t* boff,
                                BrickItem* items, uint32_t seg_cap, uint32_t item_cap,
                                uint32_t* split_slot, uint32_t split_cap, Status* st) {
  const uint32_t nb = (bv.cap_mask + 1) / 256;
  hipLaunchKernelGGL(k_brick_sums, dim3(nb), dim3(256), 0, s, bv, bsum, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_u64, dim3(1), dim3(256), 0, s, bsum, boff, nb, &st->items, 32);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_brick_offsets, dim3(nb), dim3(256), 0, s, bv, boff, items, seg_cap, item_cap,
                     split_slot, split_cap, st);
  return hipGetLastError();
}

hipError_t launch_seg_fill(hipStream_t s, int grid, const int4* rayinfo, const ScanInfo* scans,
                           const BatchCtl* ctl, const double* tbl, BrickView bv,
                           const unsigned long long* wkey, const unsigned long long* wsc, uint4* segs,
                           uint32_t seg_cap, Status* st) {
  hipLaunchKernelGGL(k_seg_fill, dim3(grid), dim3(256), 0, s, rayinfo, scans, ctl, tbl, bv, wkey, wsc, segs,
                     seg_cap, st);
  return hipGetLastError();
}

hipError_t launch_brick_walk(hipStream_t s, int grid, const BrickItem* items, const uint4* segs,
                             const int4* rayinfo, const ScanInfo* scans, const double* tbl,
                             BrickView bv, unsigned long long* shared_bits, UpdView rec,
                             unsigned long long* leaf_mask, uint32_t rec_cap, Status* st) {
  hipLaunchKernelGGL(k_brick_walk, dim3(grid), dim3(kWalkThreads), 0, s, items, segs, rayinfo, scans, tbl, bv,
                     shared_bits, rec, leaf_mask, rec_cap, st);
  return hipGetLastError();
}

// nb_cap: 256-group blocks the bsum/boff arrays hold (>= (items + splits) / 256).
hipError_t launch_rec_list(hipStream_t s, const unsigned long long* leaf_mask, uint64_t* bsum, uint64_t* boff,
                           uint32_t nb_cap, uint32_t* list, Status* st) {
  hipLaunchKernelGGL(k_mask_sums, dim3(nb_cap), dim3(256), 0, s, leaf_mask, bsum, st);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_u64, dim3(1), dim3(256), 0, s, bsum, boff, nb_cap, &st->records, 0);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mask_list, dim3(nb_cap), dim3(256), 0, s, leaf_mask, boff, list, st);
  return hipGetLastError();
}

hipError_t launch_split_emit(hipStream_t s, int grid, const uint32_t* split_slot,
                             unsigned long long* shared_bits, BrickView bv, UpdView rec,
                             unsigned long long* leaf_mask, uint32_t rec_cap, Status* st) {
  hipLaunchKernelGGL(k_split_emit, dim3(grid), dim3(256), 0, s, split_slot, shared_bits, bv, rec,
                     leaf_mask, rec_cap, st);
  return hipGetLastError();
}

}  // namespace vdbgpu
