// Shared device/host definitions of the gfx950 occupancy-integration kernels.
//
// Data layout in HBM (one map handle):
//   map leaf hash   : keys u64[mcap] (open addressing, linear probing, EMPTY = ~0),
//                     idx u32[mcap] (leaf pool index of the key)
//   leaf pool       : values f32[lcap][512] (OpenVDB LeafNode<float,3> buffer
//                     order, offset ((x&7)<<6)|((y&7)<<3)|(z&7)),
//                     active u64[lcap][8] (value mask), leaf_keys u64[lcap], dirty u8[lcap]
//   update set (per source): keys u64[ucap], masks u64[ucap][16] (visited words 0..7,
//                     hit words 8..15, same bit layout as the value mask)
//   per scan        : rayinfo int4[n] (end voxel xyz, visit count | hit flag),
//                     pre u32[n] (inclusive visit-count prefix inside each 256-ray block),
//                     bsum u32[nb], boff u64[nb] (exclusive block offsets)
//   crossing table  : tbl f64, row n (n = 1..tbl_n) at n(n-1)/2 holding the
//                     DDA crossing times of an axis with |delta| = n
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vdbgpu {

constexpr uint64_t kEmpty = ~0ull;
constexpr uint32_t kNone = 0xFFFFFFFFu;
constexpr int kCoordLimit = 1 << 23;  // voxel coords in [-2^23, 2^23): 21-bit leaf coords
constexpr uint32_t kHitBit = 0x80000000u;   // rayinfo.w: endpoint is a hit
constexpr uint32_t kValidBit = 0x40000000u; // rayinfo.w: the ray is cast (not NaN, range > 0)
constexpr uint32_t kCntMask = 0xFFFFu;      // rayinfo.w bits 0-15: visit count (<= 1 + 3 * 8192)
constexpr int kScanShift = 16;              // rayinfo.w bits 16-22: scan index in a batch
constexpr uint32_t kScanMask = 0x7Fu;
constexpr int kMaxBatchScans = 128;
constexpr int kBatchCoordLimit = 1 << 21;   // batched keys: 19-bit leaf coords + 7-bit scan
constexpr int kPrepBlock = 256;
constexpr int kMaxQuerySteps = 1 << 24;  // per-axis voxel extent bound of a raytrace query

// Status bits written by kernels.
constexpr uint32_t kErrRange = 1u;  // a voxel index out of range / non-finite
constexpr uint32_t kErrTable = 2u;  // crossing table or chunk table too short
constexpr uint32_t kErrEntry = 4u;  // mid-ray DDA entry did not converge
constexpr uint32_t kErrBrick = 8u;  // brick path inconsistency (a segment left its brick)
constexpr uint32_t kErrWalk = 16u;  // fast-mode walk exceeded its step bound
constexpr uint32_t kOvfHard = 1u;   // update-set probe found no free slot
constexpr uint32_t kMapFull = 1u;   // Status::map_full: the fold needs a larger leaf pool / hash
// Brick path overflow bits (Status::overflow): a table too small for this
// scan; the host grows it and re-runs the brick stages.
constexpr uint32_t kOvfBricks = 2u;   // brick hash
constexpr uint32_t kOvfSegs = 4u;     // segment array
constexpr uint32_t kOvfItems = 8u;    // work items
constexpr uint32_t kOvfSplit = 16u;   // shared-brick bitmaps
constexpr uint32_t kOvfRecords = 32u; // update-leaf records
constexpr uint32_t kItemSegs = 1024;  // segments per brick work item

// Direct edit ops on the marked voxels of an edit set (k_set_voxels).
constexpr int kEditOccupied = 0;  // setNodeToOccupied: (max log-odds, on)
constexpr int kEditFree = 1;      // setNodeToFree: (min log-odds, off)
constexpr int kEditState = 2;     // setNodeState: on = value > t_max
constexpr int kEditActivate = 3;  // setActiveState(on)

struct Status {
  unsigned long long total;     // visits of the scan (sum of counts)
  unsigned long long voxels;    // visited voxels in the update set (compact)
  unsigned long long rays;      // valid rays of the scan
  unsigned int error;
  unsigned int overflow;
  unsigned int entries;         // update-set leaves (compact)
  unsigned int leaf_count;      // map leaves (map insert)
  unsigned int groups;          // leaves of a batch (fold groups)
  unsigned int map_full;
  unsigned int segs;            // brick segments of the scan (brick path)
  unsigned int items;           // brick work items
  unsigned int splits;          // bricks shared by several work items
  unsigned int records;         // update-leaf records listed by the brick path
  unsigned int reserved;
  unsigned int pad;
};

// One scan of a batch: sensor origin, max range, its rays [ray0, ray0 + nrays)
// of the concatenated cloud; s = rounded origin voxel (written by k_prep).
struct ScanInfo {
  double o[3];
  double R;
  uint32_t ray0, nrays;
  int s[3];
  int pad;
};

// Sizes read by the kernels (device memory, so a captured graph can be replayed).
struct BatchCtl {
  uint32_t n_rays;
  uint32_t nb;       // 256-ray blocks
  uint32_t n_scans;
  uint32_t batched;  // 1: update-set keys are (leaf, scan)
};

struct UpdView {
  unsigned long long* keys;
  unsigned long long* masks;  // 16 words per slot
  uint32_t cap_mask;
};

// Where the brick path copies a scan's records (one record set of the ring):
// kept in device memory, so a replayed cast graph copies into the set of the
// scan it runs for.
struct RecDest {
  unsigned long long* keys;
  unsigned long long* masks;  // 16 words per record
  uint32_t cap;
  uint32_t pad;
};

// Batch fold groups: leaf (57-bit batched key) -> 128-bit scan mask + create flag.
struct GroupView {
  unsigned long long* keys;
  unsigned long long* scans;  // 2 words per slot
  unsigned int* need;
  uint32_t cap_mask;
};

// Bricks (32^3 voxels) of one scan: key -> segment count and fill cursor.
struct BrickView {
  unsigned long long* keys;  // kEmpty = free
  uint32_t* cnt;             // segments in the brick
  uint32_t* cur;             // next free segment position (fill stage)
  uint32_t cap_mask;
};
// One work item of the brick walk: segments [seg0, seg1) of brick `slot`;
// `shared` indexes the global bitmap the items of a split brick OR into
// (kNone: the item owns the brick).
struct BrickItem {
  uint32_t slot, seg0, seg1, shared;
};
struct MapView {
  unsigned long long* keys;
  uint32_t* idx;
  uint32_t cap_mask;
  float* values;
  unsigned long long* active;
  unsigned long long* leaf_keys;
  uint8_t* dirty;
  uint32_t lcap;  // leaf pool capacity
};

// Leaf-level tiles of pruneGrid (createMapFromPointCloud): leaf i of the pool
// stands for a tile of value val[i] when bit i of `bits` is set, active when
// bit i of `on` is set. bits == nullptr: the map holds no tile.
struct TileView {
  uint32_t* bits;
  uint32_t* on;
  float* val;
  uint32_t cap;  // leaves the arrays cover; a later leaf is never a tile
};

__host__ __device__ inline bool tile_at(const TileView& t, uint32_t i) {
  return t.bits && i < t.cap && ((t.bits[i >> 5] >> (i & 31)) & 1u);
}
__host__ __device__ inline bool tile_active(const TileView& t, uint32_t i) {
  return (t.on[i >> 5] >> (i & 31)) & 1u;
}

// Topology snapshot used by raytrace: hash sets of leaf, 128^3-node and
// 4096^3-node keys, plus the padded bbox of leaves with an active voxel.
struct RtView {
  unsigned long long* leaf;
  unsigned long long* n128;
  unsigned long long* n4096;
  uint32_t leaf_mask, n128_mask, n4096_mask;
  int bmin[3], bmax[3];
};

__host__ __device__ inline uint64_t leaf_key(int x, int y, int z) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)(uint32_t)(x >> 3) & m) << 42 | ((uint64_t)(uint32_t)(y >> 3) & m) << 21 |
         ((uint64_t)(uint32_t)(z >> 3) & m);
}

__host__ __device__ inline int sext21(uint64_t v) { return (int)((int64_t)(v << 43) >> 43); }

__host__ __device__ inline void key_origin(uint64_t k, int o[3]) {
  const uint64_t m = (1ull << 21) - 1;
  o[0] = sext21((k >> 42) & m) * 8;
  o[1] = sext21((k >> 21) & m) * 8;
  o[2] = sext21(k & m) * 8;
}

// MurmurHash3 64-bit finaliser.
__host__ __device__ inline uint64_t fmix64(uint64_t k) {
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

__device__ inline bool coord_ok(int c) { return c >= -kCoordLimit && c < kCoordLimit; }

// Update-set key of the leaf holding (x,y,z): the 63-bit leaf key, or in a
// batch the 57-bit leaf key (19-bit coords) shifted left 7 with the scan index.
__host__ __device__ inline uint64_t update_key(int x, int y, int z, uint32_t scan, bool batched) {
  if (!batched) return leaf_key(x, y, z);
  const uint64_t m = (1ull << 19) - 1;
  return (((uint64_t)(uint32_t)(x >> 3) & m) << 45) | (((uint64_t)(uint32_t)(y >> 3) & m) << 26) |
         (((uint64_t)(uint32_t)(z >> 3) & m) << 7) | scan;
}

// Canonical 63-bit leaf key of a batched update key (scan bits dropped).
__host__ __device__ inline uint64_t batched_to_leaf_key(uint64_t k) {
  const uint64_t m = (1ull << 19) - 1;
  const int lx = (int)((int64_t)(((k >> 45) & m) << 45) >> 45);
  const int ly = (int)((int64_t)(((k >> 26) & m) << 45) >> 45);
  const int lz = (int)((int64_t)(((k >> 7) & m) << 45) >> 45);
  return leaf_key(lx * 8, ly * 8, lz * 8);
}

// Key of the brick (32^3 voxels) with brick coordinates (bx,by,bz) = voxel >> 5
// (19-bit fields: voxel coords in [-2^23, 2^23)).
__host__ __device__ inline uint64_t brick_key(int bx, int by, int bz) {
  const uint64_t m = (1ull << 19) - 1;
  return ((uint64_t)(uint32_t)bx & m) << 38 | ((uint64_t)(uint32_t)by & m) << 19 |
         ((uint64_t)(uint32_t)bz & m);
}
__host__ __device__ inline void brick_coords(uint64_t k, int b[3]) {
  const uint64_t m = (1ull << 19) - 1;
  b[0] = (int)((int64_t)(((k >> 38) & m) << 45) >> 45);
  b[1] = (int)((int64_t)(((k >> 19) & m) << 45) >> 45);
  b[2] = (int)((int64_t)((k & m) << 45) >> 45);
}
// Key of the internal node (2^log2dim voxels per side) holding voxel (x,y,z).
__host__ __device__ inline uint64_t node_key(int x, int y, int z, int log2dim) {
  return leaf_key((x >> log2dim) * 8, (y >> log2dim) * 8, (z >> log2dim) * 8);
}

__device__ __forceinline__ bool set_insert(unsigned long long* keys, uint32_t mask, uint64_t key) {
  uint32_t h = (uint32_t)fmix64(key) & mask;
  for (uint32_t p = 0; p <= mask; ++p) {
    const unsigned long long prev = atomicCAS(&keys[h], kEmpty, (unsigned long long)key);
    if (prev == kEmpty) return true;
    if (prev == key) return false;
    h = (h + 1) & mask;
  }
  return false;
}

__device__ __forceinline__ bool set_has(const unsigned long long* __restrict__ keys, uint32_t mask,
                                        uint64_t key) {
  uint32_t h = (uint32_t)fmix64(key) & mask;
  for (uint32_t p = 0; p <= mask; ++p) {
    const unsigned long long k = keys[h];
    if (k == key) return true;
    if (k == kEmpty) return false;
    h = (h + 1) & mask;
  }
  return false;
}

// 8x8 bit-matrix transpose: bit 8r + c of x goes to bit 8c + r.
__device__ __forceinline__ uint64_t transpose8(uint64_t x) {
  uint64_t t = (x ^ (x >> 7)) & 0x00AA00AA00AA00AAull;
  x ^= t ^ (t << 7);
  t = (x ^ (x >> 14)) & 0x0000CCCC0000CCCCull;
  x ^= t ^ (t << 14);
  t = (x ^ (x >> 28)) & 0x00000000F0F0F0F0ull;
  x ^= t ^ (t << 28);
  return x;
}

}  // namespace vdbgpu
