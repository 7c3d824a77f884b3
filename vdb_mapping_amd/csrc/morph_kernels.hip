// gfx950 kernels of the morphology operators on a leaf set
// (VDBMapping.hpp:1095-1148: morphologicalDilateMap / ErodeMap over
// tools::dilateActiveValues / erodeActiveValues with 26-neighbour
// connectivity, composed by morphologicalCloseMap / OpenMap for the smoothed
// map-section applies, VDBMapping.hpp:1022-1085).
//
// The leaf set lives in a table: per leaf index its origin, its active mask
// (two buffers, read one, write the other), its 27 neighbour indices (the 3^3
// block of leaves around it, itself at 13; kNone where no leaf is) and a keep
// flag; a hash maps leaf keys to indices.
//
//   k_morph_seed   hash every leaf of the table
//   k_morph_grow   before up to 8 dilation steps: add the 26 neighbours of
//                  every leaf with an active voxel (8 steps reach no further)
//   k_morph_nbrs   the 27 neighbour indices of every leaf
//   k_morph_step   one 26-neighbour step, one lane per 64-voxel word
//   k_morph_keep   flag the leaves that hold an active voxel after dilation
//
// A step works on whole words: word x of a leaf is the 8x8 (y, z) plane at
// local x, bit y * 8 + z. Dilation ORs, over the planes x - 1, x, x + 1, the
// plane grown by one voxel in y and z (rows and columns of the neighbouring
// leaves' planes shifted in). Erosion is the complement of the dilation of the
// complement, where a voxel in no leaf is inactive.
#include "device_common.h"
#include "kernels.h"

namespace vdbgpu {

namespace {

constexpr uint64_t kColZ0 = 0x0101010101010101ull;  // z = 0 bit of every row
constexpr uint64_t kColZ7 = kColZ0 << 7;            // z = 7 bit of every row

__device__ __forceinline__ uint32_t morph_find(const MorphView& mv, uint64_t key) {
  uint32_t h = (uint32_t)fmix64(key) & mv.hmask;
  // the table is at most half full, so an empty slot ends every probe
  for (uint32_t p = 0; p <= mv.hmask; ++p) {
    const unsigned long long k = mv.keys[h];
    if (k == key) return mv.hidx[h];
    if (k == kEmpty) return kNone;
    h = (h + 1) & mv.hmask;
  }
  return kNone;
}

// Insert key; returns the slot when this call inserted it, kNone when the key
// was there already (or, never at the table's fill, no slot was found).
__device__ __forceinline__ uint32_t morph_insert(const MorphView& mv, uint64_t key) {
  uint32_t h = (uint32_t)fmix64(key) & mv.hmask;
  for (uint32_t p = 0; p <= mv.hmask; ++p) {
    const unsigned long long prev = atomicCAS(&mv.keys[h], kEmpty, (unsigned long long)key);
    if (prev == kEmpty) return h;
    if (prev == key) return kNone;
    h = (h + 1) & mv.hmask;
  }
  return kNone;
}

__device__ __forceinline__ bool any_on(const unsigned long long* m) {
  uint64_t a = 0;
#pragma unroll
  for (int w = 0; w < 8; ++w) a |= m[w];
  return a != 0;
}

}  // namespace

__global__ __launch_bounds__(256) void k_morph_seed(MorphView mv, uint32_t count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= count) return;
  const int* o = mv.org + 3 * (size_t)i;
  const uint32_t h = morph_insert(mv, leaf_key(o[0], o[1], o[2]));
  if (h != kNone) mv.hidx[h] = i;  // keys of the table are distinct
}

// One lane per (leaf, neighbour direction). New leaves are appended at
// count + atomicAdd(added) with an empty mask; cap bounds the table.
__global__ __launch_bounds__(256) void k_morph_grow(MorphView mv, unsigned long long* masks, uint8_t* keep,
                                                    uint32_t count, uint32_t cap, unsigned int* added) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint64_t)count * 27) return;
  const uint32_t i = (uint32_t)(t / 27);
  const int k = (int)(t % 27);
  if (k == 13 || !any_on(masks + (size_t)i * 8)) return;
  const int* o = mv.org + 3 * (size_t)i;
  const int n[3] = {o[0] + (k / 9 - 1) * 8, o[1] + ((k / 3) % 3 - 1) * 8, o[2] + (k % 3 - 1) * 8};
  const uint32_t h = morph_insert(mv, leaf_key(n[0], n[1], n[2]));
  if (h == kNone) return;
  const uint32_t pos = count + atomicAdd(added, 1u);
  if (pos >= cap) return;  // the host sizes cap to 27 * count: never taken
  mv.hidx[h] = pos;
  int* d = mv.org + 3 * (size_t)pos;
  d[0] = n[0];
  d[1] = n[1];
  d[2] = n[2];
#pragma unroll
  for (int w = 0; w < 8; ++w) masks[(size_t)pos * 8 + w] = 0;
  keep[pos] = 0;
}

__global__ __launch_bounds__(256) void k_morph_nbrs(MorphView mv, uint32_t count) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint64_t)count * 27) return;
  const uint32_t i = (uint32_t)(t / 27);
  const int k = (int)(t % 27);
  const int* o = mv.org + 3 * (size_t)i;
  mv.nbr[t] = k == 13 ? i
                      : morph_find(mv, leaf_key(o[0] + (k / 9 - 1) * 8, o[1] + ((k / 3) % 3 - 1) * 8,
                                                o[2] + (k % 3 - 1) * 8));
}

__global__ __launch_bounds__(256) void k_morph_step(const unsigned long long* __restrict__ in,
                                                    unsigned long long* __restrict__ out,
                                                    const uint32_t* __restrict__ nbr, uint32_t count, int erode) {
  const uint64_t t = (uint64_t)blockIdx.x * 256u + threadIdx.x;
  if (t >= (uint64_t)count * 8) return;
  const uint32_t i = (uint32_t)(t >> 3);
  const int x = (int)(t & 7);
  const uint32_t* nb = nbr + (size_t)i * 27;
  // erosion dilates the complement: flip every word, and a missing leaf
  // (all voxels inactive) reads as all ones
  const uint64_t flip = erode ? ~0ull : 0ull;
  uint64_t d = 0;
#pragma unroll
  for (int dx = -1; dx <= 1; ++dx) {
    const int xs = x + dx;
    const int lx = xs < 0 ? 0 : (xs > 7 ? 2 : 1);
    const int w = xs & 7;
    uint64_t yz[3];  // the plane at xs of the leaf column lz, grown in y
#pragma unroll
    for (int lz = 0; lz < 3; ++lz) {
      uint64_t p[3];
#pragma unroll
      for (int ly = 0; ly < 3; ++ly) {
        const uint32_t j = nb[lx * 9 + ly * 3 + lz];
        p[ly] = j == kNone ? flip : (in[(size_t)j * 8 + w] ^ flip);
      }
      // row y takes rows y - 1 and y + 1; row 0 takes row 7 of the y - 1 leaf,
      // row 7 row 0 of the y + 1 leaf
      yz[lz] = p[1] | (p[1] << 8) | (p[0] >> 56) | (p[1] >> 8) | (p[2] << 56);
    }
    // column z takes columns z - 1 and z + 1; column 0 takes column 7 of the
    // z - 1 leaf, column 7 column 0 of the z + 1 leaf
    d |= yz[1] | ((yz[1] << 1) & ~kColZ0) | ((yz[0] >> 7) & kColZ0) | ((yz[1] >> 1) & ~kColZ7) |
         ((yz[2] << 7) & kColZ7);
  }
  out[t] = erode ? (in[t] & ~d) : d;
}

__global__ __launch_bounds__(256) void k_morph_keep(const unsigned long long* __restrict__ masks, uint8_t* keep,
                                                    uint32_t count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= count) return;
  if (any_on(masks + (size_t)i * 8)) keep[i] = 1;
}

namespace {
inline unsigned int blocks(uint64_t threads) { return (unsigned int)((threads + 255) / 256); }
}  // namespace

hipError_t launch_morph_seed(hipStream_t s, MorphView mv, uint32_t count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_morph_seed, dim3(blocks(count)), dim3(256), 0, s, mv, count);
  return hipGetLastError();
}

hipError_t launch_morph_grow(hipStream_t s, MorphView mv, unsigned long long* masks, uint8_t* keep, uint32_t count,
                             uint32_t cap, unsigned int* added) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_morph_grow, dim3(blocks((uint64_t)count * 27)), dim3(256), 0, s, mv, masks, keep, count, cap,
                     added);
  return hipGetLastError();
}

hipError_t launch_morph_nbrs(hipStream_t s, MorphView mv, uint32_t count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_morph_nbrs, dim3(blocks((uint64_t)count * 27)), dim3(256), 0, s, mv, count);
  return hipGetLastError();
}

hipError_t launch_morph_step(hipStream_t s, const unsigned long long* in, unsigned long long* out,
                             const uint32_t* nbr, uint32_t count, bool erode) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_morph_step, dim3(blocks((uint64_t)count * 8)), dim3(256), 0, s, in, out, nbr, count,
                     erode ? 1 : 0);
  return hipGetLastError();
}

hipError_t launch_morph_keep(hipStream_t s, const unsigned long long* masks, uint8_t* keep, uint32_t count) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(k_morph_keep, dim3(blocks(count)), dim3(256), 0, s, masks, keep, count);
  return hipGetLastError();
}

}  // namespace vdbgpu
