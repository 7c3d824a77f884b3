// gfx950 kernels of pruneGrid (OccupancyVDBMapping.hpp:173, tools::prune with
// tolerance 0) at the leaf level, and of the rule that turns a tile back into a
// leaf.
//
//   k_prune       one wave per pool leaf: a leaf whose 512 values compare equal
//                 and whose voxels are all on or all off becomes a tile of that
//                 value and state (bit in tv.bits / tv.on, value in tv.val) and
//                 is appended to the tile list
//   k_tile_check  one wave per tile-list entry: a tile whose leaf is flagged
//                 dirty and whose voxels are no longer all (value, state) of the
//                 tile is a leaf again (OpenVDB's set / modify calls on a tile
//                 create a child exactly when they change the voxel)
//
// A tile keeps its voxels in the leaf pool (every voxel holds the tile value and
// state), so every voxel read and write works unchanged; only what iterates
// leaves or active values (exports, sections, the raytrace snapshot) reads the
// bits. Lane l of a wave reads values l, 64 + l, ..., 448 + l: eight coalesced
// 256-byte rows per leaf.
#include "device_common.h"
#include "kernels.h"

namespace vdbgpu {

namespace {

// 1 in every lane when the 512 values of leaf L all compare equal to v and its
// eight active words all equal `word`.
__device__ __forceinline__ bool leaf_uniform(const MapView& mp, uint32_t L, float v, uint64_t word) {
  const int lane = (int)__lane_id();
  const float* pv = mp.values + (size_t)L * 512;
  bool same = true;
#pragma unroll
  for (int k = 0; k < 8; ++k) same = same && pv[k * 64 + lane] == v;
  if (lane < 8) same = same && mp.active[(size_t)L * 8 + lane] == word;
  return __all(same);
}

}  // namespace

__global__ __launch_bounds__(256) void k_prune(MapView mp, uint32_t leaf_count, TileView tv,
                                               uint32_t* __restrict__ list, unsigned int* counter) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t L = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); L < leaf_count; L += waves) {
    const uint64_t w0 = mp.active[(size_t)L * 8];
    if (w0 != 0 && w0 != ~0ull) continue;
    const float v0 = mp.values[(size_t)L * 512];
    if (!leaf_uniform(mp, L, v0, w0)) continue;
    if (__lane_id() == 0) {
      atomicOr(&tv.bits[L >> 5], 1u << (L & 31));
      if (w0) atomicOr(&tv.on[L >> 5], 1u << (L & 31));
      tv.val[L] = v0;
      list[atomicAdd(counter, 1u)] = L;
    }
  }
}

__global__ __launch_bounds__(256) void k_tile_check(MapView mp, const uint32_t* __restrict__ list, uint32_t n,
                                                    TileView tv) {
  const uint32_t waves = gridDim.x * (blockDim.x >> 6);
  for (uint32_t i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < n; i += waves) {
    const uint32_t L = list[i];
    // a leaf no voxel write reached since the last check is unchanged
    if (!tile_at(tv, L) || !mp.dirty[L]) continue;
    const uint64_t word = tile_active(tv, L) ? ~0ull : 0ull;
    if (leaf_uniform(mp, L, tv.val[L], word)) continue;
    if (__lane_id() == 0) atomicAnd(&tv.bits[L >> 5], ~(1u << (L & 31)));
  }
}

// ------------------------------------------------------------- launchers --

hipError_t launch_prune(hipStream_t s, MapView mp, uint32_t leaf_count, TileView tv, uint32_t* list,
                        unsigned int* counter) {
  if (leaf_count == 0) return hipSuccess;
  const int grid = (int)std::min<uint32_t>(8192, (leaf_count + 3) / 4);
  hipLaunchKernelGGL(k_prune, dim3(grid), dim3(256), 0, s, mp, leaf_count, tv, list, counter);
  return hipGetLastError();
}

hipError_t launch_tile_check(hipStream_t s, MapView mp, const uint32_t* list, uint32_t n, TileView tv) {
  if (n == 0 || !tv.bits) return hipSuccess;
  const int grid = (int)std::min<uint32_t>(8192, (n + 3) / 4);
  hipLaunchKernelGGL(k_tile_check, dim3(grid), dim3(256), 0, s, mp, list, n, tv);
  return hipGetLastError();
}

}  // namespace vdbgpu
