// gfx950 kernels of the batch raytrace query (VDBMapping.hpp:675-721).
//
// The reference marches each ray with openvdb::tools::VolumeRayIntersector
// built by updateVolumeRayIntersectors (VDBMapping.hpp:1436-1449): a topology
// copy of the map at that moment, walked by a hierarchical DDA over node sizes
// 4096 -> 128 -> 8 (math::VolumeHDDA), then a voxel DDA that stops at the first
// active voxel of the live map. Here the topology copy is a snapshot of three
// device hash sets (leaves, 128^3 nodes, 4096^3 nodes) plus the padded
// bounding box of the leaves holding an active voxel; one lane walks one ray
// with exactly the double-precision operations of math::Ray / math::DDA.
//
// Build with -ffp-contract=off (every double operation rounds on its own).
#include <float.h>

#include "device_common.h"
#include "kernels.h"
#include "rt_common.h"

namespace vdbgpu {

__device__ __forceinline__ int rt_lane() { return (int)__lane_id(); }

// Snapshot of the map topology: every leaf, its 128^3 and 4096^3 parents, and
// the leaf-granular bounding box of active voxels (evalActiveBoundingBox with
// visitVoxels = false). One block-level min/max reduction per workgroup.
__global__ __launch_bounds__(256) void k_snapshot(MapView mp, uint32_t leaf_count,
                                                  RtView rt, int* __restrict__ bbox, TileView tv) {
  __shared__ int smin[3][4], smax[3][4];
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < leaf_count; i += G) {
    const uint64_t key = mp.leaf_keys[i];
    int o[3];
    key_origin(key, o);
    // an inactive tile is neither a leaf nor an active value for the HDDA; an active
    // tile is a hit at the leaf level, as a leaf is
    if (!tile_at(tv, i) || tile_active(tv, i)) set_insert(rt.leaf, rt.leaf_mask, key);
    set_insert(rt.n128, rt.n128_mask, node_key(o[0], o[1], o[2], 7));
    set_insert(rt.n4096, rt.n4096_mask, node_key(o[0], o[1], o[2], 12));
    uint64_t any = 0;
    for (int j = 0; j < 8; ++j) any |= mp.active[(size_t)i * 8 + j];
    if (any) {
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        lo[a] = min(lo[a], o[a]);
        hi[a] = max(hi[a], o[a] + 7);
      }
    }
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    for (int off = 32; off > 0; off >>= 1) {
      lo[a] = min(lo[a], __shfl_xor(lo[a], off, 64));
      hi[a] = max(hi[a], __shfl_xor(hi[a], off, 64));
    }
  }
  const int w = threadIdx.x >> 6;
  if (rt_lane() == 0) {
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      smin[a][w] = lo[a];
      smax[a][w] = hi[a];
    }
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int a = threadIdx.x;
    int l = INT_MAX, h = INT_MIN;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) {
      l = min(l, smin[a][k]);
      h = max(h, smax[a][k]);
    }
    if (l != INT_MAX) atomicMin(&bbox[a], l);
    if (h != INT_MIN) atomicMax(&bbox[3 + a], h);
  }
}

// Root children created by updates that created no leaf (value and state
// unchanged, e.g. probability 0.5): kept so the snapshot sees them too.
__global__ __launch_bounds__(256) void k_copy_set(const unsigned long long* __restrict__ src,
                                                  uint32_t src_cap, unsigned long long* dst,
                                                  uint32_t dst_mask) {
  const uint32_t G = gridDim.x * blockDim.x;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < src_cap; i += G)
    if (src[i] != kEmpty) set_insert(dst, dst_mask, src[i]);
}

// One lane per query ray. Mirrors VDBMapping::raytrace: normalise (multiply by
// 1/length unless length <= 1e-7), scale by the max length, continuous
// worldToIndex, clip to the padded bbox (result ignored), VolumeHDDA march,
// then a voxel DDA over the first leaf run that stops at an active voxel.
__global__ __launch_bounds__(256) void k_raytrace(RtView rt, MapView mp, double inv, double res,
                                                  int64_t n, const double* __restrict__ origins,
                                                  const double* __restrict__ dirs,
                                                  const double* __restrict__ lens,
                                                  uint8_t* __restrict__ success,
                                                  double* __restrict__ ends,
                                                  unsigned int* __restrict__ err) {
  const int64_t G = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += G) {
    double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(fabs(0.0 - len) <= 1e-7)) {
      const double r = 1.0 / len;
      d[0] *= r;
      d[1] *= r;
      d[2] *= r;
    }
    const double L = lens[i];
    double eye[3], dir[3], idir[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      d[a] *= L;
      eye[a] = origins[3 * i + a] * inv;
      dir[a] = d[a] * inv;
      idir[a] = 1.0 / dir[a];
    }
    // A query whose DDA cannot end (non-finite numbers) or would take more
    // than kMaxQuerySteps voxel steps is rejected instead of marched.
    const double span = fabs(dir[0]) + fabs(dir[1]) + fabs(dir[2]);
    if (!(isfinite(eye[0]) && isfinite(eye[1]) && isfinite(eye[2]) && isfinite(span) &&
          span < (double)kMaxQuerySteps)) {
      atomicOr(err, 1u);
      success[i] = 0;
      ends[3 * i] = ends[3 * i + 1] = ends[3 * i + 2] = NAN;
      continue;
    }
    // every DDA level takes at most span/dim + 3 steps over t in [0, 1]
    int budget = 4 * ((int)span + 8);
    double rt0 = 0.0, rt1 = 1.0;
    {  // Ray::clip to the padded active bbox
      double c0 = rt0, c1 = rt1;
      bool hit = true;
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        if (!hit) break;
        double ta = ((double)rt.bmin[a] - eye[a]) * idir[a];
        double tb = ((double)rt.bmax[a] - eye[a]) * idir[a];
        if (ta > tb) {
          const double t = ta;
          ta = tb;
          tb = t;
        }
        if (ta > c0) c0 = ta;
        if (tb < c1) c1 = tb;
        if (c0 > c1) hit = false;
      }
      if (hit) {
        rt0 = c0;
        rt1 = c1;
      }
    }
    double s0 = -1, s1 = -1;  // TimeSpan of the first leaf run
    if ((rt1 - rt0) > (double)1e-5f) {  // Ray::valid
      bool done = false;
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
                  if (span_valid(s0, s1)) {
                    done = true;
                    break;
                  }
                  s0 = s1 = -1;
                }
              } while (d3.step() && --budget > 0);
              if (done) break;
              if (s0 >= 0) s1 = d3.t1;
            } else if (s0 >= 0) {
              s1 = d7.t0;
              if (span_valid(s0, s1)) {
                done = true;
                break;
              }
              s0 = s1 = -1;
            }
          } while (d7.step() && --budget > 0);
          if (done) break;
          if (s0 >= 0) s1 = d7.t1;
        } else if (s0 >= 0) {
          s1 = d12.t0;
          if (span_valid(s0, s1)) {
            done = true;
            break;
          }
          s0 = s1 = -1;
        }
      } while (d12.step() && --budget > 0);
      if (!done && s0 >= 0) s1 = d12.t1;
    }
    double* out = ends + 3 * i;
    if (span_valid(s0, s1)) {
      RtDda fd;
      fd.init(eye, dir, idir, s0, s1, 0);
      ActiveProbe probe;
      while (fd.step() && !probe.on(mp, fd.vox[0], fd.vox[1], fd.vox[2]) && --budget > 0) {
      }
      if (budget <= 0) atomicOr(err, 2u);
#pragma unroll
      for (int a = 0; a < 3; ++a) out[a] = (double)fd.vox[a] * res;
      success[i] = 1;
    } else {
#pragma unroll
      for (int a = 0; a < 3; ++a) out[a] = (eye[a] + dir[a]) * res;
      success[i] = 0;
    }
  }
}

hipError_t launch_snapshot(hipStream_t s, MapView mp, uint32_t leaf_count, RtView rt, int* bbox, TileView tv) {
  const int grid = (int)std::min<uint32_t>(4096, (leaf_count + 255) / 256 + 1);
  hipLaunchKernelGGL(k_snapshot, dim3(grid), dim3(256), 0, s, mp, leaf_count, rt, bbox, tv);
  return hipGetLastError();
}

hipError_t launch_copy_set(hipStream_t s, const unsigned long long* src, uint32_t src_cap,
                           unsigned long long* dst, uint32_t dst_mask) {
  const int grid = (int)std::min<uint32_t>(4096, (src_cap + 255) / 256);
  hipLaunchKernelGGL(k_copy_set, dim3(grid), dim3(256), 0, s, src, src_cap, dst, dst_mask);
  return hipGetLastError();
}

hipError_t launch_raytrace(hipStream_t s, int grid, RtView rt, MapView mp, double inv, double res,
                           int64_t n, const double* origins, const double* dirs,
                           const double* lens, uint8_t* success, double* ends,
                           unsigned int* err) {
  hipLaunchKernelGGL(k_raytrace, dim3(grid), dim3(256), 0, s, rt, mp, inv, res, n, origins, dirs,
                     lens, success, ends, err);
  return hipGetLastError();
}

}  // namespace vdbgpu
