// Device pieces shared by the raytrace query and the fast-mode raycast: the
// double-precision math::DDA of openvdb (one instance per hierarchy level),
// the live-map active-state probe and the TimeSpan validity test.
#pragma once
#include <float.h>

#include "device_common.h"

namespace vdbgpu {

// math::DDA<Ray<double>, Log2Dim>.
struct RtDda {
  int vox[3];
  int stp[3];
  double nxt[3], dlt[3];
  double t0, t1;
  __device__ __forceinline__ void init(const double eye[3], const double dir[3],
                                       const double inv[3], double start, double max_t,
                                       int log2dim) {
    const int dim = 1 << log2dim;
    t0 = start;
    t1 = max_t;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      const double pos = eye[a] + dir[a] * t0;
      vox[a] = (int)floor(pos) & ~(dim - 1);
      if (dir[a] == 0.0) {
        stp[a] = 0;
        nxt[a] = DBL_MAX;
        dlt[a] = DBL_MAX;
      } else if (inv[a] > 0) {
        stp[a] = dim;
        nxt[a] = t0 + ((double)(vox[a] + dim) - pos) * inv[a];
        dlt[a] = (double)stp[a] * inv[a];
      } else {
        stp[a] = -dim;
        nxt[a] = t0 + ((double)vox[a] - pos) * inv[a];
        dlt[a] = (double)stp[a] * inv[a];
      }
    }
  }
  __device__ __forceinline__ bool step() {
    int a = 0;
    if (nxt[1] <= nxt[0]) a = 1;
    if (nxt[2] <= (a ? nxt[1] : nxt[0])) a = 2;
    double t;
    if (a == 0) {
      t = nxt[0];
      nxt[0] += dlt[0];
      vox[0] += stp[0];
    } else if (a == 1) {
      t = nxt[1];
      nxt[1] += dlt[1];
      vox[1] += stp[1];
    } else {
      t = nxt[2];
      nxt[2] += dlt[2];
      vox[2] += stp[2];
    }
    t0 = t;
    return t0 <= t1;
  }
  __device__ __forceinline__ double next() const { return fmin(t1, fmin(nxt[0], fmin(nxt[1], nxt[2]))); }
};

// Active state of a voxel in the live map; caches the last leaf looked up.
struct ActiveProbe {
  uint64_t key = kEmpty;
  uint32_t idx = kNone;
  __device__ __forceinline__ bool on(const MapView& mp, int x, int y, int z) {
    if (!coord_ok(x) || !coord_ok(y) || !coord_ok(z)) return false;
    const uint64_t k = leaf_key(x, y, z);
    if (k != key) {
      key = k;
      idx = kNone;
      uint32_t h = (uint32_t)fmix64(k) & mp.cap_mask;
      for (uint32_t p = 0; p <= mp.cap_mask; ++p) {
        const unsigned long long kk = mp.keys[h];
        if (kk == k) {
          idx = mp.idx[h];
          break;
        }
        if (kk == kEmpty) break;
        h = (h + 1) & mp.cap_mask;
      }
    }
    if (idx == kNone) return false;
    const int n = ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
    return (mp.active[(size_t)idx * 8 + (n >> 6)] >> (n & 63)) & 1;
  }
};

__device__ __forceinline__ bool span_valid(double t0, double t1) { return (t1 - t0) > 1e-9; }


// VolumeRayIntersector::setIndexRay: Ray::clip to the padded active bbox of
// the snapshot; the times change only when the ray hits the box.
__device__ __forceinline__ void rt_clip(const RtView& rt, const double eye[3], const double idir[3],
                                        double& rt0, double& rt1) {
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

}  // namespace vdbgpu
