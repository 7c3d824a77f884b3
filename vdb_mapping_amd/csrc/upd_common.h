// Device helpers shared by the kernels that build update sets and edit the
// map (integrate_kernels.hip, edit_kernels.hip).
#pragma once

#include <limits.h>

#include "device_common.h"

namespace vdbgpu {

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Rounded world->index of one component (VDBMapping.hpp:612-631): shift by
// res/2 unless fmod(c, res) == 0, multiply by the stored 1/res, floor.
// fmod(c, res) == 0 exactly when c = k res for an integer k. Then c * inv is
// within k * 2^-51 of k, so k = rint(c * inv) and fma(-k, res, c) is exactly
// 0; for any other c, c - k res is a nonzero multiple of 2^-1074 and its
// rounding is not 0. (For |c / res| >= 2^51 the index is outside int either
// way; NaN and infinities compare unequal to 0 in both forms.) Same result as
// the fmod test, without fmod's division and reduction loop.
__device__ __forceinline__ int world_to_index1(double c, double res, double inv) {
  const double k = rint(c * inv);
  if (fma(-k, res, c) != 0.0) c = c + res / 2.0;
  const double f = floor(c * inv);
  return (f >= -2147483648.0 && f < 2147483648.0) ? (int)f : INT_MIN;
}

// Find or insert `key` in an update set; returns its slot, or kNone after
// kMaxProbe probes (the host then grows the table and re-runs the scan, which
// is exact because the update set is an OR of idempotent marks).
constexpr uint32_t kMaxProbe = 512;
__device__ __forceinline__ uint32_t upd_slot(const UpdView& u, uint64_t key, unsigned int* ovf) {
  uint32_t h = (uint32_t)fmix64(key) & u.cap_mask;
  const uint32_t limit = u.cap_mask < kMaxProbe ? u.cap_mask + 1 : kMaxProbe;
  for (uint32_t probe = 0; probe < limit; ++probe) {
    const unsigned long long k =
        __hip_atomic_load(&u.keys[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (k == key) return h;
    if (k == kEmpty) {
      const unsigned long long prev = atomicCAS(&u.keys[h], kEmpty, (unsigned long long)key);
      if (prev == kEmpty || prev == key) return h;
    }
    h = (h + 1) & u.cap_mask;
  }
  atomicOr(ovf, kOvfHard);
  return kNone;
}

// OR `bits` into mask word `word` of a slot; skip the atomic when a plain
// read already shows every bit (bits are never cleared during traversal, so
// a stale read can only cause an unneeded atomic, never a lost bit).
__device__ __forceinline__ void upd_or(const UpdView& u, uint32_t slot, int word, uint64_t bits) {
  if (slot == kNone || bits == 0) return;
  unsigned long long* p = &u.masks[(size_t)slot * 16 + word];
  const unsigned long long cur = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if ((cur & bits) != bits) atomicOr(p, (unsigned long long)bits);
}

// Visited (and, for an endpoint, hit) mark of one voxel in an update set.
__device__ __forceinline__ void mark_voxel(const UpdView& u, unsigned int* ovf, int x, int y, int z,
                                           bool hit) {
  const uint32_t slot = upd_slot(u, leaf_key(x, y, z), ovf);
  const int w = x & 7;
  const uint64_t b = 1ull << (((y & 7) << 3) | (z & 7));
  upd_or(u, slot, w, b);
  if (hit) upd_or(u, slot, 8 + w, b);
}

// Workgroup-aggregated append: returns the position of this lane's element
// in a global list (valid lanes only), one atomic per workgroup call. Must be
// called by every thread of the workgroup.
__device__ __forceinline__ uint32_t block_append(bool valid, unsigned int* counter,
                                                 uint32_t* s_wave, uint32_t* s_base) {
  const uint64_t m = __ballot(valid);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) s_wave[w] = (uint32_t)__popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t tot = 0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) {
      const uint32_t c = s_wave[i];
      s_wave[i] = tot;
      tot += c;
    }
    *s_base = tot ? atomicAdd(counter, tot) : 0u;
  }
  __syncthreads();
  const uint32_t pos = *s_base + s_wave[w] + (uint32_t)__popcll(m & ((1ull << lane_id()) - 1));
  __syncthreads();
  return pos;
}

}  // namespace vdbgpu
