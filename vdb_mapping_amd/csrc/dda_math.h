// Exact integer arithmetic of the voxel walk (VDBMapping.hpp:550-566 over
// math::DDA): the order of crossing events, the walk state at any step, and
// the cut of a ray into per-brick segments. Header-only and free of HIP
// types so the same code runs in the kernels and in the CPU checks
// (tests/cpp/segments_check.cpp).
//
// An axis with |delta| = n (ray from a voxel centre) crosses voxel faces at
// the times math::DDA accumulates: T(n, 0) = 0.5*fl(1/n),
// T(n, j) = fl(T(n, j-1) + fl(1/n)); the rational value is (2j+1)/(2n). For
// n <= 8192, |T(n, j) - (2j+1)/(2n)| <= 2n * 2^-53 < 1.9e-12, while two
// different rational times differ by at least 1/(2 n_a n_b) >= 7.4e-9. So the
// double-precision order equals the rational order, which is decided exactly
// by the integer sign of (2c_b+1) n_a - (2c_a+1) n_b, except at exact rational
// ties. There the doubles decide (read from the crossing table `tbl`, row n
// at n(n-1)/2), and equal doubles go to the larger axis index
// (math::MinIndex). Axes with equal n have identical double sequences, so
// their ties need no table read.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define VDB_HD __host__ __device__ __forceinline__
#define VDB_UNROLL _Pragma("unroll")
#else
#define VDB_HD inline
#define VDB_UNROLL
#endif

namespace vdbgpu {

VDB_HD double crossing_time(const double* __restrict__ tbl, uint32_t n, uint32_t c) {
  return tbl[(((uint64_t)n * (n - 1)) >> 1) + c];
}

// Event (b, cb) precedes event (a, ca) in DDA order. Every product below is
// under (2*8192+1)*8192 < 2^31: 32-bit arithmetic (no emulated 64-bit ops).
VDB_HD bool ev_before(const double* __restrict__ tbl, int b, uint32_t nb, uint32_t cb, int a,
                      uint32_t na, uint32_t ca) {
  const int32_t d = (int32_t)((2 * cb + 1) * na) - (int32_t)((2 * ca + 1) * nb);
  if (d != 0) return d < 0;
  if (nb == na) return b > a;
  const double fb = crossing_time(tbl, nb, cb);
  const double fa = crossing_time(tbl, na, ca);
  return fb < fa || (fb == fa && b > a);
}

// Number of axis-b crossings (b has |delta| nb) that precede event (a, ca):
// #{c : (2c+1) na < (2ca+1) nb} plus the rational tie, if any, when the
// tie-break puts it first.
VDB_HD uint32_t count_before(const double* __restrict__ tbl, int b, uint32_t nb, int a, uint32_t na,
                             uint32_t ca) {
  if (nb == 0) return 0;
  const uint32_t X = (2 * ca + 1) * nb;  // < 2^28
  const uint32_t q = X / na, r = X - q * na;
  uint32_t cnt = r == 0 ? q / 2 : (q + 1) / 2;
  if (cnt > nb) cnt = nb;
  if (r == 0 && (q & 1)) {
    const uint32_t ct = (q - 1) / 2;
    if (ct < nb && ev_before(tbl, b, nb, ct, a, na, ca)) ++cnt;
  }
  return cnt;
}

// Crossing counts c[3] after k steps of a ray with |delta| n[3] (k <= sum n):
// the first k events of the merged crossing sequences. Starts from the
// rational estimate at t = k / N (off by at most one event) and repairs it
// until every included event precedes every excluded one; each repair moves
// one misplaced event, so a handful of rounds suffice. Returns false only if
// the repair did not converge (reported as a kernel error, never silently).
VDB_HD bool dda_enter(const double* __restrict__ tbl, const uint32_t n[3], uint32_t k, uint32_t c[3]) {
  const uint32_t N = n[0] + n[1] + n[2];
  int sum = 0;
  for (int a = 0; a < 3; ++a) {
    // #{j : (2j+1)/(2n) <= k/N} = floor((2 n k + N) / (2N)); 2 n k + N < 2^31
    const uint32_t g = (2 * n[a] * k + N) / (2 * N);
    c[a] = (uint32_t)(g < n[a] ? g : n[a]);
    sum += (int)c[a];
  }
  for (int round = 0; round < 64; ++round) {
    // included event that comes last (amax) and excluded event that comes first (amin)
    int amax = -1, amin = -1;
    uint32_t nmax = 0, cmax = 0, nmin = 0, cmin = 0;
    for (int a = 0; a < 3; ++a) {
      if (c[a] > 0 && (amax < 0 || ev_before(tbl, amax, nmax, cmax - 1, a, n[a], c[a] - 1))) {
        amax = a;
        nmax = n[a];
        cmax = c[a];
      }
      if (c[a] < n[a] && (amin < 0 || ev_before(tbl, a, n[a], c[a], amin, nmin, cmin))) {
        amin = a;
        nmin = n[a];
        cmin = c[a];
      }
    }
    int dec = -1, inc = -1;
    if (sum > (int)k) {
      dec = amax;
    } else if (sum < (int)k) {
      inc = amin;
    } else if (amax >= 0 && amin >= 0 && ev_before(tbl, amin, nmin, cmin, amax, nmax, cmax - 1)) {
      dec = amax;
      inc = amin;
    } else {
      return true;
    }
    for (int a = 0; a < 3; ++a) {
      if (a == dec) c[a]--;
      if (a == inc) c[a]++;
    }
    sum += (inc >= 0) - (dec >= 0);
  }
  return false;
}

// Integer DDA state of one ray: crossing counts and the pairwise order
// differences D_yx = (2c_y+1) n_x - (2c_x+1) n_y, D_zx, D_zy.
struct IntDda {
  uint32_t n0, n1, n2, c0, c1, c2;
  int dyx, dzx, dzy;
  VDB_HD void init(const uint32_t n[3], const uint32_t c[3]) {
    n0 = n[0];
    n1 = n[1];
    n2 = n[2];
    c0 = c[0];
    c1 = c[1];
    c2 = c[2];
    dyx = (int)((2 * c1 + 1) * n0) - (int)((2 * c0 + 1) * n1);
    dzx = (int)((2 * c2 + 1) * n0) - (int)((2 * c0 + 1) * n2);
    dzy = (int)((2 * c2 + 1) * n1) - (int)((2 * c1 + 1) * n2);
  }
  // Float tie-break: does (b, cb) come no later than (a, ca) under MinIndex?
  VDB_HD bool tie_le(const double* __restrict__ tbl, uint32_t nb, uint32_t cb, uint32_t na,
                     uint32_t ca) const {
    if (nb == na) return true;
    if (cb >= nb) return false;
    if (ca >= na) return true;
    return crossing_time(tbl, nb, cb) <= crossing_time(tbl, na, ca);
  }
  // Axis of the next step (0, 1, 2).
  VDB_HD int next_axis(const double* __restrict__ tbl) const {
    const bool yx = dyx < 0 || (dyx == 0 && tie_le(tbl, n1, c1, n0, c0));
    const int dzm = yx ? dzy : dzx;
    const bool zm =
        dzm < 0 || (dzm == 0 && (yx ? tie_le(tbl, n2, c2, n1, c1) : tie_le(tbl, n2, c2, n0, c0)));
    return zm ? 2 : (yx ? 1 : 0);
  }
#if defined(__HIPCC__)
  // next_axis for a wave walking many rays: without an exact tie (dyx == 0 or
  // the deciding z difference == 0) the axis follows from the signs alone; the
  // tie rule (which may read the crossing table) runs only when some lane of
  // the wave meets a tie, for all lanes at once (it is exact for every lane),
  // so the common step has no per-lane branches.
  // The step is returned as two flags: z (axis 2) or else y (axis 1) or else x.
  __device__ __forceinline__ void next_axis_wave(const double* __restrict__ tbl, bool& z, bool& y) const {
    y = dyx < 0;
    const int dzm = y ? dzy : dzx;
    z = dzm < 0;
    if (__any(dyx == 0 || dzm == 0)) {
      const int a = next_axis(tbl);
      z = a == 2;
      y = a == 1;
    }
  }
  // advance by the flags of next_axis_wave, with selects instead of branches.
  __device__ __forceinline__ void advance_sel(bool z, bool y) {
    const bool x = !z && !y;
    y = y && !z;
    c0 += x;
    c1 += y;
    c2 += z;
    dyx += z ? 0 : (y ? 2 * (int)n0 : -2 * (int)n1);
    dzx += y ? 0 : (z ? 2 * (int)n0 : -2 * (int)n2);
    dzy += x ? 0 : (z ? 2 * (int)n1 : -2 * (int)n2);
  }
#endif
  VDB_HD void advance(int a) {
    if (a == 0) {
      ++c0;
      dyx -= 2 * (int)n1;
      dzx -= 2 * (int)n2;
    } else if (a == 1) {
      ++c1;
      dyx += 2 * (int)n0;
      dzy -= 2 * (int)n2;
    } else {
      ++c2;
      dzx += 2 * (int)n0;
      dzy += 2 * (int)n1;
    }
  }
};

// ------------------------------------------------------------ brick cut --
// Space is cut into bricks of 32^3 voxels (4^3 leaves). The walk of one ray
// (visits 0 .. cnt-1, visit 0 = the origin voxel s, visit cnt-1 = the end
// voxel e) enters a new brick exactly when an axis step crosses a multiple
// of 32. The global index of the step that makes crossing ca of axis a is
// G = 1 + ca + sum over the other axes of count_before(...), so the segment
// boundaries, and the crossing counts at each boundary, follow in O(1) per
// boundary without walking.
constexpr int kBrickLog = 5;
constexpr int kBrickDim = 1 << kBrickLog;
constexpr uint32_t kNoCross = 0xFFFFFFFFu;

struct Segment {
  int bx, by, bz;      // brick coordinates (voxel >> 5)
  uint32_t c[3];       // crossing counts at the segment's first visit
  uint32_t m;          // visits in the segment (0: a zero-length ray's endpoint)
  bool last;           // holds the ray's last visit (its end voxel)
};

struct SegEnum {
  uint32_t n[3];
  int step[3];
  int b[3];
  uint32_t cur[3];     // crossing counts at the current segment's first visit
  uint32_t nxt[3];     // next brick-boundary crossing index per axis (kNoCross: none)
  uint32_t G[3];       // global step index of that crossing
  uint32_t g0;         // first visit of the current segment
  uint32_t cnt;
  bool done;

  // Global step index of crossing nxt[a] of axis a (a is a constant after
  // unrolling: the state stays in registers).
  VDB_HD uint32_t event_index(const double* __restrict__ tbl, int a) const {
    const int o1 = a == 0 ? 1 : 0, o2 = a == 2 ? 1 : 2;
    return 1 + nxt[a] + count_before(tbl, o1, n[o1], a, n[a], nxt[a]) +
           count_before(tbl, o2, n[o2], a, n[a], nxt[a]);
  }
  // Step past the brick crossing of axis A: crossing counts after it (axis A
  // made crossing nxt[A], the others theirs before it), brick, next crossing.
  template <int A>
  VDB_HD void take(const double* __restrict__ tbl) {
    constexpr int O1 = A == 0 ? 1 : 0, O2 = A == 2 ? 1 : 2;
    const uint32_t ca = nxt[A];
    cur[O1] = count_before(tbl, O1, n[O1], A, n[A], ca);
    cur[O2] = count_before(tbl, O2, n[O2], A, n[A], ca);
    cur[A] = ca + 1;
    b[A] += step[A];
    nxt[A] = ca + kBrickDim < n[A] ? ca + kBrickDim : kNoCross;
    G[A] = nxt[A] != kNoCross ? event_index(tbl, A) : kNoCross;
  }
  // s: origin voxel, e: end voxel, cnt = 1 + |e-s|_1 (0 if e == s).
  VDB_HD void init(const double* __restrict__ tbl, const int s[3], const int e[3], uint32_t count) {
    cnt = count;
    g0 = 0;
    done = false;
    VDB_UNROLL
    for (int a = 0; a < 3; ++a) {
      const int d = e[a] - s[a];
      n[a] = (uint32_t)(d < 0 ? -d : d);
      step[a] = d > 0 ? 1 : (d < 0 ? -1 : 0);
      b[a] = (count == 0 ? e[a] : s[a]) >> kBrickLog;
      cur[a] = 0;
      const uint32_t lo = (uint32_t)s[a] & (kBrickDim - 1);
      // crossing index whose step enters the next brick along this axis
      const uint32_t first = step[a] > 0 ? (kBrickDim - 1) - lo : lo;
      nxt[a] = (count != 0 && step[a] != 0 && first < n[a]) ? first : kNoCross;
    }
    VDB_UNROLL
    for (int a = 0; a < 3; ++a) G[a] = nxt[a] != kNoCross ? event_index(tbl, a) : kNoCross;
  }
  // Next segment of the ray; false when the ray has no more.
  VDB_HD bool next(const double* __restrict__ tbl, Segment& sg) {
    if (done) return false;
    sg.bx = b[0];
    sg.by = b[1];
    sg.bz = b[2];
    sg.c[0] = cur[0];
    sg.c[1] = cur[1];
    sg.c[2] = cur[2];
    // the axis whose pending brick crossing comes first (kNoCross = none)
    int a = -1;
    uint32_t ga = kNoCross;
    VDB_UNROLL
    for (int k = 0; k < 3; ++k) {
      if (G[k] < ga) {
        a = k;
        ga = G[k];
      }
    }
    if (a < 0) {
      sg.m = cnt - g0;
      sg.last = true;
      done = true;
      return true;
    }
    sg.m = ga - g0;
    sg.last = false;
    // one instance per axis: every index is a constant, the state stays in registers
    if (a == 0)
      take<0>(tbl);
    else if (a == 1)
      take<1>(tbl);
    else
      take<2>(tbl);
    g0 = ga;
    return true;
  }
};

}  // namespace vdbgpu
