// CPU check of the brick cut (vdb_mapping_amd/csrc/dda_math.h): for random
// and tie-heavy rays, the segments SegEnum derives in closed form must equal
// the brick runs of the step-by-step integer walk, with the same crossing
// counts at every segment start, and dda_enter must reach those counts from
// the segment's first visit index. The walk itself is pinned against the
// oracle's floating-point DDA by the parity tests; this program pins the cut.
//
// Build: g++ -O2 -std=c++17 -I vdb_mapping_amd/csrc tests/cpp/segments_check.cpp
// Usage: segments_check [n_rays] [seed]   (exit status 0 = all equal)
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "dda_math.h"

using namespace vdbgpu;

static std::vector<double> build_table(uint32_t nmax) {
  std::vector<double> t((size_t)nmax * (nmax + 1) / 2 + 1);
  for (uint32_t n = 1; n <= nmax; ++n) {
    const double d = 1.0 / (double)n;
    double x = 0.5 * d;
    for (uint32_t j = 0; j < n; ++j) {
      t[(((uint64_t)n * (n - 1)) >> 1) + j] = x;
      x = x + d;
    }
  }
  return t;
}

static int floor_div32(int v) { return v >> kBrickLog; }

static int check_ray(const double* tbl, const int s[3], const int e[3]) {
  uint32_t n[3];
  int st[3];
  for (int a = 0; a < 3; ++a) {
    const int d = e[a] - s[a];
    n[a] = (uint32_t)std::abs(d);
    st[a] = d > 0 ? 1 : (d < 0 ? -1 : 0);
  }
  const uint32_t cnt = (n[0] | n[1] | n[2]) ? 1 + n[0] + n[1] + n[2] : 0;
  // reference: step-by-step walk, a new segment whenever the brick changes
  std::vector<Segment> ref;
  if (cnt == 0) {
    ref.push_back(Segment{floor_div32(e[0]), floor_div32(e[1]), floor_div32(e[2]), {0, 0, 0}, 0, true});
  } else {
    IntDda d;
    const uint32_t c0[3] = {0, 0, 0};
    d.init(n, c0);
    int v[3] = {s[0], s[1], s[2]};
    Segment cur{floor_div32(v[0]), floor_div32(v[1]), floor_div32(v[2]), {0, 0, 0}, 0, false};
    for (uint32_t k = 0; k < cnt; ++k) {
      if (k > 0) {
        const int a = d.next_axis(tbl);
        d.advance(a);
        v[a] += st[a];
        const int b[3] = {floor_div32(v[0]), floor_div32(v[1]), floor_div32(v[2])};
        if (b[0] != cur.bx || b[1] != cur.by || b[2] != cur.bz) {
          ref.push_back(cur);
          cur = Segment{b[0], b[1], b[2], {d.c0, d.c1, d.c2}, 0, false};
        }
      }
      cur.m++;
    }
    if (v[0] != e[0] || v[1] != e[1] || v[2] != e[2]) {
      std::printf("walk does not end at e\n");
      return 1;
    }
    cur.last = true;
    ref.push_back(cur);
  }
  SegEnum en;
  en.init(tbl, s, e, cnt);
  Segment sg;
  size_t i = 0;
  uint32_t first = 0;
  while (en.next(tbl, sg)) {
    if (i >= ref.size()) {
      std::printf("extra segment\n");
      return 1;
    }
    const Segment& r = ref[i];
    if (sg.bx != r.bx || sg.by != r.by || sg.bz != r.bz || sg.m != r.m || sg.last != r.last ||
        sg.c[0] != r.c[0] || sg.c[1] != r.c[1] || sg.c[2] != r.c[2]) {
      std::printf("segment %zu differs: got brick (%d,%d,%d) c (%u,%u,%u) m %u last %d, walk brick "
                  "(%d,%d,%d) c (%u,%u,%u) m %u last %d\n",
                  i, sg.bx, sg.by, sg.bz, sg.c[0], sg.c[1], sg.c[2], sg.m, (int)sg.last, r.bx, r.by,
                  r.bz, r.c[0], r.c[1], r.c[2], r.m, (int)r.last);
      return 1;
    }
    if (first > 0) {
      uint32_t c[3];
      if (!dda_enter(tbl, n, first, c) || c[0] != sg.c[0] || c[1] != sg.c[1] || c[2] != sg.c[2]) {
        std::printf("dda_enter disagrees at visit %u\n", first);
        return 1;
      }
    }
    first += sg.m;
    ++i;
  }
  if (i != ref.size()) {
    std::printf("missing segments: %zu of %zu\n", i, ref.size());
    return 1;
  }
  return 0;
}

int main(int argc, char** argv) {
  const long n_rays = argc > 1 ? std::atol(argv[1]) : 200000;
  const unsigned seed = argc > 2 ? (unsigned)std::atol(argv[2]) : 12345u;
  const uint32_t nmax = 1100;
  const std::vector<double> tbl = build_table(nmax);
  std::mt19937_64 rng(seed);
  std::uniform_int_distribution<int> pos(-70000, 70000), len(-1000, 1000), small(-12, 12);
  long bad = 0, rays = 0;
  for (long r = 0; r < n_rays; ++r) {
    int s[3], e[3];
    for (int a = 0; a < 3; ++a) s[a] = pos(rng);
    const int kind = (int)(r % 6);
    if (kind == 0) {  // tie-heavy: components from a small set of multiples
      const int base = 1 + (int)(rng() % 40);
      for (int a = 0; a < 3; ++a) e[a] = s[a] + base * (int)((int)(rng() % 13) - 6);
    } else if (kind == 1) {  // equal magnitudes (diagonals)
      const int m = len(rng);
      for (int a = 0; a < 3; ++a) e[a] = s[a] + ((rng() & 1) ? m : -m);
    } else if (kind == 2) {  // short rays, often inside one brick
      for (int a = 0; a < 3; ++a) e[a] = s[a] + small(rng);
    } else if (kind == 3) {  // starts on a brick face
      for (int a = 0; a < 3; ++a) {
        s[a] = (s[a] & ~31) + ((rng() & 1) ? 0 : 31);
        e[a] = s[a] + len(rng);
      }
    } else {
      for (int a = 0; a < 3; ++a) e[a] = s[a] + len(rng);
    }
    ++rays;
    if (check_ray(tbl.data(), s, e)) {
      std::printf("  ray s (%d,%d,%d) e (%d,%d,%d)\n", s[0], s[1], s[2], e[0], e[1], e[2]);
      if (++bad > 10) break;
    }
  }
  std::printf("%ld rays checked, %ld differ\n", rays, bad);
  return bad ? 1 : 0;
}
