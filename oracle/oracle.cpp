// CPU restatement of vdb_mapping's occupancy-integration path.
//
// TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline. The
// product never links, loads or calls this file (see oracle.h).
//
// Reference files cited: include/vdb_mapping/VDBMapping.hpp (VM) and
// include/vdb_mapping/OccupancyVDBMapping.hpp (OCC). The reference delegates
// its voxel arithmetic to OpenVDB (not part of the reference tree and not
// installed here). The OpenVDB semantics restated below are marked [U#]; each
// is a choice recalled from OpenVDB's published headers and kept behind one
// macro so that a different upstream version can be selected:
//   U1 ScaleMap::applyInverseMap multiplies by a stored 1/scale.
//   U2 Vec3::unit() divides each component by length().
//   U3 Vec3::normalize() multiplies each component by 1/length().
//   U4 math::MinIndex breaks ties toward the larger axis index.
//   U5 math::DDA::init/step as restated in struct Dda below.
//   U6 Ray::clip(CoordBBox) uses the bbox corners as given (no extra +1).
//   U7 VolumeHDDA descends root-child (4096) -> 128 -> 8 and tests node existence.
//   U8 a modify on a missing leaf creates it only if the op changes value or state.
//   U9 evalActiveBoundingBox(bbox, false) expands by the node bbox of each leaf
//      that holds an active voxel.
//   U10 Ray::valid() uses eps = Delta<float> (1e-5f); TimeSpan::valid() uses
//      Delta<double> (1e-9).
//   U11 VolumeHDDA::hits (fast-mode raycast) walks like the march without its
//      Ray::valid() pre-check and early exit, pushes every valid span, and
//      closes an open span at the root level's max time.
//   U12 Grid::cbeginValueAll visits the inactive tiles of internal nodes (one
//      value per tile, at the tile's origin) besides the leaf voxels.
//   U13 pcl::transformPointCloud with a double matrix computes each output
//      coordinate as ((m0 p0 + m1 p1) + m2 p2) + m3 in double, rounded to float.
//   U14 tools::dilateActiveValues / erodeActiveValues (NN_FACE_EDGE_VERTEX,
//      EXPAND_TILES, no pruning): a voxel outside every leaf is inactive; a
//      dilation step creates a leaf exactly where it activates a voxel in none;
//      erosion keeps every leaf, emptied or not.
//   U15 pruneGrid (tools::prune, tolerance 0) is restated at the leaf level: a
//      leaf of 512 equal values, all on or all off, becomes a tile; coarser
//      internal-node tiles are not formed (they change no leaf set and no
//      raytrace run, only the saveMapToPCD list of a fully occupied 128^3 block).
//      A tile becomes a leaf again at the first write that changes a voxel's
//      value or state. cbeginValueOn yields an active tile once, at its origin
//      (saveMapToPCD, applyMapSectionUpdateGrid), placed where its leaf would be.
//
// Compile with -ffp-contract=off (no FMA contraction): the reference is built
// for baseline x86-64 (Release, no -march flags, VM CMakeLists.txt:8-11),
// where no fused multiply-add exists, so every double op rounds on its own.

#include "oracle.h"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#ifndef ORC_U2_UNIT_DIVIDES
#define ORC_U2_UNIT_DIVIDES 1
#endif
#ifndef ORC_U6_CLIP_PAD
#define ORC_U6_CLIP_PAD 0
#endif

namespace {

constexpr int kCoordLimit = 1 << 23;  // |voxel coord| must stay below this (21-bit leaf keys)

inline bool in_range(const int32_t c[3]) {
  for (int a = 0; a < 3; ++a)
    if (c[a] < -kCoordLimit || c[a] >= kCoordLimit) return false;
  return true;
}

inline uint64_t leaf_key(int32_t x, int32_t y, int32_t z) {
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)(uint32_t)(x >> 3) & m) << 42 | ((uint64_t)(uint32_t)(y >> 3) & m) << 21 |
         ((uint64_t)(uint32_t)(z >> 3) & m);
}

inline int32_t sext21(uint64_t v) { return (int32_t)((int64_t)(v << 43) >> 43); }

inline void key_origin(uint64_t k, int32_t o[3]) {
  const uint64_t m = (1ull << 21) - 1;
  o[0] = sext21((k >> 42) & m) * 8;
  o[1] = sext21((k >> 21) & m) * 8;
  o[2] = sext21(k & m) * 8;
}

// Voxel offset inside an 8^3 leaf, OpenVDB LeafNode::coordToOffset layout.
inline int leaf_offset(int32_t x, int32_t y, int32_t z) {
  return ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
}

// Rounded world->index (VM:612-631). fmod != 0 (NaN counts as non-zero)
// shifts by res/2, then Coord::floor(ScaleMap::applyInverseMap) [U1].
inline void world_to_index(double res, double inv, const double w[3], int32_t out[3]) {
  for (int a = 0; a < 3; ++a) {
    double c = w[a];
    if (std::fmod(c, res) != 0.0) c = c + (res / 2.0);
    const double f = std::floor(c * inv);
    out[a] = (f >= -2147483648.0 && f < 2147483648.0) ? (int32_t)f : INT32_MIN;
  }
}

// math::MinIndex [U4]: ties go to the larger index.
inline int min_index(const double v[3]) {
  int r = 0;
  if (v[1] <= v[r]) r = 1;
  if (v[2] <= v[r]) r = 2;
  return r;
}

// math::DDA<Ray<double>, Log2Dim> [U5].
struct Dda {
  int32_t vox[3];
  int32_t step_[3];
  double next_[3], delta_[3];
  double t0, t1;
  void init(const double eye[3], const double dir[3], const double inv[3], double start,
            double max_t, int log2dim) {
    const int dim = 1 << log2dim;
    t0 = start;
    t1 = max_t;
    double pos[3];
    for (int a = 0; a < 3; ++a) pos[a] = eye[a] + dir[a] * t0;  // Ray::operator()(t0)
    for (int a = 0; a < 3; ++a) vox[a] = (int32_t)std::floor(pos[a]) & ~(dim - 1);
    for (int a = 0; a < 3; ++a) {
      if (dir[a] == 0.0) {
        step_[a] = 0;
        next_[a] = DBL_MAX;
        delta_[a] = DBL_MAX;
      } else if (inv[a] > 0) {
        step_[a] = dim;
        next_[a] = t0 + ((double)(vox[a] + dim) - pos[a]) * inv[a];
        delta_[a] = (double)step_[a] * inv[a];
      } else {
        step_[a] = -dim;
        next_[a] = t0 + ((double)vox[a] - pos[a]) * inv[a];
        delta_[a] = (double)step_[a] * inv[a];
      }
    }
  }
  bool step() {
    const int a = min_index(next_);
    t0 = next_[a];
    next_[a] += delta_[a];
    vox[a] += step_[a];
    return t0 <= t1;
  }
  double time() const { return t0; }
  double max_time() const { return t1; }
  double next() const { return std::min(t1, std::min(next_[0], std::min(next_[1], next_[2]))); }
};

// castRayIntoGrid (VM:550-566): Ray(s + 0.5, e - s, 0, 1), DDA<.,0> from t=0,
// visit, then step while the crossing time stays <= 1.
template <class Visit>
inline void cast_ray(const int32_t s[3], const int32_t e[3], Visit&& visit) {
  if (s[0] == e[0] && s[1] == e[1] && s[2] == e[2]) return;
  double eye[3], dir[3], inv[3];
  for (int a = 0; a < 3; ++a) {
    eye[a] = (double)s[a] + 0.5;
    dir[a] = (double)e[a] - (double)s[a];
    inv[a] = 1.0 / dir[a];
  }
  Dda d;
  d.init(eye, dir, inv, 0.0, 1.0, 0);
  do {
    visit(d.vox[0], d.vox[1], d.vox[2]);
  } while (d.step());
}

struct UpdLeaf {
  uint64_t vis[8];
  uint64_t hit[8];
};

// UpdateGridT (VM:89): active = visited, value = hit (VM:535, :563).
struct UpdateSet {
  std::unordered_map<uint64_t, UpdLeaf> leaves;
  uint64_t last_key = ~0ull;
  UpdLeaf* last = nullptr;  // one-entry cache, like a ValueAccessor
  UpdLeaf& get(uint64_t k) {
    if (k == last_key) return *last;
    auto it = leaves.find(k);
    if (it == leaves.end()) {
      UpdLeaf z;
      std::memset(&z, 0, sizeof z);
      it = leaves.emplace(k, z).first;
    }
    last_key = k;
    last = &it->second;
    return *last;
  }
  void clear() {
    leaves.clear();
    last_key = ~0ull;
    last = nullptr;
  }
};

struct Leaf {
  float v[512];
  uint64_t act[8];
  // pruneGrid (OCC:173) turned this leaf into a tile of value tv and state ts:
  // the tree holds no leaf here, but every voxel reads (tv, ts) [U15]
  bool tile;
  float tv;
  bool ts;
};

struct Source {
  std::string id;
  double max_range = 0;
  UpdateSet upd;
  bool has_isect = false;  // volume_ray_intersector set (VM:1440-1446), null after addInputSource
};

// Snapshot taken by updateVolumeRayIntersectors (VM:1436-1449): topology of
// the map plus the padded active bbox [U9].
struct Intersector {
  std::unordered_set<uint64_t> leaves, n128, n4096;
  int32_t bmin[3], bmax[3];
};

struct Map {
  double res, inv;
  // Config state (VM:1456-1469, OCC:59-89). Zero-initialised where the
  // reference leaves members uninitialised.
  bool config_set = false;
  double max_range = 0;
  bool fast_mode = false;
  float l_hit = 0, l_miss = 0, t_min = 0, t_max = 0, c_max = 0, c_min = 0;
  std::map<std::string, Source> sources;  // std::map: integrate order = id order (VM:380)
  std::unordered_map<uint64_t, Leaf> leaves;
  std::unordered_set<uint64_t> n4096;  // root children (created even by no-op modifies)
  std::unique_ptr<Intersector> isect;
  UpdateSet art;             // m_artificial_area_grid (VM:132): visited = artificial voxel
  bool art_present = false;  // m_artificial_areas_present (VM:1529)
  std::unordered_set<uint64_t> tiles;  // keys of leaves that stand for pruned tiles
};

// Keys of the 4096^3 (root child) and 128^3 internal nodes holding voxel (x,y,z).
uint64_t key4096(int32_t x, int32_t y, int32_t z) {
  return leaf_key((x >> 12) * 8, (y >> 12) * 8, (z >> 12) * 8);
}
uint64_t key128(int32_t x, int32_t y, int32_t z) {
  return leaf_key((x >> 7) * 8, (y >> 7) * 8, (z >> 7) * 8);
}

// updateOccupiedNode / updateFreeNode (OCC:92-117).
inline void op_hit(const Map& m, float& x, bool& act) {
  x += m.l_hit;
  if (x > m.t_max) {
    act = true;
    if (x > m.c_max) x = m.c_max;
  }
}
inline void op_miss(const Map& m, float& x, bool& act) {
  x += m.l_miss;
  if (x < m.t_min) {
    act = false;
    if (x < m.c_min) x = m.c_min;
  }
}

Leaf& touch_leaf(Map& m, uint64_t k) {
  auto it = m.leaves.find(k);
  if (it != m.leaves.end()) return it->second;
  Leaf z;
  std::memset(&z, 0, sizeof z);
  return m.leaves.emplace(k, z).first->second;
}

// A voxel write into a tile replaces the tile by a leaf exactly when it changes
// the voxel's value or state (OpenVDB's InternalNode set / modify calls compare
// with the tile first); the leaf then holds the tile value everywhere else, as
// here. So after each operation that writes voxels once, a tile whose voxels are
// no longer all (tv, ts) has become a leaf [U15].
void tile_check(Map& m) {
  for (auto it = m.tiles.begin(); it != m.tiles.end();) {
    Leaf& l = m.leaves.at(*it);
    bool same = true;
    const uint64_t word = l.ts ? ~0ull : 0ull;
    for (int w = 0; w < 8 && same; ++w) same = l.act[w] == word;
    for (int n = 0; n < 512 && same; ++n) same = l.v[n] == l.tv;
    if (same) {
      ++it;
    } else {
      l.tile = false;
      it = m.tiles.erase(it);
    }
  }
}

// pruneGrid (OCC:173: tools::prune, tolerance 0) at the leaf level: every leaf
// whose 512 values compare equal and whose voxels are all on or all off becomes
// a tile of that value and state [U15].
void prune_leaves(Map& m) {
  for (auto& kv : m.leaves) {
    Leaf& l = kv.second;
    const bool on = l.act[0] == ~0ull, off = l.act[0] == 0;
    bool uni = on || off;
    for (int w = 1; w < 8 && uni; ++w) uni = l.act[w] == l.act[0];
    for (int n = 1; n < 512 && uni; ++n) uni = l.v[n] == l.v[0];
    l.tile = uni;
    if (uni) {
      l.tv = l.v[0];
      l.ts = on;
      m.tiles.insert(kv.first);
    } else {
      m.tiles.erase(kv.first);
    }
  }
}

// setActiveState(xyz, true) over every artificial voxel (VM:785-789): a
// missing leaf is created (value 0) since the state changes.
void activate_artificial(Map& m) {
  for (auto& kv : m.art.leaves) {
    int32_t o[3];
    key_origin(kv.first, o);
    m.n4096.insert(key4096(o[0], o[1], o[2]));
    Leaf& leaf = touch_leaf(m, kv.first);
    for (int w = 0; w < 8; ++w) leaf.act[w] |= kv.second.vis[w];
  }
}

// updateMap (VM:731-792) over one source's update set. Leaves are created
// only when the op would change a background tile [U8]. An empty update set
// returns at once (VM:735-738); otherwise the artificial voxels are set
// active after the update (VM:785-789).
void apply_update(Map& m, UpdateSet& u) {
  if (u.leaves.empty()) return;
  for (auto& kv : u.leaves) {
    const uint64_t k = kv.first;
    const UpdLeaf& ul = kv.second;
    int32_t o[3];
    key_origin(k, o);
    m.n4096.insert(key4096(o[0], o[1], o[2]));
    auto it = m.leaves.find(k);
    Leaf* leaf = nullptr;
    if (it != m.leaves.end()) {
      leaf = &it->second;
    } else {
      bool need = false;
      for (int w = 0; w < 8 && !need; ++w) {
        const uint64_t hv = ul.vis[w] & ul.hit[w], mv = ul.vis[w] & ~ul.hit[w];
        float x = 0.0f;
        bool a = false;
        if (hv) {
          op_hit(m, x, a);
          need |= (x != 0.0f) || a;
        }
        x = 0.0f;
        a = false;
        if (mv) {
          op_miss(m, x, a);
          need |= (x != 0.0f) || a;
        }
      }
      if (!need) continue;
      Leaf z;
      std::memset(&z, 0, sizeof z);
      leaf = &m.leaves.emplace(k, z).first->second;
    }
    for (int w = 0; w < 8; ++w) {
      uint64_t vis = ul.vis[w];
      while (vis) {
        const int b = __builtin_ctzll(vis);
        vis &= vis - 1;
        const int n = w * 64 + b;
        bool a = (leaf->act[w] >> b) & 1;
        if ((ul.hit[w] >> b) & 1)
          op_hit(m, leaf->v[n], a);
        else
          op_miss(m, leaf->v[n], a);
        leaf->act[w] = (leaf->act[w] & ~(1ull << b)) | ((uint64_t)a << b);
      }
    }
  }
  tile_check(m);
  activate_artificial(m);
  tile_check(m);
}

// updateVolumeRayIntersectors (VM:1436-1449) for the state that raytrace needs.
void build_intersector(Map& m) {
  auto is = std::make_unique<Intersector>();
  bool any = false;
  for (auto& kv : m.leaves) {
    int32_t o[3];
    key_origin(kv.first, o);
    // an inactive tile is neither a leaf nor an active value for VolumeHDDA; an
    // active tile is a hit at the leaf level, like a leaf [U15]
    if (!kv.second.tile || kv.second.ts) is->leaves.insert(kv.first);
    is->n128.insert(key128(o[0], o[1], o[2]));
    bool active = false;
    for (int w = 0; w < 8; ++w) active |= kv.second.act[w] != 0;
    if (!active) continue;
    for (int a = 0; a < 3; ++a) {
      if (!any || o[a] < is->bmin[a]) is->bmin[a] = o[a];
      if (!any || o[a] + 7 > is->bmax[a]) is->bmax[a] = o[a] + 7;
    }
    any = true;
  }
  is->n4096 = m.n4096;
  if (!any) {
    // CoordBBox default: min = INT_MAX, max = INT_MIN (an empty box).
    for (int a = 0; a < 3; ++a) {
      is->bmin[a] = INT32_MAX;
      is->bmax[a] = INT32_MIN;
    }
  }
  for (int a = 0; a < 3; ++a) is->bmax[a] += 1;  // mBBox.max().offset(1)
  m.isect = std::move(is);
}

template <class Visit>
void cast_ray_fast(const Map& m, const Intersector& is, const int32_t s[3], const int32_t e[3],
                   Visit&& visit);

// raycastPointCloud (VM:466-539); fast mode casts with castRayIntoGridFast
// (VM:577-602) and only when the map is not empty (VM:495, :520-527).
int raycast(Map& m, Source& src, const float* xyz, int64_t n, int64_t stride,
            const double origin[3]) {
  if (!m.config_set) return ORC_NOT_CONFIGURED;  // VM:478-482
  const bool grid_empty = m.n4096.empty();
  // the reference dereferences a null intersector here (crash); reported instead
  if (m.fast_mode && !grid_empty && !(src.has_isect && m.isect)) return ORC_NO_INTERSECTOR;
  const double R = src.max_range;
  for (int a = 0; a < 3; ++a)
    if (std::isnan(origin[a])) return ORC_OK;  // every point is skipped (VM:505-510)
  int32_t s[3];
  world_to_index(m.res, m.inv, origin, s);
  // Pass 1: validate so that an out-of-range scan leaves the update set untouched.
  std::vector<int32_t> ends((size_t)n * 3);
  std::vector<uint8_t> flags((size_t)n);
  const char* base = (const char*)xyz;
  for (int64_t i = 0; i < n; ++i) {
    const float* pt = (const float*)(base + i * stride);
    double p[3] = {(double)pt[0], (double)pt[1], (double)pt[2]};
    if (std::isnan(p[0]) || std::isnan(p[1]) || std::isnan(p[2])) {
      flags[i] = 0;
      continue;
    }
    bool trunc = false;
    const double d[3] = {p[0] - origin[0], p[1] - origin[1], p[2] - origin[2]};
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (R > 0.0 && len > R) {  // VM:512-517
      for (int a = 0; a < 3; ++a) {
#if ORC_U2_UNIT_DIVIDES
        const double u = d[a] / len;
#else
        const double u = d[a] * (1.0 / len);
#endif
        p[a] = origin[a] + u * R;
      }
      trunc = true;
    }
    int32_t* e = &ends[(size_t)i * 3];
    world_to_index(m.res, m.inv, p, e);
    if (!in_range(e) || !in_range(s)) return ORC_OUT_OF_RANGE;
    flags[i] = trunc ? 1 : 2;
  }
  // Pass 2: traverse (VM:529) and mark endpoints (VM:533-536).
  UpdateSet& u = src.upd;
  for (int64_t i = 0; i < n; ++i) {
    if (!flags[i]) continue;
    const int32_t* e = &ends[(size_t)i * 3];
    auto mark = [&](int32_t x, int32_t y, int32_t z) {
      UpdLeaf& l = u.get(leaf_key(x, y, z));
      const int off = leaf_offset(x, y, z);
      l.vis[off >> 6] |= 1ull << (off & 63);
    };
    if (!m.fast_mode)
      cast_ray(s, e, mark);
    else if (!grid_empty)
      cast_ray_fast(m, *m.isect, s, e, mark);
    if (flags[i] == 2) {
      UpdLeaf& l = u.get(leaf_key(e[0], e[1], e[2]));
      const int off = leaf_offset(e[0], e[1], e[2]);
      l.vis[off >> 6] |= 1ull << (off & 63);
      l.hit[off >> 6] |= 1ull << (off & 63);
    }
  }
  return ORC_OK;
}

int64_t export_sorted_keys(const std::vector<uint64_t>& keys_in, std::vector<uint64_t>& out) {
  std::vector<std::pair<std::array<int32_t, 3>, uint64_t>> v;
  v.reserve(keys_in.size());
  for (uint64_t k : keys_in) {
    std::array<int32_t, 3> o;
    key_origin(k, o.data());
    v.emplace_back(o, k);
  }
  std::sort(v.begin(), v.end());
  out.clear();
  for (auto& p : v) out.push_back(p.second);
  return (int64_t)out.size();
}

// VolumeHDDA (openvdb/math/DDA.h) over the intersector snapshot [U7].
struct Span {
  double t0 = -1, t1 = -1;
  bool valid() const { return (t1 - t0) > 1e-9; }  // TimeSpan::valid [U10]
};

struct RayState {
  double eye[3], dir[3], inv[3];
  double t0, t1;
};

bool hdda_leaf(const Intersector& is, RayState& ray, Span& t) {
  Dda d;
  d.init(ray.eye, ray.dir, ray.inv, ray.t0, ray.t1, 3);
  do {
    if (is.leaves.count(leaf_key(d.vox[0], d.vox[1], d.vox[2]))) {
      if (t.t0 < 0) t.t0 = d.time();
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) return true;
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
  return false;
}

bool hdda_128(const Intersector& is, RayState& ray, Span& t) {
  Dda d;
  d.init(ray.eye, ray.dir, ray.inv, ray.t0, ray.t1, 7);
  do {
    if (is.n128.count(key128(d.vox[0], d.vox[1], d.vox[2]))) {
      ray.t0 = d.time();
      ray.t1 = d.next();
      if (hdda_leaf(is, ray, t)) return true;
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) return true;
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
  return false;
}

bool hdda_4096(const Intersector& is, RayState& ray, Span& t) {
  Dda d;
  d.init(ray.eye, ray.dir, ray.inv, ray.t0, ray.t1, 12);
  do {
    if (is.n4096.count(key4096(d.vox[0], d.vox[1], d.vox[2]))) {
      ray.t0 = d.time();
      ray.t1 = d.next();
      if (hdda_128(is, ray, t)) return true;
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) return true;
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
  return false;
}

// Ray::clip to the padded active bbox (VolumeRayIntersector::setIndexRay) [U6];
// the times change only when the ray hits the box.
void clip_ray(const Intersector& is, RayState& ray) {
  double t0 = ray.t0, t1 = ray.t1;
  bool hit = true;
  for (int a = 0; a < 3 && hit; ++a) {
    double lo = (double)is.bmin[a], hi = (double)is.bmax[a] + ORC_U6_CLIP_PAD;
    double ta = (lo - ray.eye[a]) * ray.inv[a];
    double tb = (hi - ray.eye[a]) * ray.inv[a];
    if (ta > tb) std::swap(ta, tb);
    if (ta > t0) t0 = ta;
    if (tb < t1) t1 = tb;
    if (t0 > t1) hit = false;
  }
  if (hit) {
    ray.t0 = t0;
    ray.t1 = t1;
  }
}

// VolumeHDDA::hits (openvdb/math/DDA.h): every run of existing leaves along
// the ray, as time spans [U7]; the same walk as the march above without its
// early exit.
void hits_leaf(const Intersector& is, RayState& ray, std::vector<Span>& out, Span& t) {
  Dda d;
  d.init(ray.eye, ray.dir, ray.inv, ray.t0, ray.t1, 3);
  do {
    if (is.leaves.count(leaf_key(d.vox[0], d.vox[1], d.vox[2]))) {
      if (t.t0 < 0) t.t0 = d.time();
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) out.push_back(t);
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
}

void hits_128(const Intersector& is, RayState& ray, std::vector<Span>& out, Span& t) {
  Dda d;
  d.init(ray.eye, ray.dir, ray.inv, ray.t0, ray.t1, 7);
  do {
    if (is.n128.count(key128(d.vox[0], d.vox[1], d.vox[2]))) {
      ray.t0 = d.time();
      ray.t1 = d.next();
      hits_leaf(is, ray, out, t);
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) out.push_back(t);
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
}

void hits_4096(const Intersector& is, RayState& ray, std::vector<Span>& out, Span& t) {
  Dda d;
  d.init(ray.eye, ray.dir, ray.inv, ray.t0, ray.t1, 12);
  do {
    if (is.n4096.count(key4096(d.vox[0], d.vox[1], d.vox[2]))) {
      ray.t0 = d.time();
      ray.t1 = d.next();
      hits_128(is, ray, out, t);
    } else if (t.t0 >= 0) {
      t.t1 = d.time();
      if (t.valid()) out.push_back(t);
      t.t0 = t.t1 = -1;
    }
  } while (d.step());
  if (t.t0 >= 0) t.t1 = d.max_time();
}

bool map_is_on(const Map& m, int32_t x, int32_t y, int32_t z) {
  if (x < -kCoordLimit || x >= kCoordLimit || y < -kCoordLimit || y >= kCoordLimit ||
      z < -kCoordLimit || z >= kCoordLimit)
    return false;
  auto it = m.leaves.find(leaf_key(x, y, z));
  if (it == m.leaves.end()) return false;
  const int off = leaf_offset(x, y, z);
  return (it->second.act[off >> 6] >> (off & 63)) & 1;
}

// castRayIntoGridFast (VM:577-602): Ray(s + 0.5, e - s, 0, 1) through the
// intersector (clip, hits), then for every span a voxel DDA over
// [t0, t1] that marks the voxels active in the live map. No e == s guard:
// a zero-length ray still probes the origin voxel.
template <class Visit>
void cast_ray_fast(const Map& m, const Intersector& is, const int32_t s[3], const int32_t e[3],
                   Visit&& visit) {
  RayState ray;
  for (int a = 0; a < 3; ++a) {
    ray.eye[a] = (double)s[a] + 0.5;
    ray.dir[a] = (double)e[a] - (double)s[a];
    ray.inv[a] = 1.0 / ray.dir[a];
  }
  ray.t0 = 0.0;
  ray.t1 = 1.0;
  clip_ray(is, ray);  // setIndexRay; its result is ignored (VM:590)
  std::vector<Span> spans;
  Span t;
  RayState r = ray;
  hits_4096(is, r, spans, t);
  if (t.valid()) spans.push_back(t);
  for (const Span& sp : spans) {
    Dda d;
    d.init(ray.eye, ray.dir, ray.inv, sp.t0, sp.t1, 0);
    do {
      if (map_is_on(m, d.vox[0], d.vox[1], d.vox[2])) visit(d.vox[0], d.vox[1], d.vox[2]);
    } while (d.step());
  }
}

}  // namespace

extern "C" {

void* orc_create(double resolution) {
  Map* m = new Map();
  m->res = resolution;
  m->inv = 1.0 / resolution;  // UniformScaleMap stores 1/scale [U1]
  return m;
}

void orc_destroy(void* h) { delete (Map*)h; }

// OccupancyVDBMapping::setConfig (OCC:59-89) calling VDBMapping::setConfig
// (VM:1456-1469). The base rejects max_range < 0 but the derived part still
// runs; a failed probability check returns after the base has already marked
// the map configured.
int orc_set_config(void* h, const orc_config* c) {
  Map& m = *(Map*)h;
  int status = ORC_OK;
  if (c->max_range < 0.0) {
    status = ORC_BAD_MAX_RANGE;
  } else {
    m.max_range = c->max_range;
    m.fast_mode = c->fast_mode != 0;
    m.config_set = true;
  }
  if (c->prob_miss > 0.5) return ORC_BAD_PROB_MISS;
  if (c->prob_hit < 0.5) return ORC_BAD_PROB_HIT;
  m.l_miss = (float)(std::log(c->prob_miss) - std::log(1 - c->prob_miss));
  m.l_hit = (float)(std::log(c->prob_hit) - std::log(1 - c->prob_hit));
  m.t_min = (float)(std::log(c->prob_thres_min) - std::log(1 - c->prob_thres_min));
  m.t_max = (float)(std::log(c->prob_thres_max) - std::log(1 - c->prob_thres_max));
  m.c_max = (float)(std::log(0.99) - std::log(0.01));
  m.c_min = (float)(std::log(0.01) - std::log(0.99));
  m.config_set = true;
  return status;
}

// addInputSource (VM:1352-1375): max_range 0 means the config's max_range.
int orc_add_source(void* h, const char* id, double max_range, double max_rate) {
  (void)max_rate;
  Map& m = *(Map*)h;
  Source& s = m.sources[id];
  s.id = id;
  s.max_range = (max_range == 0) ? m.max_range : max_range;
  s.upd.clear();
  s.has_isect = false;  // a new InputSource has no intersector yet
  return ORC_OK;
}

// accumulateUpdate (VM:316-346).
int orc_accumulate(void* h, const char* id, const float* xyz, int64_t n, int64_t stride,
                   const double origin[3]) {
  Map& m = *(Map*)h;
  auto it = m.sources.find(id);
  if (it == m.sources.end()) return ORC_UNKNOWN_SOURCE;
  if (!(it->second.max_range > 0)) return ORC_OK;
  return raycast(m, it->second, xyz, n, stride, origin);
}

// integrateUpdate (VM:375-387).
int orc_integrate(void* h) {
  Map& m = *(Map*)h;
  for (auto& kv : m.sources) {
    apply_update(m, kv.second.upd);
    kv.second.upd.clear();
  }
  if (m.fast_mode && !m.n4096.empty()) {
    build_intersector(m);  // updateVolumeRayIntersectors: every source gets a copy
    for (auto& kv : m.sources) kv.second.has_isect = true;
  }
  return ORC_OK;
}

// insertPointCloud (VM:399-406).
int orc_insert(void* h, const char* id, const float* xyz, int64_t n, int64_t stride,
               const double origin[3]) {
  const int s = orc_accumulate(h, id, xyz, n, stride, origin);
  orc_integrate(h);
  return s;
}

// resetMap (VM:174-186).
void orc_reset(void* h) {
  Map& m = *(Map*)h;
  m.leaves.clear();
  m.tiles.clear();
  m.n4096.clear();
  for (auto& kv : m.sources) kv.second.upd.clear();
}

int orc_empty(void* h) { return ((Map*)h)->n4096.empty() ? 1 : 0; }

void orc_logodds(void* h, float out[6]) {
  Map& m = *(Map*)h;
  out[0] = m.l_hit;
  out[1] = m.l_miss;
  out[2] = m.t_min;
  out[3] = m.t_max;
  out[4] = m.c_max;
  out[5] = m.c_min;
}

int64_t orc_leaf_count(void* h) {
  const Map& m = *(Map*)h;
  return (int64_t)(m.leaves.size() - m.tiles.size());
}

// Leaves only: tiles are not leaves of the tree.
int64_t orc_export_leaves(void* h, int32_t* origins, float* values, uint64_t* masks,
                          int64_t cap) {
  Map& m = *(Map*)h;
  std::vector<uint64_t> keys, sorted;
  for (auto& kv : m.leaves)
    if (!kv.second.tile) keys.push_back(kv.first);
  const int64_t n = export_sorted_keys(keys, sorted);
  if (n > cap) return -n;
  for (int64_t i = 0; i < n; ++i) {
    const Leaf& l = m.leaves.at(sorted[i]);
    key_origin(sorted[i], origins + 3 * i);
    std::memcpy(values + 512 * i, l.v, sizeof l.v);
    std::memcpy(masks + 8 * i, l.act, sizeof l.act);
  }
  return n;
}

// The leaf-level tiles pruneGrid made, sorted by origin: origin, value, state.
int64_t orc_export_tiles(void* h, int32_t* origins, float* values, uint8_t* active, int64_t cap) {
  Map& m = *(Map*)h;
  std::vector<uint64_t> keys(m.tiles.begin(), m.tiles.end()), sorted;
  const int64_t n = export_sorted_keys(keys, sorted);
  if (!origins && !values && !active) return n;
  if (n > cap) return -n;
  for (int64_t i = 0; i < n; ++i) {
    const Leaf& l = m.leaves.at(sorted[i]);
    key_origin(sorted[i], origins + 3 * i);
    values[i] = l.tv;
    active[i] = l.ts ? 1 : 0;
  }
  return n;
}

int64_t orc_update_leaf_count(void* h, const char* id) {
  Map& m = *(Map*)h;
  auto it = m.sources.find(id);
  return it == m.sources.end() ? -1 : (int64_t)it->second.upd.leaves.size();
}

int64_t orc_export_update(void* h, const char* id, int32_t* origins, uint64_t* vis,
                          uint64_t* hit, int64_t cap) {
  Map& m = *(Map*)h;
  auto it = m.sources.find(id);
  if (it == m.sources.end()) return -1;
  std::vector<uint64_t> keys, sorted;
  for (auto& kv : it->second.upd.leaves) keys.push_back(kv.first);
  const int64_t n = export_sorted_keys(keys, sorted);
  if (n > cap) return -n;
  for (int64_t i = 0; i < n; ++i) {
    const UpdLeaf& l = it->second.upd.leaves.at(sorted[i]);
    key_origin(sorted[i], origins + 3 * i);
    std::memcpy(vis + 8 * i, l.vis, sizeof l.vis);
    std::memcpy(hit + 8 * i, l.hit, sizeof l.hit);
  }
  return n;
}

float orc_get_value(void* h, const int32_t ijk[3], int32_t* active) {
  Map& m = *(Map*)h;
  *active = 0;
  if (!in_range(ijk)) return 0.0f;
  auto it = m.leaves.find(leaf_key(ijk[0], ijk[1], ijk[2]));
  if (it == m.leaves.end()) return 0.0f;
  const int off = leaf_offset(ijk[0], ijk[1], ijk[2]);
  *active = (int32_t)((it->second.act[off >> 6] >> (off & 63)) & 1);
  return it->second.v[off];
}

void orc_world_to_index(double res, const double w[3], int32_t out[3]) {
  world_to_index(res, 1.0 / res, w, out);
}

int64_t orc_cast_ray(const int32_t s[3], const int32_t e[3], int32_t* out, int64_t cap) {
  int64_t n = 0;
  cast_ray(s, e, [&](int32_t x, int32_t y, int32_t z) {
    if (n < cap) {
      out[3 * n] = x;
      out[3 * n + 1] = y;
      out[3 * n + 2] = z;
    }
    ++n;
  });
  return n;
}

// castRayIntoGridFast (VM:577-602) on its own: the map voxels the fast-mode
// cast from voxel s to voxel e marks, in visiting order (a voxel can repeat
// when spans share it). -1 without an intersector.
int64_t orc_cast_ray_fast(void* h, const int32_t s[3], const int32_t e[3], int32_t* out,
                          int64_t cap) {
  const Map& m = *(const Map*)h;
  if (!m.isect) return -1;
  int64_t n = 0;
  cast_ray_fast(m, *m.isect, s, e, [&](int32_t x, int32_t y, int32_t z) {
    if (n < cap) {
      out[3 * n] = x;
      out[3 * n + 1] = y;
      out[3 * n + 2] = z;
    }
    ++n;
  });
  return n;
}

// raytrace (VM:675-721).
int orc_raytrace(void* h, int64_t n, const double* origins, const double* dirs,
                 const double* lens, uint8_t* success, double* ends) {
  Map& m = *(Map*)h;
  if (!m.isect) return ORC_NO_INTERSECTOR;
  const Intersector& is = *m.isect;
  for (int64_t i = 0; i < n; ++i) {
    double d[3] = {dirs[3 * i], dirs[3 * i + 1], dirs[3 * i + 2]};
    // Vec3::normalize [U3] (eps 1e-7), then *= max length.
    const double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (!(std::fabs(0.0 - len) <= 1e-7)) {
      const double r = 1.0 / len;
      for (int a = 0; a < 3; ++a) d[a] *= r;
    }
    for (int a = 0; a < 3; ++a) d[a] *= lens[i];
    RayState ray;
    for (int a = 0; a < 3; ++a) {
      ray.eye[a] = origins[3 * i + a] * m.inv;  // continuous worldToIndex (VM:692)
      ray.dir[a] = d[a] * m.inv;
      ray.inv[a] = 1.0 / ray.dir[a];
    }
    ray.t0 = 0.0;
    ray.t1 = 1.0;
    clip_ray(is, ray);  // setIndexRay: the result is ignored (VM:698) [U6]
    Span t;
    if ((ray.t1 - ray.t0) > (double)1e-5f) hdda_4096(is, ray, t);  // Ray::valid [U10]
    double* out = ends + 3 * i;
    if (t.valid()) {
      double eye[3], dir[3], inv[3];
      for (int a = 0; a < 3; ++a) {
        eye[a] = origins[3 * i + a] * m.inv;
        dir[a] = d[a] * m.inv;
        inv[a] = 1.0 / dir[a];
      }
      Dda dd;
      dd.init(eye, dir, inv, t.t0, t.t1, 0);
      while (dd.step() && !map_is_on(m, dd.vox[0], dd.vox[1], dd.vox[2])) {
      }
      for (int a = 0; a < 3; ++a) out[a] = (double)dd.vox[a] * m.res;
      success[i] = 1;
    } else {
      for (int a = 0; a < 3; ++a) {
        const double oi = origins[3 * i + a] * m.inv, di = d[a] * m.inv;
        out[a] = (oi + di) * m.res;
      }
      success[i] = 0;
    }
  }
  return ORC_OK;
}

// ------------------------------------------------------------ map edits --

// addPointsToGrid / removePointsFromGrid (VM:413-447): each point's voxel
// Coord::floor(worldToIndex(p)) [U1; no half-voxel shift, unlike VM:612-631]
// gets setNodeToOccupied (OCC:124-129: max log-odds, active) or setNodeToFree
// (OCC:118-123: min log-odds, inactive) through modifyValueAndActiveState
// [U8]. NaN or out-of-range points reject the call before any change (the
// reference converts them to undefined coordinates).
int orc_set_points(void* h, const float* xyz, int64_t n, int64_t stride, int32_t occupied) {
  Map& m = *(Map*)h;
  if (n < 0 || stride < 12 || (stride & 3) || (n && !xyz)) return ORC_BAD_ARGUMENT;
  std::vector<int32_t> vox((size_t)n * 3);
  const char* base = (const char*)xyz;
  for (int64_t i = 0; i < n; ++i) {
    const float* pt = (const float*)(base + i * stride);
    int32_t* c = &vox[(size_t)i * 3];
    for (int a = 0; a < 3; ++a) {
      const double f = std::floor((double)pt[a] * m.inv);
      if (!(f >= -(double)kCoordLimit && f < (double)kCoordLimit)) return ORC_OUT_OF_RANGE;
      c[a] = (int32_t)f;
    }
  }
  const float val = occupied ? m.c_max : m.c_min;
  const bool need = occupied || val != 0.0f;  // does the op change a background voxel?
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* c = &vox[(size_t)i * 3];
    m.n4096.insert(key4096(c[0], c[1], c[2]));
    const uint64_t k = leaf_key(c[0], c[1], c[2]);
    if (!need && !m.leaves.count(k)) continue;
    Leaf& leaf = touch_leaf(m, k);
    const int off = leaf_offset(c[0], c[1], c[2]);
    leaf.v[off] = val;
    if (occupied)
      leaf.act[off >> 6] |= 1ull << (off & 63);
    else
      leaf.act[off >> 6] &= ~(1ull << (off & 63));
  }
  tile_check(m);
  return ORC_OK;
}

namespace {

// restoreMapIntegrity (VM:1152-1166): setNodeState (OCC:130-134: active =
// value > t_max) through modifyValueAndActiveState on every artificial voxel
// [U8], then the artificial grid is cleared.
void restore_integrity(Map& m) {
  const bool need = 0.0f > m.t_max;  // would a background voxel change?
  for (auto& kv : m.art.leaves) {
    int32_t o[3];
    key_origin(kv.first, o);
    m.n4096.insert(key4096(o[0], o[1], o[2]));
    if (!need && !m.leaves.count(kv.first)) continue;
    Leaf& leaf = touch_leaf(m, kv.first);
    for (int w = 0; w < 8; ++w) {
      uint64_t vis = kv.second.vis[w];
      while (vis) {
        const int b = __builtin_ctzll(vis);
        vis &= vis - 1;
        if (leaf.v[w * 64 + b] > m.t_max)
          leaf.act[w] |= 1ull << b;
        else
          leaf.act[w] &= ~(1ull << b);
      }
    }
  }
  tile_check(m);
  m.art.clear();
  m.art_present = false;
}

}  // namespace

int orc_restore_map_integrity(void* h) {
  restore_integrity(*(Map*)h);
  return ORC_OK;
}

namespace {

// (int)(height / m_resolution) of addArtificialWall (VM:1227-1228).
bool wall_layers(const Map& m, double neg, double pos, int32_t* k0, int32_t* k1) {
  const double qn = neg / m.res, qp = pos / m.res;
  if (!(qn > -2147483649.0 && qn < 2147483648.0) || !(qp > -2147483649.0 && qp < 2147483648.0)) return false;
  *k0 = (int32_t)qn;
  *k1 = (int32_t)qp;
  return true;
}

// Rounded wall ends (VM:1222-1225, :612-631), checked for every layer.
int wall_ends(const Map& m, const double* a, const double* b, int32_t k0, int32_t k1,
              std::array<int32_t, 6>& w) {
  world_to_index(m.res, m.inv, a, w.data());
  world_to_index(m.res, m.inv, b, w.data() + 3);
  for (int j = 0; j < 6; ++j) {
    const int64_t lo = (int64_t)w[j] + ((j % 3 == 2) ? k0 : 0);
    const int64_t hi = (int64_t)w[j] + ((j % 3 == 2) ? k1 - 1 : 0);
    if (w[j] == INT32_MIN || lo < -kCoordLimit || hi >= kCoordLimit) return ORC_OUT_OF_RANGE;
  }
  return ORC_OK;
}

// castRayIntoGrid of every layer of the walls into the artificial grid
// (VM:1230-1235).
void cast_walls(Map& m, const std::vector<std::array<int32_t, 6>>& walls, int32_t k0, int32_t k1) {
  for (const auto& w : walls) {
    for (int32_t k = k0; k < k1; ++k) {
      const int32_t s[3] = {w[0], w[1], w[2] + k}, e[3] = {w[3], w[4], w[5] + k};
      cast_ray(s, e, [&](int32_t x, int32_t y, int32_t z) {
        UpdLeaf& l = m.art.get(leaf_key(x, y, z));
        const int off = leaf_offset(x, y, z);
        l.vis[off >> 6] |= 1ull << (off & 63);
      });
    }
  }
}

}  // namespace

// addArtificialAreas (VM:1175-1190) -> addArtificialPolygon (VM:1198-1207)
// -> addArtificialWall (VM:1217-1236): restore first, then for every polygon
// edge (i, i+1 mod size) a castRayIntoGrid between the rounded edge ends
// (VM:612-631) shifted by (0, 0, k) for k in [(int)(neg / res), (int)(pos /
// res)), marking the artificial grid. Polygons are `counts[p]` consecutive
// (x, y, z) vertices of `xyz`. Arguments are validated before any change.
int orc_add_artificial_areas(void* h, int32_t n_polygons, const int32_t* counts, const double* xyz,
                             double negative_height, double positive_height) {
  Map& m = *(Map*)h;
  if (n_polygons < 0 || (n_polygons && !counts)) return ORC_BAD_ARGUMENT;
  int32_t k0, k1;
  if (!wall_layers(m, negative_height, positive_height, &k0, &k1)) return ORC_BAD_ARGUMENT;
  std::vector<std::array<int32_t, 6>> walls;
  int64_t v0 = 0;
  for (int32_t p = 0; p < n_polygons; ++p) {
    const int32_t cnt = counts[p];
    if (cnt < 0 || (cnt && !xyz)) return ORC_BAD_ARGUMENT;
    for (int32_t i = 0; i < cnt && k0 < k1; ++i) {
      std::array<int32_t, 6> w;
      const int st = wall_ends(m, xyz + 3 * (v0 + i), xyz + 3 * (v0 + (i + 1) % cnt), k0, k1, w);
      if (st) return st;
      walls.push_back(w);
    }
    v0 += cnt;
  }
  restore_integrity(m);
  m.art_present = true;
  cast_walls(m, walls, k0, k1);
  return ORC_OK;
}

// addArtificialWall (VM:1217-1236) for n walls (start xyz, end xyz), added to
// the artificial grid without a restore.
int orc_add_artificial_walls(void* h, int32_t n_walls, const double* xyz, double negative_height,
                             double positive_height) {
  Map& m = *(Map*)h;
  if (n_walls < 0 || (n_walls && !xyz)) return ORC_BAD_ARGUMENT;
  int32_t k0, k1;
  if (!wall_layers(m, negative_height, positive_height, &k0, &k1)) return ORC_BAD_ARGUMENT;
  std::vector<std::array<int32_t, 6>> walls;
  for (int32_t i = 0; i < n_walls && k0 < k1; ++i) {
    std::array<int32_t, 6> w;
    const int st = wall_ends(m, xyz + 6 * i, xyz + 6 * i + 3, k0, k1, w);
    if (st) return st;
    walls.push_back(w);
  }
  cast_walls(m, walls, k0, k1);
  return ORC_OK;
}

// The artificial grid (visited mask per leaf), sorted by origin.
int64_t orc_export_artificial(void* h, int32_t* origins, uint64_t* masks, int64_t cap) {
  Map& m = *(Map*)h;
  std::vector<uint64_t> keys, sorted;
  for (auto& kv : m.art.leaves) keys.push_back(kv.first);
  const int64_t n = export_sorted_keys(keys, sorted);
  if (!origins && !masks) return n;
  if (n > cap) return -n;
  for (int64_t i = 0; i < n; ++i) {
    key_origin(sorted[i], origins + 3 * i);
    std::memcpy(masks + 8 * i, m.art.leaves.at(sorted[i]).vis, 64);
  }
  return n;
}

// ----------------------------------------------------------- map sections --

namespace {

// Voxels of the leaf at origin o inside the inclusive index box b.
void inside_mask(const int32_t o[3], const int32_t b[6], uint64_t out[8]) {
  for (int w = 0; w < 8; ++w) out[w] = 0;
  for (int n = 0; n < 512; ++n) {
    const int32_t x = o[0] + (n >> 6), y = o[1] + ((n >> 3) & 7), z = o[2] + (n & 7);
    if (x >= b[0] && x <= b[3] && y >= b[1] && y <= b[4] && z >= b[2] && z <= b[5])
      out[n >> 6] |= 1ull << (n & 63);
  }
}

}  // namespace

// getMapSection (VM:922-969) with extractFullLeaf (VM:971-992) or
// extractSparseLeaf (VM:1000-1014) for every map leaf overlapping the box.
// The section grid gets a leaf when a voxel inside the box is set on, or set
// off with a value other than its background 0 (setValueOff on a tile keeps
// the tile when the value equals it); voxels of that leaf visited before its
// creation keep the background +0. Full: inside voxels keep value and state.
// Sparse: inside active voxels are set on with value true (1.0 in a float
// grid). Everything else in a section leaf is background (0, off). Sorted by
// origin; a null output returns the count; a short cap returns -count.
int64_t orc_export_section(void* h, const int32_t bbox[6], int32_t full, int32_t* origins,
                           float* values, uint64_t* masks, int64_t cap) {
  Map& m = *(Map*)h;
  std::vector<uint64_t> keys, sorted;
  std::vector<Leaf> outs;
  std::unordered_map<uint64_t, size_t> at;
  for (auto& kv : m.leaves) {
    if (kv.second.tile) continue;  // cbeginLeaf (VM:937) does not visit tiles
    int32_t o[3];
    key_origin(kv.first, o);
    uint64_t in[8];
    inside_mask(o, bbox, in);
    Leaf out;
    std::memset(&out, 0, sizeof out);
    bool created = false;
    for (int n = 0; n < 512; ++n) {
      if (!((in[n >> 6] >> (n & 63)) & 1)) continue;
      const bool on = (kv.second.act[n >> 6] >> (n & 63)) & 1;
      const float v = kv.second.v[n];
      if (full) {
        if (!created && !on && v == 0.0f) continue;  // setValueOff(xyz, 0) on the background tile
        created = true;
        out.v[n] = v;
        if (on) out.act[n >> 6] |= 1ull << (n & 63);
      } else if (on) {
        created = true;
        out.v[n] = 1.0f;
        out.act[n >> 6] |= 1ull << (n & 63);
      }
    }
    if (!created) continue;
    at[kv.first] = outs.size();
    outs.push_back(out);
    keys.push_back(kv.first);
  }
  const int64_t n = export_sorted_keys(keys, sorted);
  if (!origins && !values && !masks) return n;
  if (n > cap) return -n;
  for (int64_t i = 0; i < n; ++i) {
    const Leaf& l = outs[at[sorted[i]]];
    key_origin(sorted[i], origins + 3 * i);
    std::memcpy(values + 512 * i, l.v, sizeof l.v);
    std::memcpy(masks + 8 * i, l.act, sizeof l.act);
  }
  return n;
}

// applyMapSectionGrid without smoothing (VM:1022-1050): every value of the
// section grid, cbeginValueAll, is written to the map with its state
// (setValueOn / setValueOff). For a section leaf that is all 512 voxels; a
// missing map leaf is created at the first voxel that is on or non-zero (the
// voxels before it stay background). The iteration also visits the inactive
// tiles of the section's internal nodes [U12]: every 8^3 slot of a 128^3
// node without a section leaf, and every 128^3 slot of a 4096^3 node without
// a 128^3 child, yields setValueOff(slot origin, 0), which sets that voxel of
// an existing map leaf to (0, off) and changes nothing where the map has no
// leaf. `origins` are the section's leaf origins (multiples of 8, distinct).
int orc_apply_section(void* h, int64_t n, const int32_t* origins, const float* values,
                      const uint64_t* masks) {
  Map& m = *(Map*)h;
  if (n < 0 || (n && (!origins || !values || !masks))) return ORC_BAD_ARGUMENT;
  std::unordered_set<uint64_t> sleaf, s128, s4096;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    if ((o[0] | o[1] | o[2]) & 7) return ORC_BAD_ARGUMENT;
    if (!in_range(o)) return ORC_OUT_OF_RANGE;
    if (!sleaf.insert(leaf_key(o[0], o[1], o[2])).second) return ORC_BAD_ARGUMENT;
  }
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    const uint64_t k = leaf_key(o[0], o[1], o[2]);
    s128.insert(key128(o[0], o[1], o[2]));
    s4096.insert(key4096(o[0], o[1], o[2]));
    const float* v = values + 512 * i;
    const uint64_t* a = masks + 8 * i;
    int first = 0;
    if (!m.leaves.count(k)) {
      first = 512;
      for (int j = 0; j < 512; ++j)
        if (((a[j >> 6] >> (j & 63)) & 1) || v[j] != 0.0f) {
          first = j;
          break;
        }
      if (first == 512) continue;
      m.n4096.insert(key4096(o[0], o[1], o[2]));
    }
    Leaf& leaf = touch_leaf(m, k);
    for (int j = first; j < 512; ++j) {
      leaf.v[j] = v[j];
      const uint64_t bit = 1ull << (j & 63);
      leaf.act[j >> 6] = (leaf.act[j >> 6] & ~bit) | (a[j >> 6] & bit);
    }
  }
  auto clobber = [&](int32_t x, int32_t y, int32_t z) {
    auto it = m.leaves.find(leaf_key(x, y, z));
    if (it == m.leaves.end()) return;
    it->second.v[leaf_offset(x, y, z)] = 0.0f;
    const int off = leaf_offset(x, y, z);
    it->second.act[off >> 6] &= ~(1ull << (off & 63));
  };
  for (uint64_t k : s128) {
    int32_t o[3];
    key_origin(k, o);  // key128 packs origin / 128 * 8
    for (int a = 0; a < 3; ++a) o[a] = o[a] / 8 * 128;
    for (int j = 0; j < 4096; ++j) {
      const int32_t x = o[0] + 8 * (j >> 8), y = o[1] + 8 * ((j >> 4) & 15), z = o[2] + 8 * (j & 15);
      if (!sleaf.count(leaf_key(x, y, z))) clobber(x, y, z);
    }
  }
  for (uint64_t k : s4096) {
    int32_t o[3];
    key_origin(k, o);
    for (int a = 0; a < 3; ++a) o[a] = o[a] / 8 * 4096;
    for (int j = 0; j < 32768; ++j) {
      const int32_t x = o[0] + 128 * (j >> 10), y = o[1] + 128 * ((j >> 5) & 31),
                    z = o[2] + 128 * (j & 31);
      if (!s128.count(key128(x, y, z))) clobber(x, y, z);
    }
  }
  tile_check(m);
  return ORC_OK;
}

// applyMapSectionUpdateGrid without smoothing (VM:1058-1085): every active
// map voxel inside the inclusive box (the section's bb_min / bb_max) is set
// inactive, then every active section voxel is set active (a missing map leaf
// is created with value 0).
int orc_apply_section_update(void* h, const int32_t bbox[6], int64_t n, const int32_t* origins,
                             const uint64_t* masks) {
  Map& m = *(Map*)h;
  if (!bbox || n < 0 || (n && (!origins || !masks))) return ORC_BAD_ARGUMENT;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    if ((o[0] | o[1] | o[2]) & 7) return ORC_BAD_ARGUMENT;
    if (!in_range(o)) return ORC_OUT_OF_RANGE;
  }
  for (auto& kv : m.leaves) {
    int32_t o[3];
    key_origin(kv.first, o);
    uint64_t in[8];
    inside_mask(o, bbox, in);
    if (kv.second.tile) {
      // cbeginValueOn visits an active tile once, at its origin: only that voxel
      // is switched off, which turns the tile into a leaf [U15]
      if (kv.second.ts && (in[0] & 1)) {
        kv.second.act[0] &= ~1ull;
        kv.second.tile = false;
        m.tiles.erase(kv.first);
      }
      continue;
    }
    for (int w = 0; w < 8; ++w) kv.second.act[w] &= ~in[w];
  }
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    const uint64_t* a = masks + 8 * i;
    uint64_t any = 0;
    for (int w = 0; w < 8; ++w) any |= a[w];
    if (!any) continue;
    m.n4096.insert(key4096(o[0], o[1], o[2]));
    Leaf& leaf = touch_leaf(m, leaf_key(o[0], o[1], o[2]));
    for (int w = 0; w < 8; ++w) leaf.act[w] |= a[w];
  }
  tile_check(m);
  return ORC_OK;
}

// createIndexBoundingBox (VM:858-874) over createWorldBoundingBox (VM:811-850):
// the 8 corners of [min, max] as float, transformed by the row-major 4x4
// matrix in double and stored as float [U13: pcl::transformPointCloud], their
// float min / max, then Coord::floor(worldToIndex) [U1] of each.
void orc_index_bbox(double res, const double mn[3], const double mx[3], const double tf[16],
                    int32_t out[6]) {
  float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
  for (int c = 0; c < 8; ++c) {
    const double p[3] = {(double)(float)((c & 4) ? mx[0] : mn[0]), (double)(float)((c & 2) ? mx[1] : mn[1]),
                         (double)(float)((c & 1) ? mx[2] : mn[2])};
    for (int r = 0; r < 3; ++r) {
      const float q = (float)(tf[4 * r] * p[0] + tf[4 * r + 1] * p[1] + tf[4 * r + 2] * p[2] + tf[4 * r + 3]);
      lo[r] = std::min(lo[r], q);
      hi[r] = std::max(hi[r], q);
    }
  }
  const double inv = 1.0 / res;
  for (int r = 0; r < 3; ++r) {
    out[r] = (int32_t)std::floor((double)lo[r] * inv);
    out[3 + r] = (int32_t)std::floor((double)hi[r] * inv);
  }
}

// ------------------------------------------------------------ persistence --

namespace {

// Position of a leaf in OpenVDB's tree iteration order: root children by
// origin (x, then y, then z), then the 128^3 nodes and the leaves inside
// each by offset (x-major).
std::array<int32_t, 9> tree_order(uint64_t k) {
  int32_t o[3];
  key_origin(k, o);
  return {o[0] >> 12, o[1] >> 12, o[2] >> 12, o[0] >> 7, o[1] >> 7, o[2] >> 7, o[0] >> 3, o[1] >> 3, o[2] >> 3};
}

}  // namespace

// The point list of saveMapToPCD (VM:218-256): every active voxel in
// cbeginValueOn order (tree order, then voxel offset within the leaf), at
// indexToWorld(ijk) + 0.5 res (ScaleMap::applyMap multiplies by res [U1]),
// rounded to float. A null output returns the count; a short cap -count.
int64_t orc_export_active_points(void* h, float* xyz, int64_t cap) {
  Map& m = *(Map*)h;
  std::vector<std::pair<std::array<int32_t, 9>, uint64_t>> order;
  int64_t n = 0;
  for (auto& kv : m.leaves) {
    order.emplace_back(tree_order(kv.first), kv.first);
    if (kv.second.tile)
      n += kv.second.ts ? 1 : 0;  // cbeginValueOn yields an active tile once [U15]
    else
      for (int w = 0; w < 8; ++w) n += __builtin_popcountll(kv.second.act[w]);
  }
  if (!xyz) return n;
  if (n > cap) return -n;
  std::sort(order.begin(), order.end());
  int64_t i = 0;
  for (auto& ok : order) {
    const Leaf& l = m.leaves.at(ok.second);
    int32_t o[3];
    key_origin(ok.second, o);
    for (int v = 0; v < (l.tile ? 1 : 512); ++v) {
      if (!((l.act[v >> 6] >> (v & 63)) & 1)) continue;
      const int32_t c[3] = {o[0] + (v >> 6), o[1] + ((v >> 3) & 7), o[2] + (v & 7)};
      for (int a = 0; a < 3; ++a) {
        double w = (double)c[a] * m.res;
        w += 0.5 * m.res;
        xyz[3 * i + a] = (float)w;
      }
      ++i;
    }
  }
  return n;
}

// createMapFromPointCloud (OCC:136-174): clear the grid (not the update grids)
// if asked, setValueOn(floor(worldToIndex(p)), max log-odds) per point [U1],
// then with set_background every voxel of the active-voxel bounding box that is
// not active gets setValueOff(min log-odds) (a missing leaf is created when that
// value differs from the background 0), then pruneGrid at the leaf level [U15].
// NaN or out-of-range points reject the call before any change.
int orc_create_from_points(void* h, const float* xyz, int64_t n, int64_t stride, int32_t set_background,
                           int32_t clear_map) {
  Map& m = *(Map*)h;
  if (n < 0 || stride < 12 || (stride & 3) || (n && !xyz)) return ORC_BAD_ARGUMENT;
  const char* base = (const char*)xyz;
  for (int64_t i = 0; i < n; ++i) {
    const float* pt = (const float*)(base + i * stride);
    for (int a = 0; a < 3; ++a) {
      const double f = std::floor((double)pt[a] * m.inv);
      if (!(f >= -(double)kCoordLimit && f < (double)kCoordLimit)) return ORC_OUT_OF_RANGE;
    }
  }
  if (clear_map) {
    m.leaves.clear();
    m.tiles.clear();
    m.n4096.clear();
  }
  for (int64_t i = 0; i < n; ++i) {
    const float* pt = (const float*)(base + i * stride);
    int32_t c[3];
    for (int a = 0; a < 3; ++a) c[a] = (int32_t)std::floor((double)pt[a] * m.inv);
    m.n4096.insert(key4096(c[0], c[1], c[2]));
    Leaf& leaf = touch_leaf(m, leaf_key(c[0], c[1], c[2]));
    const int off = leaf_offset(c[0], c[1], c[2]);
    leaf.v[off] = m.c_max;
    leaf.act[off >> 6] |= 1ull << (off & 63);
  }
  if (!set_background) {
    prune_leaves(m);
    return ORC_OK;
  }
  int32_t lo[3] = {INT32_MAX, INT32_MAX, INT32_MAX}, hi[3] = {INT32_MIN, INT32_MIN, INT32_MIN};
  for (auto& kv : m.leaves) {
    int32_t o[3];
    key_origin(kv.first, o);
    for (int v = 0; v < 512; ++v) {
      if (!((kv.second.act[v >> 6] >> (v & 63)) & 1)) continue;
      const int32_t c[3] = {o[0] + (v >> 6), o[1] + ((v >> 3) & 7), o[2] + (v & 7)};
      for (int a = 0; a < 3; ++a) {
        lo[a] = std::min(lo[a], c[a]);
        hi[a] = std::max(hi[a], c[a]);
      }
    }
  }
  if (lo[0] > hi[0]) {  // no active voxel: an empty box
    prune_leaves(m);
    return ORC_OK;
  }
  for (int32_t x = lo[0]; x <= hi[0]; ++x)
    for (int32_t y = lo[1]; y <= hi[1]; ++y)
      for (int32_t z = lo[2]; z <= hi[2]; ++z) {
        const uint64_t k = leaf_key(x, y, z);
        const int off = leaf_offset(x, y, z);
        auto it = m.leaves.find(k);
        if (it != m.leaves.end()) {
          if ((it->second.act[off >> 6] >> (off & 63)) & 1) continue;
          it->second.v[off] = m.c_min;
        } else if (m.c_min != 0.0f) {
          m.n4096.insert(key4096(x, y, z));
          touch_leaf(m, k).v[off] = m.c_min;
        }
      }
  prune_leaves(m);
  return ORC_OK;
}

}  // extern "C"

// Morphology of a leaf set (VM:1095-1148: morphologicalDilateMap and
// morphologicalErodeMap, composed by morphologicalCloseMap / OpenMap), voxel by
// voxel as the textbook operators with a 3x3x3 structuring element [U14].
namespace {

struct MorphSet {
  std::unordered_map<uint64_t, std::array<uint64_t, 8>> leaves;
  std::vector<uint64_t> created;  // leaves dilation added, in creation order
};

bool morph_on(const MorphSet& s, int32_t x, int32_t y, int32_t z) {
  auto it = s.leaves.find(leaf_key(x, y, z));
  if (it == s.leaves.end()) return false;
  const int off = leaf_offset(x, y, z);
  return (it->second[off >> 6] >> (off & 63)) & 1;
}

template <class F>
void morph_for_each_on(const MorphSet& s, F&& f) {
  for (const auto& kv : s.leaves) {
    int32_t o[3];
    key_origin(kv.first, o);
    for (int v = 0; v < 512; ++v)
      if ((kv.second[v >> 6] >> (v & 63)) & 1) f(o[0] + (v >> 6), o[1] + ((v >> 3) & 7), o[2] + (v & 7));
  }
}

// One step of dilateActiveValues: every voxel with an active 26-neighbour (or
// itself active) becomes active; a missing leaf is created on first activation.
void morph_dilate1(MorphSet& s) {
  MorphSet next = s;
  morph_for_each_on(s, [&](int32_t x, int32_t y, int32_t z) {
    for (int dx = -1; dx <= 1; ++dx)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dz = -1; dz <= 1; ++dz) {
          const int32_t c[3] = {x + dx, y + dy, z + dz};
          const uint64_t k = leaf_key(c[0], c[1], c[2]);
          auto it = next.leaves.find(k);
          if (it == next.leaves.end()) {
            it = next.leaves.emplace(k, std::array<uint64_t, 8>{}).first;
            next.created.push_back(k);
          }
          const int off = leaf_offset(c[0], c[1], c[2]);
          it->second[off >> 6] |= 1ull << (off & 63);
        }
  });
  s = std::move(next);
}

// One step of erodeActiveValues: an active voxel stays active only if all 26
// neighbours are active (a voxel in no leaf counts as inactive).
void morph_erode1(MorphSet& s) {
  MorphSet next = s;
  morph_for_each_on(s, [&](int32_t x, int32_t y, int32_t z) {
    bool all = true;
    for (int dx = -1; dx <= 1 && all; ++dx)
      for (int dy = -1; dy <= 1 && all; ++dy)
        for (int dz = -1; dz <= 1 && all; ++dz) all = morph_on(s, x + dx, y + dy, z + dz);
    if (!all) {
      const int off = leaf_offset(x, y, z);
      next.leaves[leaf_key(x, y, z)][off >> 6] &= ~(1ull << (off & 63));
    }
  });
  s = std::move(next);
}

}  // namespace

extern "C" {

int64_t orc_morphology(int32_t op, int32_t iterations, int64_t n, const int32_t* origins,
                       const uint64_t* masks, int32_t* out_origins, uint64_t* out_masks, int64_t cap) {
  if (op < 0 || op > 3 || n < 0 || (n && (!origins || !masks))) return -1;
  MorphSet s;
  std::vector<uint64_t> input((size_t)n);
  // dilation must not carry a leaf out of the key range: 8 voxels per 8 steps
  const int64_t margin = op == 1 || iterations <= 0 ? 0 : 8 * (((int64_t)iterations + 7) / 8);
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* o = origins + 3 * i;
    if ((o[0] | o[1] | o[2]) & 7) return -1;
    for (int a = 0; a < 3; ++a)
      if (o[a] - margin < -kCoordLimit || o[a] + 8 + margin > kCoordLimit) return -1;
    input[(size_t)i] = leaf_key(o[0], o[1], o[2]);
    std::array<uint64_t, 8> w;
    std::memcpy(w.data(), masks + 8 * i, 64);
    if (!s.leaves.emplace(input[(size_t)i], w).second) return -1;  // repeated origin
  }
  const bool dilate_first = op == 0 || op == 2;
  const int phases = op >= 2 ? 2 : 1;
  for (int p = 0; p < phases; ++p) {
    const bool dilate = (p == 0) == dilate_first;
    for (int32_t it = 0; it < iterations; ++it) dilate ? morph_dilate1(s) : morph_erode1(s);
  }
  std::vector<std::array<int32_t, 3>> fresh;
  for (uint64_t k : s.created) {
    std::array<int32_t, 3> o;
    key_origin(k, o.data());
    fresh.push_back(o);
  }
  std::sort(fresh.begin(), fresh.end());
  const int64_t total = n + (int64_t)fresh.size();
  if (!out_origins && !out_masks) return total;
  if (total > cap) return -total;
  for (int64_t i = 0; i < total; ++i) {
    int32_t o[3];
    if (i < n) {
      key_origin(input[(size_t)i], o);
    } else {
      std::memcpy(o, fresh[(size_t)(i - n)].data(), 12);
    }
    if (out_origins) std::memcpy(out_origins + 3 * i, o, 12);
    if (out_masks) std::memcpy(out_masks + 8 * i, s.leaves[leaf_key(o[0], o[1], o[2])].data(), 64);
  }
  return total;
}

}  // extern "C"
