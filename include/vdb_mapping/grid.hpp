// Host-side grids of the drop-in API.
//
// FloatGrid is what VDBMapping::getGrid() returns (GridT): a live view of the
// device map in openvdb::FloatTree leaf layout (8^3 float leaves, offset
// ((x&7)<<6)|((y&7)<<3)|(z&7), value mask bit n&63 of word n>>6). It keeps a
// host copy of the leaves and, when the map has integrated since the last
// read, pulls only the changed leaves (vdbgpu_export_dirty_leaves). An
// accessor taken before an insert therefore sees the new values, as the
// reference's shared grid pointer does (tests/mapping.cpp:13 -> :28-29).
// The leaf-level tiles of createMapFromPointCloud's pruneGrid are kept beside
// the leaves (value and state per 8^3 tile); a changed tile comes back as a
// dirty leaf and leaves the tile list.
//
// BoolGrid is UpdateGridT: per leaf, visited (active) and hit (value true)
// masks, the content of the reference's Tree4<bool,1,4,3> update grid.
//
// With OpenVDB installed, FloatGrid::toOpenVDB() copies the map into a new
// openvdb::FloatGrid (transform, GRID_LEVEL_SET class, leaf values and masks).
#ifndef VDB_MAPPING_GRID_HPP
#define VDB_MAPPING_GRID_HPP

#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../vdbgpu.h"
#include "vdb_mapping/compat.hpp"

namespace vdb_mapping {

inline uint64_t host_leaf_key(int32_t x, int32_t y, int32_t z)
{
  const uint64_t m = (1ull << 21) - 1;
  return ((uint64_t)(uint32_t)(x >> 3) & m) << 42 | ((uint64_t)(uint32_t)(y >> 3) & m) << 21 |
         ((uint64_t)(uint32_t)(z >> 3) & m);
}

inline int host_leaf_offset(int32_t x, int32_t y, int32_t z)
{
  return ((x & 7) << 6) | ((y & 7) << 3) | (z & 7);
}

struct HostLeaf
{
  int32_t origin[3];
  float values[512];
  uint64_t mask[8];
};

// An 8^3 tile of a pruned grid: one value and one state for all its voxels.
struct HostTile
{
  int32_t origin[3];
  float value;
  bool active;
};

class FloatGrid
{
public:
  using Ptr       = std::shared_ptr<FloatGrid>;
  using ConstPtr  = std::shared_ptr<const FloatGrid>;
  using ValueType = float;

  explicit FloatGrid(double voxel_size)
    : m_voxel_size(voxel_size)
  {
  }

  class Accessor
  {
  public:
    explicit Accessor(const FloatGrid* g = nullptr)
      : m_grid(g)
    {
    }
    float getValue(const openvdb::Coord& c) const
    {
      const HostTile* t = nullptr;
      const HostLeaf* l = m_grid->leaf(c, &t);
      if (t)
      {
        return t->value;
      }
      return l ? l->values[host_leaf_offset(c.x(), c.y(), c.z())] : 0.0f;
    }
    bool isValueOn(const openvdb::Coord& c) const
    {
      const HostTile* t = nullptr;
      const HostLeaf* l = m_grid->leaf(c, &t);
      if (t)
      {
        return t->active;
      }
      if (!l)
      {
        return false;
      }
      const int n = host_leaf_offset(c.x(), c.y(), c.z());
      return (l->mask[n >> 6] >> (n & 63)) & 1;
    }

  private:
    const FloatGrid* m_grid;
  };

  Accessor getAccessor() const { return Accessor(this); }
  Accessor getConstAccessor() const { return Accessor(this); }

  double voxelSize() const { return m_voxel_size; }

  bool empty() const
  {
    sync();
    return m_leaves.empty() && m_tiles.empty();
  }

  uint64_t leafCount() const
  {
    sync();
    return m_leaves.size();
  }

  openvdb::Index64 activeVoxelCount() const
  {
    sync();
    openvdb::Index64 n = 0;
    for (const auto& kv : m_leaves)
    {
      for (int w = 0; w < 8; ++w)
      {
        n += (openvdb::Index64)__builtin_popcountll(kv.second.mask[w]);
      }
    }
    for (const auto& kv : m_tiles)
    {
      n += kv.second.active ? 512 : 0;  // a tile counts all its voxels
    }
    return n;
  }

  void clear()
  {
    std::lock_guard<std::mutex> lk(m_mu);
    m_leaves.clear();
    m_tiles.clear();
    m_stale   = false;
    m_rebuild = false;
  }

  // f(const HostTile&) for every leaf-level tile (unordered).
  template <typename F>
  void forEachTile(F&& f) const
  {
    sync();
    for (const auto& kv : m_tiles)
    {
      f(kv.second);
    }
  }
  uint64_t tileCount() const
  {
    sync();
    return m_tiles.size();
  }

  // f(const HostLeaf&) for every leaf (unordered).
  template <typename F>
  void forEachLeaf(F&& f) const
  {
    sync();
    for (const auto& kv : m_leaves)
    {
      f(kv.second);
    }
  }

#ifdef VDBGPU_HAVE_OPENVDB
  openvdb::FloatGrid::Ptr toOpenVDB() const
  {
    sync();
    openvdb::FloatGrid::Ptr g = openvdb::FloatGrid::create(0.0f);
    g->setTransform(openvdb::math::Transform::createLinearTransform(m_voxel_size));
    g->setGridClass(openvdb::GRID_LEVEL_SET);
    auto& tree = g->tree();
    for (const auto& kv : m_leaves)
    {
      const HostLeaf& l = kv.second;
      auto* leaf = tree.touchLeaf(openvdb::Coord(l.origin[0], l.origin[1], l.origin[2]));
      for (int n = 0; n < 512; ++n)
      {
        leaf->setValueOnly(n, l.values[n]);
        leaf->setActiveState(n, (l.mask[n >> 6] >> (n & 63)) & 1);
      }
    }
    for (const auto& kv : m_tiles)
    {
      const HostTile& t = kv.second;
      tree.addTile(1, openvdb::Coord(t.origin[0], t.origin[1], t.origin[2]), t.value, t.active);
    }
    return g;
  }
#endif

  // Detached grids (map sections): leaves set directly, Vec3d metadata.
  static Ptr create(double voxel_size) { return std::make_shared<FloatGrid>(voxel_size); }
  void insertLeaf(const HostLeaf& leaf)
  {
    std::lock_guard<std::mutex> lk(m_mu);
    m_leaves[host_leaf_key(leaf.origin[0], leaf.origin[1], leaf.origin[2])] = leaf;
  }
  void insertMeta(const std::string& name, const openvdb::Vec3d& value) { m_meta[name] = value; }
  template <typename T>
  T metaValue(const std::string& name) const
  {
    static_assert(std::is_same<T, openvdb::Vec3d>::value, "grid metadata holds Vec3d values");
    auto it = m_meta.find(name);
    if (it == m_meta.end())
    {
      throw std::runtime_error("no grid metadata named " + name);
    }
    return it->second;
  }
  bool hasMeta(const std::string& name) const { return m_meta.count(name) != 0; }
  // Active mask of the leaf at origin; a missing leaf is created with the
  // background value 0 in every voxel (detached grids only).
  void setLeafMask(const int32_t origin[3], const uint64_t* mask)
  {
    std::lock_guard<std::mutex> lk(m_mu);
    auto it = m_leaves.find(host_leaf_key(origin[0], origin[1], origin[2]));
    if (it == m_leaves.end())
    {
      HostLeaf l;
      std::memset(&l, 0, sizeof l);
      std::memcpy(l.origin, origin, sizeof l.origin);
      it = m_leaves.emplace(host_leaf_key(origin[0], origin[1], origin[2]), l).first;
    }
    std::memcpy(it->second.mask, mask, sizeof it->second.mask);
  }

  // Live binding to a device map (used by VDBMapping).
  void attach(vdbgpu_map* map)
  {
    std::lock_guard<std::mutex> lk(m_mu);
    m_map   = map;
    m_stale = true;
  }
  void detach()
  {
    std::lock_guard<std::mutex> lk(m_mu);
    m_map = nullptr;
  }
  void markStale()
  {
    std::lock_guard<std::mutex> lk(m_mu);
    m_stale = true;
  }
  // The next read pulls every leaf and tile (after a prune turned leaves into tiles).
  void markRebuild()
  {
    std::lock_guard<std::mutex> lk(m_mu);
    m_stale   = true;
    m_rebuild = true;
  }
  bool attached() const
  {
    std::lock_guard<std::mutex> lk(m_mu);
    return m_map != nullptr;
  }

private:
  const HostLeaf* leaf(const openvdb::Coord& c, const HostTile** tile = nullptr) const
  {
    sync();
    const uint64_t k = host_leaf_key(c.x(), c.y(), c.z());
    if (tile)
    {
      auto t = m_tiles.find(k);
      *tile  = t == m_tiles.end() ? nullptr : &t->second;
    }
    auto it = m_leaves.find(k);
    return it == m_leaves.end() ? nullptr : &it->second;
  }

  void sync() const
  {
    std::lock_guard<std::mutex> lk(m_mu);
    if (!m_stale || !m_map)
    {
      return;
    }
    const int64_t cap = vdbgpu_leaf_count(m_map);
    if (m_rebuild)
    {
      m_leaves.clear();
      m_tiles.clear();
      const int64_t nt = vdbgpu_export_tiles(m_map, nullptr, nullptr, nullptr, 0);
      if (nt > 0)
      {
        std::vector<int32_t> o(3 * nt);
        std::vector<float> v(nt);
        std::vector<uint8_t> a(nt);
        const int64_t got = vdbgpu_export_tiles(m_map, o.data(), v.data(), a.data(), nt);
        for (int64_t i = 0; i < got; ++i)
        {
          HostTile& t = m_tiles[host_leaf_key(o[3 * i], o[3 * i + 1], o[3 * i + 2])];
          std::memcpy(t.origin, &o[3 * i], sizeof t.origin);
          t.value  = v[i];
          t.active = a[i] != 0;
        }
      }
    }
    if (cap > 0)
    {
      std::vector<int32_t> o(3 * cap);
      std::vector<float> v(512 * cap);
      std::vector<uint64_t> msk(8 * cap);
      const int64_t n = m_rebuild ? vdbgpu_export_leaves(m_map, o.data(), v.data(), msk.data(), cap)
                                  : vdbgpu_export_dirty_leaves(m_map, o.data(), v.data(), msk.data(), cap);
      for (int64_t i = 0; i < n; ++i)
      {
        const uint64_t k = host_leaf_key(o[3 * i], o[3 * i + 1], o[3 * i + 2]);
        HostLeaf& l      = m_leaves[k];
        std::memcpy(l.origin, &o[3 * i], sizeof l.origin);
        std::memcpy(l.values, &v[512 * i], sizeof l.values);
        std::memcpy(l.mask, &msk[8 * i], sizeof l.mask);
        m_tiles.erase(k);  // a changed tile is a leaf again
      }
    }
    m_stale   = false;
    m_rebuild = false;
  }

  double m_voxel_size;
  std::map<std::string, openvdb::Vec3d> m_meta;
  vdbgpu_map* m_map = nullptr;
  mutable std::mutex m_mu;
  mutable bool m_stale   = false;
  mutable bool m_rebuild = false;
  mutable std::unordered_map<uint64_t, HostLeaf> m_leaves;
  mutable std::unordered_map<uint64_t, HostTile> m_tiles;
};

class BoolGrid
{
public:
  using Ptr      = std::shared_ptr<BoolGrid>;
  using ConstPtr = std::shared_ptr<const BoolGrid>;

  struct Leaf
  {
    int32_t origin[3];
    uint64_t active[8];
    uint64_t value[8];
  };

  static Ptr create(bool = false) { return std::make_shared<BoolGrid>(); }

  class Accessor
  {
  public:
    explicit Accessor(BoolGrid* g = nullptr)
      : m_grid(g)
    {
    }
    void setActiveState(const openvdb::Coord& c, bool on)
    {
      Leaf& l           = m_grid->touch(c);
      const int n       = host_leaf_offset(c.x(), c.y(), c.z());
      const uint64_t b  = 1ull << (n & 63);
      l.active[n >> 6]  = on ? (l.active[n >> 6] | b) : (l.active[n >> 6] & ~b);
    }
    void setValueOn(const openvdb::Coord& c, bool value)
    {
      Leaf& l          = m_grid->touch(c);
      const int n      = host_leaf_offset(c.x(), c.y(), c.z());
      const uint64_t b = 1ull << (n & 63);
      l.active[n >> 6] |= b;
      l.value[n >> 6] = value ? (l.value[n >> 6] | b) : (l.value[n >> 6] & ~b);
    }
    bool isValueOn(const openvdb::Coord& c) const { return bit(c, true); }
    bool getValue(const openvdb::Coord& c) const { return bit(c, false); }
    // OR one leaf of visited / hit masks into the grid.
    void mergeLeaf(const int32_t origin[3], const uint64_t* visited, const uint64_t* hit)
    {
      Leaf& l = m_grid->touch(openvdb::Coord(origin[0], origin[1], origin[2]));
      for (int w = 0; w < 8; ++w)
      {
        l.active[w] |= visited[w];
        l.value[w] |= hit[w];
      }
    }

  private:
    bool bit(const openvdb::Coord& c, bool active) const
    {
      auto it = m_grid->m_leaves.find(host_leaf_key(c.x(), c.y(), c.z()));
      if (it == m_grid->m_leaves.end())
      {
        return false;
      }
      const int n = host_leaf_offset(c.x(), c.y(), c.z());
      const uint64_t w = active ? it->second.active[n >> 6] : it->second.value[n >> 6];
      return (w >> (n & 63)) & 1;
    }
    BoolGrid* m_grid;
  };

  Accessor getAccessor() { return Accessor(this); }
  bool empty() const { return m_leaves.empty(); }
  // the grid's linear transform (voxel size); 1 unless set
  double voxelSize() const { return m_voxel_size; }
  void setVoxelSize(double voxel_size) { m_voxel_size = voxel_size; }
  bool hasMeta(const std::string& name) const { return m_meta.count(name) != 0; }
  void insertMeta(const std::string& name, const openvdb::Vec3d& value) { m_meta[name] = value; }
  template <typename T>
  T metaValue(const std::string& name) const
  {
    static_assert(std::is_same<T, openvdb::Vec3d>::value, "grid metadata holds Vec3d values");
    auto it = m_meta.find(name);
    if (it == m_meta.end())
    {
      throw std::runtime_error("no grid metadata named " + name);
    }
    return it->second;
  }
  openvdb::Index64 activeVoxelCount() const
  {
    openvdb::Index64 n = 0;
    for (const auto& kv : m_leaves)
    {
      for (int w = 0; w < 8; ++w)
      {
        n += (openvdb::Index64)__builtin_popcountll(kv.second.active[w]);
      }
    }
    return n;
  }
  template <typename F>
  void forEachLeaf(F&& f) const
  {
    for (const auto& kv : m_leaves)
    {
      f(kv.second);
    }
  }
  // Active mask of the leaf at origin; a missing leaf is created with every
  // value false (the background).
  void setLeafMask(const int32_t origin[3], const uint64_t* active)
  {
    Leaf& l = touch(openvdb::Coord(origin[0], origin[1], origin[2]));
    std::memcpy(l.active, active, sizeof l.active);
  }
  bool attached() const { return false; }

private:
  Leaf& touch(const openvdb::Coord& c)
  {
    const uint64_t k = host_leaf_key(c.x(), c.y(), c.z());
    auto it          = m_leaves.find(k);
    if (it == m_leaves.end())
    {
      Leaf l;
      std::memset(&l, 0, sizeof l);
      l.origin[0] = c.x() & ~7;
      l.origin[1] = c.y() & ~7;
      l.origin[2] = c.z() & ~7;
      it          = m_leaves.emplace(k, l).first;
    }
    return it->second;
  }
  std::unordered_map<uint64_t, Leaf> m_leaves;
  std::map<std::string, openvdb::Vec3d> m_meta;
  double m_voxel_size = 1.0;
};

// Active-mask words of a leaf of either grid type.
inline const uint64_t* leafActiveWords(const HostLeaf& l) { return l.mask; }
inline const uint64_t* leafActiveWords(const BoolGrid::Leaf& l) { return l.active; }

} // namespace vdb_mapping

#endif // VDB_MAPPING_GRID_HPP
