// The OpenVDB archive layout written by io::File (saveMap, VDBMapping.hpp:193-211)
// and io::Stream (gridToByteArray, :1310-1317), for the two grid types of the
// mapping: the map (Tree_float_5_4_3, class "level set") and update grids
// (Tree_bool_5_4_3). Host code only.
//
// Layout, file format version 224: header (magic, format and library versions,
// grid-offset flag, 36-character UUID), file metadata, grid count; per grid its
// descriptor (name, type, instance parent, three stream offsets), compression
// flags, grid metadata, transform (UniformScaleMap), topology (root tiles and
// children, 32^3 and 16^3 internal nodes with child and value masks and their
// tile values, leaf value masks) and buffers (leaf values in tree order).
// Writing uses active-mask compression (a level-set map takes no zlib), zlib
// on request, never blosc; the 8^3 tiles of a pruned float grid go into the
// value tables of the 16^3 nodes. Reading accepts none / active-mask / zlib
// compression, voxelises tiles that are active or hold a value other than the
// background (a root-level tile of that kind is refused), and refuses blosc and
// half-float grids.
//
// This restates OpenVDB's published layout; it has not been checked against
// OpenVDB itself (not available to this build): files written here read back
// here, byte for byte the same grid.
#ifndef VDB_MAPPING_VDB_IO_HPP
#define VDB_MAPPING_VDB_IO_HPP

#include <zlib.h>

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <istream>
#include <map>
#include <memory>
#include <ostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

namespace vdb_mapping {
namespace vdbio {

constexpr int64_t kMagic             = 0x56444220;
constexpr uint32_t kFileVersion      = 224;
constexpr uint32_t kCompressZip      = 0x1;
constexpr uint32_t kCompressMask     = 0x2;
constexpr uint32_t kCompressBlosc    = 0x4;
constexpr char kFloatTree[]          = "Tree_float_5_4_3";
constexpr char kBoolTree[]           = "Tree_bool_5_4_3";
enum : int8_t
{
  NO_MASK_OR_INACTIVE_VALS     = 0,
  NO_MASK_AND_MINUS_BG         = 1,
  NO_MASK_AND_ONE_INACTIVE_VAL = 2,
  MASK_AND_NO_INACTIVE_VALS    = 3,
  MASK_AND_ONE_INACTIVE_VAL    = 4,
  MASK_AND_TWO_INACTIVE_VALS   = 5,
  NO_MASK_AND_ALL_VALS         = 6
};

// One 8^3 leaf: origin, active mask, and float values or bool value bits.
struct Leaf
{
  int32_t origin[3];
  uint64_t active[8];
  float values[512];
  uint64_t bits[8];
};

// One 8^3 tile of a pruned float grid (a 16^3 node table entry without a child).
struct Tile
{
  int32_t origin[3];
  float value;
  bool active;
};

struct Grid
{
  bool is_float   = true;
  std::string name;
  std::string grid_class;  // "level set", or empty (no class metadata)
  double voxel_size = 1.0;
  float background  = 0.0f;
  std::map<std::string, std::array<double, 3>> vec3d_meta;
  std::vector<Leaf> leaves;
  std::vector<Tile> tiles;  // written only for float grids; read back as leaves
};

namespace detail {

inline bool on(const uint64_t* m, uint32_t i) { return (m[i >> 6] >> (i & 63)) & 1; }

class Out
{
public:
  explicit Out(std::string& b) : buf(b) {}
  template <typename T>
  void put(const T& v)
  {
    buf.append(reinterpret_cast<const char*>(&v), sizeof(T));
  }
  void bytes(const void* p, size_t n) { buf.append(static_cast<const char*>(p), n); }
  void str(const std::string& s)
  {
    put<uint32_t>((uint32_t)s.size());
    buf.append(s);
  }
  std::string& buf;
};

class In
{
public:
  In(const std::string& b) : buf(b) {}
  template <typename T>
  bool get(T& v)
  {
    return bytes(&v, sizeof(T));
  }
  bool bytes(void* p, size_t n)
  {
    if (pos + n > buf.size())
    {
      return false;
    }
    std::memcpy(p, buf.data() + pos, n);
    pos += n;
    return true;
  }
  bool str(std::string& s)
  {
    uint32_t n = 0;
    if (!get(n) || pos + n > buf.size())
    {
      return false;
    }
    s.assign(buf.data() + pos, n);
    pos += n;
    return true;
  }
  const std::string& buf;
  size_t pos = 0;
};

inline std::string uuid()
{
  std::random_device rd;
  std::mt19937_64 gen(((uint64_t)rd() << 32) ^ rd());
  const char* hex = "0123456789abcdef";
  std::string s(36, '-');
  for (int i = 0; i < 36; ++i)
  {
    if (i == 8 || i == 13 || i == 18 || i == 23)
    {
      continue;
    }
    s[i] = hex[gen() & 15];
  }
  s[14] = '4';
  return s;
}

// A data block: raw, or with zlib a signed byte count (negative: stored
// uncompressed because compression did not shrink it) and the bytes.
inline void put_data(Out& o, const void* p, size_t bytes, uint32_t comp)
{
  if (!(comp & kCompressZip))
  {
    o.bytes(p, bytes);
    return;
  }
  uLongf len = compressBound((uLong)bytes);
  std::vector<Bytef> z(len);
  if (compress2(z.data(), &len, static_cast<const Bytef*>(p), (uLong)bytes, 1) == Z_OK && len < bytes)
  {
    o.put<int64_t>((int64_t)len);
    o.bytes(z.data(), len);
  }
  else
  {
    o.put<int64_t>(-(int64_t)bytes);
    o.bytes(p, bytes);
  }
}

// Values of n slots whose active slots are given by mask (child slots, if any,
// are never active) with active-mask compression: the header byte, the
// inactive values it needs, the selection mask, then the active values (all
// values without active-mask compression).
template <typename V>
void put_values(Out& o, const V* vals, uint32_t n, const uint64_t* active, const uint64_t* child, V bg, uint32_t comp)
{
  if (!(comp & kCompressMask))
  {
    o.put<int8_t>(NO_MASK_AND_ALL_VALS);
    put_data(o, vals, sizeof(V) * n, comp);
    return;
  }
  V inact[2] = {bg, bg};
  int unique = 0;
  for (uint32_t i = 0; i < n && unique < 3; ++i)
  {
    if (on(active, i) || (child && on(child, i)))
    {
      continue;
    }
    const V v = vals[i];
    // == as OpenVDB compares: -0.0 counts as the background 0.0 (and is read
    // back as 0.0; no map value is ever -0.0)
    const bool seen = (unique > 0 && v == inact[0]) || (unique > 1 && v == inact[1]);
    if (!seen)
    {
      if (unique < 2)
      {
        inact[unique] = v;
      }
      ++unique;
    }
  }
  int8_t meta = NO_MASK_OR_INACTIVE_VALS;
  const V neg = (V)(-bg);
  if (unique == 1)
  {
    if (!(inact[0] == bg))
    {
      meta = (inact[0] == neg) ? NO_MASK_AND_MINUS_BG : NO_MASK_AND_ONE_INACTIVE_VAL;
    }
  }
  else if (unique == 2)
  {
    if (!(inact[0] == bg) && !(inact[1] == bg))
    {
      meta = MASK_AND_TWO_INACTIVE_VALS;
    }
    else if (inact[1] == bg)
    {
      meta = (inact[0] == neg) ? MASK_AND_NO_INACTIVE_VALS : MASK_AND_ONE_INACTIVE_VAL;
    }
    else
    {
      std::swap(inact[0], inact[1]);
      meta = (inact[0] == neg) ? MASK_AND_NO_INACTIVE_VALS : MASK_AND_ONE_INACTIVE_VAL;
    }
  }
  else if (unique > 2)
  {
    meta = NO_MASK_AND_ALL_VALS;
  }
  o.put(meta);
  if (meta == NO_MASK_AND_ONE_INACTIVE_VAL || meta == MASK_AND_ONE_INACTIVE_VAL ||
      meta == MASK_AND_TWO_INACTIVE_VALS)
  {
    o.put(inact[0]);
    if (meta == MASK_AND_TWO_INACTIVE_VALS)
    {
      o.put(inact[1]);
    }
  }
  if (meta == NO_MASK_AND_ALL_VALS)
  {
    put_data(o, vals, sizeof(V) * n, comp);
    return;
  }
  if (meta == MASK_AND_NO_INACTIVE_VALS || meta == MASK_AND_ONE_INACTIVE_VAL || meta == MASK_AND_TWO_INACTIVE_VALS)
  {
    std::vector<uint64_t> sel((n + 63) / 64, 0);
    for (uint32_t i = 0; i < n; ++i)
    {
      if (!on(active, i) && !(child && on(child, i)) && vals[i] == inact[1])
      {
        sel[i >> 6] |= 1ull << (i & 63);
      }
    }
    o.bytes(sel.data(), sel.size() * 8);
  }
  std::string act;
  Out a(act);
  for (uint32_t i = 0; i < n; ++i)
  {
    if (on(active, i))
    {
      a.put(vals[i]);
    }
  }
  put_data(o, act.data(), act.size(), comp);
}

// Raw, zlib or (refused) blosc data block of `bytes` bytes into dst.
inline bool get_data(In& in, void* dst, size_t bytes, uint32_t comp, std::string& err)
{
  if (comp & kCompressBlosc)
  {
    err = "blosc-compressed grid data is not supported";
    return false;
  }
  if (!(comp & kCompressZip))
  {
    return in.bytes(dst, bytes);
  }
  int64_t n = 0;
  if (!in.get(n))
  {
    return false;
  }
  if (n <= 0)
  {
    if ((size_t)(-n) != bytes)
    {
      err = "uncompressed block size mismatch";
      return false;
    }
    return in.bytes(dst, bytes);
  }
  if (in.pos + (size_t)n > in.buf.size())
  {
    return false;
  }
  uLongf out = (uLongf)bytes;
  if (uncompress(static_cast<Bytef*>(dst), &out, reinterpret_cast<const Bytef*>(in.buf.data() + in.pos), (uLong)n) !=
        Z_OK ||
      out != bytes)
  {
    err = "zlib block does not decompress";
    return false;
  }
  in.pos += (size_t)n;
  return true;
}

template <typename V>
bool get_values(In& in, V* dst, uint32_t n, const uint64_t* active, uint32_t comp, V bg, std::string& err)
{
  int8_t meta = NO_MASK_AND_ALL_VALS;
  if (!in.get(meta))
  {
    return false;
  }
  V v1 = bg, v0 = (meta == NO_MASK_OR_INACTIVE_VALS) ? bg : (V)(-bg);
  if (meta == NO_MASK_AND_ONE_INACTIVE_VAL || meta == MASK_AND_ONE_INACTIVE_VAL || meta == MASK_AND_TWO_INACTIVE_VALS)
  {
    if (!in.get(v0) || (meta == MASK_AND_TWO_INACTIVE_VALS && !in.get(v1)))
    {
      return false;
    }
  }
  std::vector<uint64_t> sel((n + 63) / 64, 0);
  if (meta == MASK_AND_NO_INACTIVE_VALS || meta == MASK_AND_ONE_INACTIVE_VAL || meta == MASK_AND_TWO_INACTIVE_VALS)
  {
    if (!in.bytes(sel.data(), sel.size() * 8))
    {
      return false;
    }
  }
  uint32_t count = n;
  if ((comp & kCompressMask) && meta != NO_MASK_AND_ALL_VALS)
  {
    count = 0;
    for (uint32_t i = 0; i < n; ++i) count += on(active, i) ? 1 : 0;
  }
  std::unique_ptr<V[]> tmp(new V[count ? count : 1]);
  if (!get_data(in, tmp.get(), sizeof(V) * count, comp, err))  // an empty block still has its header
  {
    return false;
  }
  if (count == n)
  {
    std::copy(tmp.get(), tmp.get() + n, dst);
    return true;
  }
  for (uint32_t i = 0, t = 0; i < n; ++i)
  {
    dst[i] = on(active, i) ? tmp[t++] : (on(sel.data(), i) ? v1 : v0);
  }
  return true;
}

// Internal nodes of a tree: 32^3 (4096^3 voxels) holding 16^3 (128^3 voxels)
// holding leaves. Children sorted by slot index.
struct Node16
{
  std::map<uint32_t, size_t> leaves;  // slot -> leaf index
  std::map<uint32_t, size_t> tiles;   // slot -> tile index
};
struct Node32
{
  std::map<uint32_t, Node16> kids;
};

inline uint32_t slot32(const int32_t* o) { return (((o[0] & 4095) >> 7) << 10) | (((o[1] & 4095) >> 7) << 5) | ((o[2] & 4095) >> 7); }
inline uint32_t slot16(const int32_t* o) { return (((o[0] & 127) >> 3) << 8) | (((o[1] & 127) >> 3) << 4) | ((o[2] & 127) >> 3); }

template <typename V>
void put_internal_values(Out& o, uint32_t n, const std::vector<uint64_t>& child, V bg, uint32_t comp)
{
  std::vector<uint64_t> none((n + 63) / 64, 0);
  std::unique_ptr<V[]> vals(new V[n]);
  for (uint32_t i = 0; i < n; ++i) vals[i] = on(child.data(), i) ? V() : bg;  // child slots hold zero
  put_values<V>(o, vals.get(), n, none.data(), child.data(), bg, comp);
}

inline void put_meta(Out& o, const Grid& g)
{
  uint32_t count = (uint32_t)g.vec3d_meta.size() + (g.grid_class.empty() ? 0 : 1) + (g.name.empty() ? 0 : 1);
  o.put(count);
  // MetaMap order: by name
  std::map<std::string, std::pair<std::string, std::string>> items;
  for (const auto& kv : g.vec3d_meta)
  {
    items[kv.first] = {"vec3d", std::string(reinterpret_cast<const char*>(kv.second.data()), 24)};
  }
  if (!g.grid_class.empty())
  {
    items["class"] = {"string", g.grid_class};
  }
  if (!g.name.empty())
  {
    items["name"] = {"string", g.name};
  }
  for (const auto& kv : items)
  {
    o.str(kv.first);
    o.str(kv.second.first);
    o.put<uint32_t>((uint32_t)kv.second.second.size());
    o.bytes(kv.second.second.data(), kv.second.second.size());
  }
}

} // namespace detail

// Writes one grid as an io::File (seekable: grid offsets filled in) or an
// io::Stream archive; comp: the compression flags (active-mask by default,
// zlib optional).
inline std::string write_archive(const Grid& g, bool seekable, uint32_t comp = kCompressMask)
{
  using namespace detail;
  std::string buf;
  Out o(buf);
  o.put(kMagic);
  o.put(kFileVersion);
  o.put<uint32_t>(9);  // library version written as 9.1
  o.put<uint32_t>(1);
  o.put<uint8_t>(seekable ? 1 : 0);
  const std::string id = uuid();
  o.bytes(id.data(), 36);
  o.put<uint32_t>(0);  // file metadata
  o.put<int32_t>(1);   // grids
  o.str(g.name.empty() ? std::string("\x1e") + "0" : g.name);
  o.str(g.is_float ? kFloatTree : kBoolTree);
  o.str("");
  const size_t offsets_at = buf.size();
  o.put<int64_t>(0);
  o.put<int64_t>(0);
  o.put<int64_t>(0);
  const int64_t grid_pos = (int64_t)buf.size();
  o.put<uint32_t>(comp);
  put_meta(o, g);
  o.str("UniformScaleMap");
  const double s = g.voxel_size;
  const double map[15] = {s, s, s, s, s, s, 1 / s, 1 / s, 1 / s, 1 / (s * s), 1 / (s * s), 1 / (s * s),
                          0.5 / s, 0.5 / s, 0.5 / s};
  o.bytes(map, sizeof map);
  // topology
  std::map<std::array<int32_t, 3>, Node32> roots;
  for (size_t i = 0; i < g.leaves.size(); ++i)
  {
    const int32_t* lo = g.leaves[i].origin;
    const std::array<int32_t, 3> ro = {lo[0] & ~4095, lo[1] & ~4095, lo[2] & ~4095};
    roots[ro].kids[slot32(lo)].leaves[slot16(lo)] = i;
  }
  for (size_t i = 0; g.is_float && i < g.tiles.size(); ++i)
  {
    const int32_t* to = g.tiles[i].origin;
    const std::array<int32_t, 3> ro = {to[0] & ~4095, to[1] & ~4095, to[2] & ~4095};
    roots[ro].kids[slot32(to)].tiles[slot16(to)] = i;
  }
  o.put<int32_t>(1);  // buffer count
  if (g.is_float)
  {
    o.put(g.background);
  }
  else
  {
    o.put<uint8_t>(0);
  }
  o.put<uint32_t>(0);  // tiles
  o.put<uint32_t>((uint32_t)roots.size());
  std::vector<size_t> order;  // leaves in tree order
  for (const auto& r : roots)
  {
    o.bytes(r.first.data(), 12);
    std::vector<uint64_t> c32(512, 0), v32(512, 0);
    for (const auto& k : r.second.kids) c32[k.first >> 6] |= 1ull << (k.first & 63);
    o.bytes(c32.data(), 4096);
    o.bytes(v32.data(), 4096);
    if (g.is_float)
    {
      put_internal_values<float>(o, 32768, c32, g.background, comp);
    }
    else
    {
      put_internal_values<bool>(o, 32768, c32, false, comp);
    }
    for (const auto& k : r.second.kids)
    {
      std::vector<uint64_t> c16(64, 0), v16(64, 0);
      for (const auto& l : k.second.leaves) c16[l.first >> 6] |= 1ull << (l.first & 63);
      for (const auto& t : k.second.tiles)
      {
        if (g.tiles[t.second].active)
        {
          v16[t.first >> 6] |= 1ull << (t.first & 63);
        }
      }
      o.bytes(c16.data(), 512);
      o.bytes(v16.data(), 512);
      if (g.is_float && !k.second.tiles.empty())
      {
        // the table: zero at child slots, the tile values, background elsewhere
        std::unique_ptr<float[]> vals(new float[4096]);
        for (uint32_t i = 0; i < 4096; ++i) vals[i] = on(c16.data(), i) ? 0.0f : g.background;
        for (const auto& t : k.second.tiles) vals[t.first] = g.tiles[t.second].value;
        put_values<float>(o, vals.get(), 4096, v16.data(), c16.data(), g.background, comp);
      }
      else if (g.is_float)
      {
        put_internal_values<float>(o, 4096, c16, g.background, comp);
      }
      else
      {
        put_internal_values<bool>(o, 4096, c16, false, comp);
      }
      for (const auto& l : k.second.leaves)
      {
        o.bytes(g.leaves[l.second].active, 64);
        order.push_back(l.second);
      }
    }
  }
  // buffers
  const int64_t block_pos = (int64_t)buf.size();
  for (size_t i : order)
  {
    const Leaf& l = g.leaves[i];
    o.bytes(l.active, 64);
    if (g.is_float)
    {
      put_values<float>(o, l.values, 512, l.active, nullptr, g.background, comp);
    }
    else
    {
      o.bytes(l.origin, 12);
      o.bytes(l.bits, 64);
    }
  }
  if (seekable)
  {
    const int64_t end_pos = (int64_t)buf.size();
    std::memcpy(&buf[offsets_at], &grid_pos, 8);
    std::memcpy(&buf[offsets_at + 8], &block_pos, 8);
    std::memcpy(&buf[offsets_at + 16], &end_pos, 8);
  }
  return buf;
}


namespace detail {

template <typename V>
void tile_leaf(Grid& g, int32_t x, int32_t y, int32_t z, V value, bool active)
{
  Leaf l;
  std::memset(&l, 0, sizeof l);
  l.origin[0] = x;
  l.origin[1] = y;
  l.origin[2] = z;
  for (int w = 0; w < 8; ++w)
  {
    l.active[w] = active ? ~0ull : 0ull;
    l.bits[w]   = (!g.is_float && value) ? ~0ull : 0ull;
  }
  for (int v = 0; v < 512; ++v) l.values[v] = g.is_float ? (float)value : 0.0f;
  g.leaves.push_back(l);
}

template <typename V>
bool read_tree(In& in, Grid& g, uint32_t comp, std::string& err)
{
  int32_t nbuf = 0;
  if (!in.get(nbuf) || nbuf != 1)
  {
    err = "unexpected buffer count";
    return false;
  }
  V bg{};
  uint32_t ntiles = 0, nkids = 0;
  if (!in.get(bg) || !in.get(ntiles) || !in.get(nkids))
  {
    return false;
  }
  g.background = g.is_float ? (float)bg : 0.0f;
  for (uint32_t t = 0; t < ntiles; ++t)
  {
    int32_t o[3];
    V v{};
    uint8_t act = 0;
    if (!in.bytes(o, 12) || !in.get(v) || !in.get(act))
    {
      return false;
    }
    if (act || !(v == bg))
    {
      err = "a root-level tile (4096^3 voxels) that is active or not background";
      return false;
    }
  }
  std::vector<size_t> order;  // leaves with buffers, tree order
  for (uint32_t c = 0; c < nkids; ++c)
  {
    int32_t ro[3];
    if (!in.bytes(ro, 12))
    {
      return false;
    }
    std::vector<uint64_t> c32(512), v32(512);
    std::unique_ptr<V[]> t32(new V[32768]);
    if (!in.bytes(c32.data(), 4096) || !in.bytes(v32.data(), 4096) ||
        !get_values<V>(in, t32.get(), 32768, v32.data(), comp, bg, err))
    {
      return false;
    }
    for (uint32_t n = 0; n < 32768; ++n)
    {
      const int32_t x = ro[0] + (int32_t)((n >> 10) << 7), y = ro[1] + (int32_t)(((n >> 5) & 31) << 7),
                    z = ro[2] + (int32_t)((n & 31) << 7);
      if (!on(c32.data(), n))
      {
        if (on(v32.data(), n) || !(t32[n] == bg))
        {
          for (int i = 0; i < 4096; ++i)
          {
            tile_leaf<V>(g, x + ((i >> 8) << 3), y + (((i >> 4) & 15) << 3), z + ((i & 15) << 3), t32[n],
                         on(v32.data(), n));
          }
        }
        continue;
      }
      std::vector<uint64_t> c16(64), v16(64);
      std::unique_ptr<V[]> t16(new V[4096]);
      if (!in.bytes(c16.data(), 512) || !in.bytes(v16.data(), 512) ||
          !get_values<V>(in, t16.get(), 4096, v16.data(), comp, bg, err))
      {
        return false;
      }
      for (uint32_t k = 0; k < 4096; ++k)
      {
        const int32_t lx = x + (int32_t)((k >> 8) << 3), ly = y + (int32_t)(((k >> 4) & 15) << 3),
                      lz = z + (int32_t)((k & 15) << 3);
        if (!on(c16.data(), k))
        {
          if (on(v16.data(), k) || !(t16[k] == bg))
          {
            tile_leaf<V>(g, lx, ly, lz, t16[k], on(v16.data(), k));
          }
          continue;
        }
        Leaf l;
        std::memset(&l, 0, sizeof l);
        l.origin[0] = lx;
        l.origin[1] = ly;
        l.origin[2] = lz;
        if (!in.bytes(l.active, 64))
        {
          return false;
        }
        order.push_back(g.leaves.size());
        g.leaves.push_back(l);
      }
    }
  }
  for (size_t i : order)
  {
    Leaf& l = g.leaves[i];
    if (!in.bytes(l.active, 64))
    {
      return false;
    }
    if (g.is_float)
    {
      if (!get_values<float>(in, l.values, 512, l.active, comp, (float)bg, err))
      {
        return false;
      }
    }
    else
    {
      int32_t o[3];
      if (!in.bytes(o, 12) || !in.bytes(l.bits, 64))
      {
        return false;
      }
    }
  }
  return true;
}

} // namespace detail

// Reads every grid of an io::File or io::Stream archive (either layout);
// false with a reason in err for anything outside what this reader takes.
inline bool read_archive(const std::string& buf, std::vector<Grid>& grids, std::string& err)
{
  using namespace detail;
  In in(buf);
  int64_t magic = 0;
  uint32_t ver = 0, maj = 0, min = 0;
  uint8_t offsets = 0;
  char id[36];
  if (!in.get(magic) || magic != kMagic)
  {
    err = "not a VDB archive";
    return false;
  }
  if (!in.get(ver) || ver < 222)
  {
    err = "VDB file format older than 222";
    return false;
  }
  uint32_t nmeta = 0;
  if (!in.get(maj) || !in.get(min) || !in.get(offsets) || !in.bytes(id, 36) || !in.get(nmeta))
  {
    err = "truncated header";
    return false;
  }
  for (uint32_t i = 0; i < nmeta; ++i)
  {
    std::string name, type;
    uint32_t size = 0;
    if (!in.str(name) || !in.str(type) || !in.get(size) || in.pos + size > buf.size())
    {
      err = "truncated file metadata";
      return false;
    }
    in.pos += size;
  }
  int32_t ngrids = 0;
  if (!in.get(ngrids) || ngrids < 1)
  {
    err = "no grid in the archive";
    return false;
  }
  for (int32_t gi = 0; gi < ngrids; ++gi)
  {
    Grid g;
    std::string uname, type, parent;
    int64_t pos[3];
    uint32_t comp = 0;
    if (!in.str(uname) || !in.str(type) || !in.str(parent) || !in.bytes(pos, 24) || !in.get(comp))
    {
      err = "truncated grid descriptor";
      return false;
    }
    if (!parent.empty())
    {
      err = "instanced grids are not supported";
      return false;
    }
    if (type == kFloatTree)
    {
      g.is_float = true;
    }
    else if (type == kBoolTree)
    {
      g.is_float = false;
    }
    else
    {
      err = "grid type " + type + " is not supported (float or bool 5-4-3 trees only)";
      return false;
    }
    const size_t sep = uname.find('\x1e');
    g.name = sep == std::string::npos ? uname : uname.substr(0, sep);
    uint32_t ngm = 0;
    if (!in.get(ngm))
    {
      return false;
    }
    for (uint32_t i = 0; i < ngm; ++i)
    {
      std::string name, mtype;
      uint32_t size = 0;
      if (!in.str(name) || !in.str(mtype) || !in.get(size) || in.pos + size > buf.size())
      {
        err = "truncated grid metadata";
        return false;
      }
      if (mtype == "vec3d" && size == 24)
      {
        std::array<double, 3> v;
        std::memcpy(v.data(), buf.data() + in.pos, 24);
        g.vec3d_meta[name] = v;
      }
      else if (mtype == "string" && name == "class")
      {
        g.grid_class.assign(buf.data() + in.pos, size);
      }
      else if (mtype == "string" && name == "name")
      {
        g.name.assign(buf.data() + in.pos, size);
      }
      in.pos += size;
    }
    std::string map_type;
    double m[15];
    if (!in.str(map_type) || (map_type != "UniformScaleMap" && map_type != "ScaleMap") || !in.bytes(m, sizeof m))
    {
      err = "transform " + map_type + " is not supported (scale maps only)";
      return false;
    }
    g.voxel_size = m[3];
    if (!(g.is_float ? read_tree<float>(in, g, comp, err) : read_tree<bool>(in, g, comp, err)))
    {
      if (err.empty())
      {
        err = "truncated grid data";
      }
      return false;
    }
    grids.push_back(std::move(g));
  }
  return true;
}

} // namespace vdbio
} // namespace vdb_mapping

#endif // VDB_MAPPING_VDB_IO_HPP
