// Host-only checks of the OpenVDB archive code (include/vdb_mapping/vdb_io.hpp):
// grids written as io::File and io::Stream archives, with active-mask, zlib
// and no compression, read back identical; every inactive-value encoding of a
// leaf (background only, minus background, one or two other values, background
// and minus background, more than two) occurs; the 8^3 tiles of a pruned float
// grid, written into the 16^3 node tables, read back as leaves of the tile
// value and state (a background-valued inactive tile as nothing). No GPU.
// Prints one line per case and exits non-zero on failure.
#include <cstdio>
#include <cstring>
#include <random>
#include <vdb_mapping/vdb_io.hpp>

using namespace vdb_mapping::vdbio;

static int g_fail = 0;

static Leaf make_leaf(std::mt19937& rng, int32_t x, int32_t y, int32_t z, int mode, bool is_float, float bg)
{
  Leaf l;
  std::memset(&l, 0, sizeof l);
  l.origin[0] = x;
  l.origin[1] = y;
  l.origin[2] = z;
  std::uniform_real_distribution<float> u(-4.0f, 4.0f);
  for (int v = 0; v < 512; ++v)
  {
    const bool act = (rng() % 3) == 0;
    if (act) l.active[v >> 6] |= 1ull << (v & 63);
    float val = bg;
    switch (mode)
    {
      case 0: break;                                         // inactive = background
      case 1: val = -bg; break;                              // minus background
      case 2: val = 1.25f; break;                            // one other value
      case 3: val = (v & 1) ? bg : 2.5f; break;              // background and one other
      case 4: val = (v & 1) ? -3.0f : 2.5f; break;           // two other values
      case 5: val = (v & 1) ? bg : -bg; break;               // background and minus background
      default: val = u(rng); break;                          // many values
    }
    l.values[v] = act ? u(rng) : val;
    if (!is_float && (rng() & 1)) l.bits[v >> 6] |= 1ull << (v & 63);
  }
  return l;
}

static bool same(const Grid& a, const Grid& b)
{
  if (a.leaves.size() != b.leaves.size() || a.is_float != b.is_float || a.voxel_size != b.voxel_size ||
      a.background != b.background ||
      a.grid_class != b.grid_class || a.vec3d_meta != b.vec3d_meta)
    return false;
  for (const Leaf& l : a.leaves)
  {
    bool found = false;
    for (const Leaf& r : b.leaves)
    {
      if (std::memcmp(l.origin, r.origin, 12)) continue;
      found = std::memcmp(l.active, r.active, 64) == 0 &&
              (a.is_float ? std::memcmp(l.values, r.values, 2048) == 0 : std::memcmp(l.bits, r.bits, 64) == 0);
    }
    if (!found) return false;
  }
  return true;
}

int main()
{
  std::mt19937 rng(7);
  for (int is_float = 0; is_float < 2; ++is_float)
  {
    Grid g;
    g.is_float   = is_float;
    g.grid_class = is_float ? "level set" : "";
    g.voxel_size = 0.05;
    g.background = is_float ? 1.5f : 0.0f;  // distinct from -background, unlike 0
    g.vec3d_meta["bb_min"] = {-1.0, 2.0, -3.5};
    g.vec3d_meta["bb_max"] = {4.0, 5.0, 6.5};
    int k = 0;
    for (int32_t x : {-8200, -16, 0, 4088, 131072})
      for (int32_t y : {-8, 120})
        for (int32_t z : {0, 4096})
          g.leaves.push_back(make_leaf(rng, x, y, z, (k++) % 7, is_float, g.background));
    for (uint32_t comp : {kCompressMask, kCompressZip | kCompressMask, kCompressZip, 0u})
      for (int seekable = 0; seekable < 2; ++seekable)
      {
        const std::string bytes = write_archive(g, seekable, comp);
        std::vector<Grid> back;
        std::string err;
        const bool ok = read_archive(bytes, back, err) && back.size() == 1 && same(g, back[0]);
        std::printf("[%s] %s grid, %s, compression %u: %zu bytes %s\n", ok ? "  OK  " : "FAILED",
                    is_float ? "float" : "bool", seekable ? "file" : "stream", comp, bytes.size(), err.c_str());
        g_fail += ok ? 0 : 1;
      }
    // truncated and foreign input are refused, not read past the end
    const std::string bytes = write_archive(g, false);
    std::vector<Grid> back;
    std::string err;
    bool refused = !read_archive(bytes.substr(0, bytes.size() / 2), back, err);
    refused = refused && !read_archive(std::string("not a vdb file at all"), back, err);
    std::printf("[%s] %s grid: truncated / foreign input refused\n", refused ? "  OK  " : "FAILED",
                is_float ? "float" : "bool");
    g_fail += refused ? 0 : 1;
  }
  {
    // tiles beside leaves: in a node of their own, next to leaves, active and inactive
    Grid g;
    g.grid_class = "level set";
    g.voxel_size = 0.1;
    g.background = 0.0f;
    g.leaves.push_back(make_leaf(rng, 0, 0, 0, 2, true, 0.0f));
    g.leaves.push_back(make_leaf(rng, 4096, -128, 8, 6, true, 0.0f));
    const Tile tiles[] = {{{8, 0, 0}, 4.59512f, true}, {{16, 0, 0}, -4.59512f, false}, {{24, 8, 0}, 0.0f, false},
                          {{-136, 64, 120}, 1.0f, true}, {{4096, -128, 0}, 0.0f, true}};
    Grid want = g;
    for (const Tile& t : tiles)
    {
      g.tiles.push_back(t);
      if (!t.active && t.value == g.background) continue;
      Leaf l;
      std::memset(&l, 0, sizeof l);
      std::memcpy(l.origin, t.origin, 12);
      for (int w = 0; w < 8; ++w) l.active[w] = t.active ? ~0ull : 0ull;
      for (int v = 0; v < 512; ++v) l.values[v] = t.value;
      want.leaves.push_back(l);
    }
    for (uint32_t comp : {kCompressMask, kCompressZip | kCompressMask})
    {
      const std::string bytes = write_archive(g, true, comp);
      std::vector<Grid> back;
      std::string err;
      const bool ok = read_archive(bytes, back, err) && back.size() == 1 && same(want, back[0]);
      std::printf("[%s] float grid with 8^3 tiles, compression %u: %zu bytes %s\n", ok ? "  OK  " : "FAILED", comp,
                  bytes.size(), err.c_str());
      g_fail += ok ? 0 : 1;
    }
  }
  std::printf("%d failure(s)\n", g_fail);
  return g_fail ? 1 : 0;
}
