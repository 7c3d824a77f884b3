// The reference's unit tests (tests/mapping.cpp:6-158) re-expressed against
// the drop-in headers (include/vdb_mapping/), plus checks of the live grid,
// raycastPointCloud and raytrace. Plain asserts: googletest is not required.
// Needs a gfx950 GPU; prints one line per test and exits non-zero on failure.
#include <dirent.h>
#include <stdlib.h>
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <map>
#include <array>
#include <vdb_mapping/OccupancyVDBMapping.hpp>

using namespace vdb_mapping;

static int g_failures = 0;
#define CHECK(cond)                                                         \
  do {                                                                      \
    if (!(cond)) {                                                          \
      std::printf("  FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);       \
      ++g_failures;                                                         \
    }                                                                       \
  } while (0)

static Config make_config(double max_range)
{
  Config conf;
  conf.max_range           = max_range;
  conf.fast_mode           = false;
  conf.accumulation_period = 0;
  conf.prob_hit            = 0.9;
  conf.prob_miss           = 0.1;
  conf.prob_thres_max      = 0.51;
  conf.prob_thres_min      = 0.49;
  return conf;
}

static float logodds(double p) { return static_cast<float>(std::log(p) - std::log(1 - p)); }

static void set_config_test()
{
  OccupancyVDBMapping map(1);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(0, 0, 1);
  Eigen::Matrix<double, 3, 1> origin(0, 0, 0);
  map.insertPointCloud(cloud, origin, "test");  // before setConfig / addInputSource: no-op
  OccupancyVDBMapping::GridT::Accessor acc = map.getGrid()->getAccessor();
  CHECK(acc.getValue(openvdb::Coord(0, 0, 1)) == 0.0);
  Config conf = make_config(10);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  map.insertPointCloud(cloud, origin, "test");
  // the accessor taken before the insert sees the new values (live grid)
  CHECK(acc.getValue(openvdb::Coord(0, 0, 0)) == logodds(conf.prob_miss));
  CHECK(acc.getValue(openvdb::Coord(0, 0, 1)) == logodds(conf.prob_hit));
}

static void insert_point_test(int sign)
{
  const double resolution = 0.1;
  OccupancyVDBMapping map(resolution);
  Config conf = make_config(10);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(0, 0, sign * 5 * resolution);
  Eigen::Matrix<double, 3, 1> origin(0, 0, 0);
  map.insertPointCloud(cloud, origin, "test");
  OccupancyVDBMapping::GridT::Accessor acc = map.getGrid()->getAccessor();
  for (int i = 0; i < 5; ++i)
  {
    openvdb::Coord coord(0, 0, sign * i);
    CHECK(acc.getValue(coord) == logodds(conf.prob_miss));
    CHECK(!acc.isValueOn(coord));
  }
  openvdb::Coord coord(0, 0, sign * 5);
  CHECK(acc.getValue(coord) == logodds(conf.prob_hit));
  CHECK(acc.isValueOn(coord));
}

static void insert_max_range_test()
{
  const double resolution = 0.1;
  OccupancyVDBMapping map(resolution);
  Config conf = make_config(0.5);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(0, 0, 7 * resolution);
  Eigen::Matrix<double, 3, 1> origin(0, 0, 0);
  map.insertPointCloud(cloud, origin, "test");
  OccupancyVDBMapping::GridT::Accessor acc = map.getGrid()->getAccessor();
  for (int i = 0; i <= (int)(conf.max_range / resolution); ++i)
  {
    openvdb::Coord coord(0, 0, i);
    CHECK(acc.getValue(coord) == logodds(conf.prob_miss));
    CHECK(!acc.isValueOn(coord));
  }
  openvdb::Coord coord(0, 0, (int)(conf.max_range / resolution) + 1);
  CHECK(acc.getValue(coord) == 0);
  CHECK(!acc.isValueOn(coord));
}

static void reset_map_test()
{
  OccupancyVDBMapping map(1);
  Config conf = make_config(10);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(0, 0, 1);
  Eigen::Matrix<double, 3, 1> origin(0, 0, 0);
  map.insertPointCloud(cloud, origin, "test");
  OccupancyVDBMapping::GridT::Accessor acc = map.getGrid()->getAccessor();
  CHECK(acc.getValue(openvdb::Coord(0, 0, 1)) == logodds(conf.prob_hit));
  map.resetMap();
  acc = map.getGrid()->getAccessor();
  CHECK(acc.getValue(openvdb::Coord(0, 0, 1)) == 0.0);
  CHECK(map.getGrid()->empty());
}

static void raycast_point_cloud_test()
{
  OccupancyVDBMapping map(1);
  Config conf = make_config(10);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(0, 0, 3);
  cloud->points.emplace_back(0, 0, 20);  // beyond the 5 m cast range: misses only
  Eigen::Matrix<double, 3, 1> origin(0, 0, 0);
  auto update = OccupancyVDBMapping::UpdateGridT::create(false);
  auto uacc   = update->getAccessor();
  CHECK(map.raycastPointCloud(cloud, origin, 5.0, uacc));
  CHECK(uacc.isValueOn(openvdb::Coord(0, 0, 0)) && !uacc.getValue(openvdb::Coord(0, 0, 0)));
  CHECK(uacc.isValueOn(openvdb::Coord(0, 0, 3)) && uacc.getValue(openvdb::Coord(0, 0, 3)));
  CHECK(uacc.isValueOn(openvdb::Coord(0, 0, 5)) && !uacc.getValue(openvdb::Coord(0, 0, 5)));
  CHECK(!uacc.isValueOn(openvdb::Coord(0, 0, 6)));
  CHECK(update->activeVoxelCount() == 6);
  CHECK(map.getGrid()->empty());  // the map itself is untouched
  map.insertPointCloud(cloud, origin, "test");
  CHECK(map.getGrid()->getAccessor().isValueOn(openvdb::Coord(0, 0, 3)));
}

static void raytrace_test()
{
  OccupancyVDBMapping map(1);
  Config conf = make_config(50);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  for (int y = -4; y <= 4; ++y)
    for (int z = -4; z <= 4; ++z) cloud->points.emplace_back(10, y, z);
  Eigen::Matrix<double, 3, 1> origin(0, 0, 0);
  map.insertPointCloud(cloud, origin, "test");
  conf.fast_mode = true;
  map.setConfig(conf);
  map.integrateUpdate();  // builds the raytrace snapshot (updateVolumeRayIntersectors)
  bool success = false;
  openvdb::Vec3d end;
  map.raytrace(openvdb::Vec3d(0.5, 0.5, 0.5), openvdb::Vec3d(1, 0, 0), 20.0, success, end);
  CHECK(success);
  CHECK(end[0] == 10.0 && end[1] == 0.0 && end[2] == 0.0);
  map.raytrace(openvdb::Vec3d(0.5, 30.5, 0.5), openvdb::Vec3d(0, 1, 0), 4.0, success, end);
  CHECK(!success);
  CHECK(end[0] == 0.5 && end[1] == 34.5 && end[2] == 0.5);
}

static void add_data_to_accumulate_test()
{
  // the asynchronous front end: one cloud through addDataToAccumulate, then a
  // flush, gives the same map as insertPointCloud of that cloud
  OccupancyVDBMapping a(1), b(1);
  Config conf              = make_config(20);
  conf.accumulation_period = 60.0;  // the integration thread idles; the flush integrates
  a.setConfig(conf);
  b.setConfig(conf);
  a.addInputSource("test", conf.max_range, 0);
  b.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  for (int k = -3; k <= 3; ++k) cloud->points.emplace_back(6, k, 2 * k);
  Eigen::Matrix<double, 3, 1> origin(0.2, 0.1, 0.3);
  a.addDataToAccumulate(cloud, origin, "test");
  a.flushAccumulation();
  b.insertPointCloud(cloud, origin, "test");
  auto ga = a.getGrid()->getAccessor();
  auto gb = b.getGrid()->getAccessor();
  int same = 0;
  for (int x = -1; x <= 7; ++x)
    for (int y = -4; y <= 4; ++y)
      for (int z = -7; z <= 7; ++z)
      {
        const openvdb::Coord c(x, y, z);
        same += ga.getValue(c) == gb.getValue(c) && ga.isValueOn(c) == gb.isValueOn(c);
      }
  CHECK(same == 9 * 9 * 15);
  CHECK(ga.isValueOn(openvdb::Coord(6, 3, 6)));
  a.addDataToAccumulate(cloud, origin, "missing");  // prints, does nothing
}

static void point_edits_test()
{
  // addPointsToGrid / removePointsFromGrid (reference :413-447)
  OccupancyVDBMapping map(0.1);
  Config conf = make_config(10);
  map.setConfig(conf);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(0.05f, 0.05f, 0.05f);
  cloud->points.emplace_back(1.23f, -0.47f, 0.31f);
  CHECK(map.addPointsToGrid(cloud));
  auto acc           = map.getGrid()->getAccessor();
  const float c_max  = static_cast<float>(std::log(0.99) - std::log(0.01));
  const openvdb::Coord a(0, 0, 0), b(12, -5, 3);
  CHECK(acc.getValue(a) == c_max && acc.isValueOn(a));
  CHECK(acc.getValue(b) == c_max && acc.isValueOn(b));
  cloud->points.pop_back();
  CHECK(map.removePointsFromGrid(cloud));
  CHECK(acc.getValue(a) == -c_max && !acc.isValueOn(a));
  CHECK(acc.getValue(b) == c_max && acc.isValueOn(b));
}

static void map_section_test()
{
  // getMapSectionGrid -> applyMapSectionGrid into an empty map copies the
  // voxels inside the box; the update-grid section carries the active states
  OccupancyVDBMapping src(1), dst(1), upd(1);
  Config conf = make_config(30);
  for (auto* m : {&src, &dst, &upd})
  {
    m->setConfig(conf);
    m->addInputSource("test", conf.max_range, 0);
  }
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  for (int k = -6; k <= 6; ++k) cloud->points.emplace_back(12, k, k / 2);
  Eigen::Matrix<double, 3, 1> origin(0.5, 0.5, 0.5);
  src.insertPointCloud(cloud, origin, "test");
  const Eigen::Matrix<double, 3, 1> lo(2, -3, -2), hi(12.5, 3, 2);
  const Eigen::Matrix<double, 4, 4> tf = Eigen::Matrix<double, 4, 4>::Identity();
  const openvdb::CoordBBox box = src.createIndexBoundingBox(lo, hi, tf);
  CHECK(box.min() == openvdb::Coord(2, -3, -2) && box.max() == openvdb::Coord(12, 3, 2));
  auto section = src.getMapSectionGrid(lo, hi, tf, true);
  CHECK(section->metaValue<openvdb::Vec3d>("bb_max").x() == 12.0);
  dst.applyMapSectionGrid(section);
  auto sa = src.getGrid()->getAccessor();
  auto da = dst.getGrid()->getAccessor();
  int inside = 0, equal = 0, nonzero = 0;
  for (int x = 0; x <= 14; ++x)
    for (int y = -5; y <= 5; ++y)
      for (int z = -4; z <= 4; ++z)
      {
        const openvdb::Coord c(x, y, z);
        if (!box.isInside(c)) continue;
        ++inside;
        equal += sa.getValue(c) == da.getValue(c) && sa.isValueOn(c) == da.isValueOn(c);
        nonzero += da.getValue(c) != 0.0f;
      }
  CHECK(inside == equal && nonzero > 20);
  auto us = src.getMapSectionUpdateGrid(lo, hi, tf, false);
  CHECK(us->activeVoxelCount() == 7);  // the hit voxels (12, k, k / 2) with |k| <= 3
  upd.applyMapSectionUpdateGrid(us);
  CHECK(upd.getGrid()->activeVoxelCount() == 7);
  CHECK(upd.getGrid()->getAccessor().isValueOn(openvdb::Coord(12, 6, 3)) == false);  // outside the box
  // smooth_map closes the section first (in place, as the reference does):
  // the same as closing it by hand and applying it unsmoothed
  OccupancyVDBMapping sm(1), sm2(1);
  for (auto* m : {&sm, &sm2})
  {
    m->setConfig(conf);
    m->addInputSource("test", conf.max_range, 0);
  }
  auto us2 = src.getMapSectionUpdateGrid(lo, hi, tf, false);
  auto us3 = src.getMapSectionUpdateGrid(lo, hi, tf, false);
  sm.applyMapSectionUpdateGrid(us2, true, 1);
  sm2.morphologicalCloseMap<OccupancyVDBMapping::UpdateGridT>(us3, 1);
  sm2.applyMapSectionUpdateGrid(us3);
  CHECK(us2->activeVoxelCount() == us3->activeVoxelCount() && us2->activeVoxelCount() >= 7);
  CHECK(sm.getGrid()->activeVoxelCount() == us2->activeVoxelCount());
  CHECK(sm2.getGrid()->activeVoxelCount() == us2->activeVoxelCount());
  auto fs = src.getMapSectionGrid(lo, hi, tf, true);
  const openvdb::Index64 before = fs->activeVoxelCount();
  dst.applyMapSectionGrid(fs, true);
  CHECK(fs->activeVoxelCount() >= before);
}

static void morphology_test()
{
  // closed forms on one voxel: a 3^3 / 5^3 cube after 1 / 2 dilation steps,
  // back to one voxel after erosion; opening removes it; created leaves of a
  // float grid hold the background value
  OccupancyVDBMapping map(1);
  map.setConfig(make_config(30));
  auto g = OccupancyVDBMapping::UpdateGridT::create(false);
  g->getAccessor().setValueOn(openvdb::Coord(0, 0, 0), true);
  map.morphologicalDilateMap<OccupancyVDBMapping::UpdateGridT>(g, 1);
  CHECK(g->activeVoxelCount() == 27);
  CHECK(g->getAccessor().isValueOn(openvdb::Coord(-1, -1, -1)));
  map.morphologicalErodeMap<OccupancyVDBMapping::UpdateGridT>(g, 1);
  CHECK(g->activeVoxelCount() == 1 && g->getAccessor().isValueOn(openvdb::Coord(0, 0, 0)));
  map.morphologicalDilateMap<OccupancyVDBMapping::UpdateGridT>(g, 2);
  CHECK(g->activeVoxelCount() == 125);
  map.morphologicalCloseMap<OccupancyVDBMapping::UpdateGridT>(g, 2);
  CHECK(g->activeVoxelCount() == 125);
  map.morphologicalOpenMap<OccupancyVDBMapping::UpdateGridT>(g, 2);  // a 5^3 cube survives a 5^3 opening
  CHECK(g->activeVoxelCount() == 125);
  map.morphologicalOpenMap<OccupancyVDBMapping::UpdateGridT>(g, 3);
  CHECK(g->activeVoxelCount() == 0);
  // closing fills a one-voxel gap
  auto gap = OccupancyVDBMapping::UpdateGridT::create(false);
  gap->getAccessor().setValueOn(openvdb::Coord(7, 0, 0), true);
  gap->getAccessor().setValueOn(openvdb::Coord(9, 0, 0), true);
  map.morphologicalCloseMap<OccupancyVDBMapping::UpdateGridT>(gap, 1);
  CHECK(gap->activeVoxelCount() == 3 && gap->getAccessor().isValueOn(openvdb::Coord(8, 0, 0)));

  auto f = OccupancyVDBMapping::GridT::create(1.0);
  HostLeaf l;
  std::memset(&l, 0, sizeof l);
  l.values[0] = 2.5f;
  l.values[1] = -1.0f;  // inactive voxel (0, 0, 1) keeps its value when activated
  l.mask[0]   = 1;
  f->insertLeaf(l);
  map.morphologicalDilateMap<OccupancyVDBMapping::GridT>(f, 1);
  CHECK(f->activeVoxelCount() == 27 && f->leafCount() == 8);
  auto fa = f->getAccessor();
  CHECK(fa.getValue(openvdb::Coord(0, 0, 0)) == 2.5f && fa.getValue(openvdb::Coord(0, 0, 1)) == -1.0f);
  CHECK(fa.isValueOn(openvdb::Coord(-1, 0, 0)) && fa.getValue(openvdb::Coord(-1, 0, 0)) == 0.0f);
  map.morphologicalErodeMap<OccupancyVDBMapping::GridT>(f, 1);
  CHECK(f->activeVoxelCount() == 1 && f->leafCount() == 8);  // emptied leaves stay
}

static void artificial_areas_test()
{
  // addArtificialAreas: the walls' voxels are forced active after every
  // non-empty update; restoreMapIntegrity puts back value > t_max
  OccupancyVDBMapping map(1);
  Config conf = make_config(30);
  map.setConfig(conf);
  map.addInputSource("test", conf.max_range, 0);
  std::vector<std::vector<Eigen::Matrix<double, 4, 1> > > areas(1);
  areas[0].emplace_back(5.5, -5.5, 0.5, 1.0);
  areas[0].emplace_back(5.5, 5.5, 0.5, 1.0);
  areas[0].emplace_back(-5.5, 5.5, 0.5, 1.0);
  map.addArtificialAreas(areas, 0.0, 2.0);
  auto acc = map.getGrid()->getAccessor();
  const openvdb::Coord wall(6, 1, 2);  // on the first wall (x = 6), layer 1
  CHECK(!acc.isValueOn(wall));         // nothing until the next update
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  cloud->points.emplace_back(9.5f, 0.5f, 1.5f);  // its ray passes the wall voxel (a miss)
  Eigen::Matrix<double, 3, 1> origin(0.5, 0.5, 1.5);
  map.insertPointCloud(cloud, origin, "test");
  CHECK(acc.isValueOn(wall) && acc.getValue(wall) == logodds(conf.prob_miss));
  CHECK(acc.isValueOn(openvdb::Coord(0, 6, 2)));  // second wall, layer 1
  map.restoreMapIntegrity();
  CHECK(!acc.isValueOn(wall));
  CHECK(!acc.isValueOn(openvdb::Coord(0, 6, 2)));
  map.addArtificialWall(areas[0][0], areas[0][1], 0.0, 2.0);
  map.insertPointCloud(cloud, origin, "test");
  CHECK(acc.isValueOn(wall));
}

static void pcd_persistence_test()
{
  // saveMapToPCD -> loadMapFromPCD(clear) reproduces the active voxels
  // (reference :218-256, :295-307); a binary PCD file loads as well
  char dir_tmpl[] = "/tmp/vdbgpu_pcd_XXXXXX";
  const char* dir = mkdtemp(dir_tmpl);
  CHECK(dir != nullptr);
  if (!dir) return;
  OccupancyVDBMapping a(0.1), b(0.1);
  Config conf             = make_config(10);
  conf.map_directory_path = std::string(dir) + "/";
  a.setConfig(conf);
  b.setConfig(conf);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  for (int k = 0; k < 50; ++k) cloud->points.emplace_back(0.1f * k - 2.45f, 0.03f * k + 0.005f, -0.02f * k - 0.005f);
  CHECK(a.addPointsToGrid(cloud));
  CHECK(a.saveMapToPCD());
  std::string file;
  if (DIR* d = opendir(dir))
  {
    while (dirent* e = readdir(d))
    {
      const std::string name = e->d_name;
      if (name.size() > 4 && name.find("_active_values_map.pcd") != std::string::npos) file = std::string(dir) + "/" + name;
    }
    closedir(d);
  }
  CHECK(!file.empty());
  CHECK(b.loadMapFromPCD(file, false, true));
  CHECK(b.getGrid()->activeVoxelCount() == a.getGrid()->activeVoxelCount());
  CHECK(b.getGrid()->activeVoxelCount() == 50);
  auto aa = a.getGrid()->getAccessor();
  auto ba = b.getGrid()->getAccessor();
  int same = 0;
  a.getGrid()->forEachLeaf([&](const HostLeaf& l) {
    for (int n = 0; n < 512; ++n)
    {
      if (!((l.mask[n >> 6] >> (n & 63)) & 1)) continue;
      const openvdb::Coord c(l.origin[0] + (n >> 6), l.origin[1] + ((n >> 3) & 7), l.origin[2] + (n & 7));
      same += ba.isValueOn(c) && ba.getValue(c) == aa.getValue(c);
    }
  });
  CHECK(same == 50);
  // binary PCD with an extra field: x y z intensity
  const std::string bin = std::string(dir) + "/bin.pcd";
  {
    std::ofstream f(bin, std::ios::binary);
    f << "# .PCD v0.7\nVERSION 0.7\nFIELDS x y z intensity\nSIZE 4 4 4 4\nTYPE F F F F\nCOUNT 1 1 1 1\n"
         "WIDTH 2\nHEIGHT 1\nVIEWPOINT 0 0 0 1 0 0 0\nPOINTS 2\nDATA binary\n";
    const float pts[8] = {0.05f, 0.05f, 0.05f, 9.0f, 1.25f, -0.35f, 0.45f, 7.0f};
    f.write(reinterpret_cast<const char*>(pts), sizeof pts);
  }
  OccupancyVDBMapping c(0.1);
  c.setConfig(conf);
  CHECK(c.loadMapFromPCD(bin, true, true));
  auto ca = c.getGrid()->getAccessor();
  CHECK(ca.isValueOn(openvdb::Coord(0, 0, 0)) && ca.isValueOn(openvdb::Coord(12, -4, 4)));
  CHECK(!ca.isValueOn(openvdb::Coord(5, -2, 2)) && ca.getValue(openvdb::Coord(5, -2, 2)) < -4.5f);  // background
  CHECK(!c.loadMapFromPCD(std::string(dir) + "/missing.pcd", false, false));
  std::remove(file.c_str());
  std::remove(bin.c_str());
  rmdir(dir);
}

// Every leaf of a equals a leaf of b (origin, values bit for bit, mask) and
// the leaf counts agree.
static bool same_leaves(const OccupancyVDBMapping::GridT& a, const OccupancyVDBMapping::GridT& b)
{
  std::map<std::array<int32_t, 3>, const HostLeaf*> idx;
  b.forEachLeaf([&](const HostLeaf& l) { idx[{l.origin[0], l.origin[1], l.origin[2]}] = &l; });
  size_t n = 0;
  bool ok  = true;
  a.forEachLeaf([&](const HostLeaf& l) {
    ++n;
    auto it = idx.find({l.origin[0], l.origin[1], l.origin[2]});
    ok = ok && it != idx.end() && std::memcmp(l.values, it->second->values, sizeof l.values) == 0 &&
         std::memcmp(l.mask, it->second->mask, sizeof l.mask) == 0;
  });
  return ok && n == idx.size();
}

static void vdb_persistence_test()
{
  // saveMap -> loadMap into a second map gives the same grid (reference
  // :193-211, :263-284); byte arrays of section grids and update grids
  // round-trip through gridToByteArray / byteArrayToGrid (:1246-1338)
  char dir_tmpl[] = "/tmp/vdbgpu_vdb_XXXXXX";
  const char* dir = mkdtemp(dir_tmpl);
  CHECK(dir != nullptr);
  if (!dir) return;
  OccupancyVDBMapping a(0.1), b(0.1);
  Config conf             = make_config(10);
  conf.map_directory_path = std::string(dir) + "/";
  a.setConfig(conf);
  b.setConfig(conf);
  a.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  for (int k = 0; k < 400; ++k)
    cloud->points.emplace_back(3.0f * std::cos(0.05f * k), 2.0f * std::sin(0.07f * k), 0.01f * k - 2.0f);
  Eigen::Matrix<double, 3, 1> origin(0.05, -0.15, 0.25);
  a.insertPointCloud(cloud, origin, "test");
  a.insertPointCloud(cloud, origin, "test");
  CHECK(a.saveMap());
  std::string file;
  if (DIR* d = opendir(dir))
  {
    while (dirent* e = readdir(d))
    {
      const std::string name = e->d_name;
      if (name.find("_map.vdb") != std::string::npos) file = std::string(dir) + "/" + name;
    }
    closedir(d);
  }
  CHECK(!file.empty());
  CHECK(b.loadMap(file));
  CHECK(a.getGrid()->leafCount() > 100);
  CHECK(same_leaves(*a.getGrid(), *b.getGrid()));
  CHECK(!b.loadMap(std::string(dir) + "/missing.vdb"));
  OccupancyVDBMapping c(0.2);  // another resolution: refused
  c.setConfig(conf);
  CHECK(!c.loadMap(file));
  // byte arrays
  const Eigen::Matrix<double, 3, 1> lo(-1.5, -1.0, -2.0), hi(2.5, 1.5, 1.0);
  const Eigen::Matrix<double, 4, 4> tf = Eigen::Matrix<double, 4, 4>::Identity();
  auto section = a.getMapSectionGrid(lo, hi, tf, true);
  const std::vector<uint8_t> bytes = a.gridToByteArray<OccupancyVDBMapping::GridT>(section);
  auto back = b.byteArrayToGrid<OccupancyVDBMapping::GridT>(bytes);
  CHECK(back != nullptr);
  if (back)
  {
    CHECK(same_leaves(*section, *back));
    CHECK(back->metaValue<openvdb::Vec3d>("bb_max").x() == section->metaValue<openvdb::Vec3d>("bb_max").x());
    CHECK(back->voxelSize() == 0.1);
  }
  CHECK(b.byteArrayToGrid<OccupancyVDBMapping::UpdateGridT>(bytes) == nullptr);  // a float grid
  auto us       = a.getMapSectionUpdateGrid(lo, hi, tf, false);
  auto us_bytes = a.gridToByteArray<OccupancyVDBMapping::UpdateGridT>(us);
  auto us_back  = b.byteArrayToGrid<OccupancyVDBMapping::UpdateGridT>(us_bytes);
  CHECK(us_back != nullptr && us_back->activeVoxelCount() == us->activeVoxelCount() && us->activeVoxelCount() > 0);
  if (us_back)
  {
    int same = 0, total = 0;
    auto ua = us->getAccessor();
    auto ub = us_back->getAccessor();
    for (int x = -20; x <= 30; ++x)
      for (int y = -12; y <= 17; ++y)
        for (int z = -22; z <= 12; ++z)
        {
          const openvdb::Coord q(x, y, z);
          ++total;
          same += ua.isValueOn(q) == ub.isValueOn(q) && ua.getValue(q) == ub.getValue(q);
        }
    CHECK(same == total);
    // an update grid from a byte array applies like the original
    OccupancyVDBMapping d(0.1), e(0.1);
    for (auto* m : {&d, &e})
    {
      m->setConfig(conf);
    }
    d.applyMapSectionUpdateGrid(us);
    e.applyMapSectionUpdateGrid(us_back);
    CHECK(same_leaves(*d.getGrid(), *e.getGrid()));
  }
  const std::string text = "vdb byte array";
  CHECK(a.decompressByteArray(a.compressString(text)) == text);
  std::remove(file.c_str());
  rmdir(dir);
}

static std::string find_file(const char* dir, const char* suffix)
{
  std::string file;
  if (DIR* d = opendir(dir))
  {
    while (dirent* e = readdir(d))
    {
      const std::string name = e->d_name;
      if (name.find(suffix) != std::string::npos) file = std::string(dir) + "/" + name;
    }
    closedir(d);
  }
  return file;
}

static void pruned_map_test()
{
  // createMapFromPointCloud ends with pruneGrid (reference
  // OccupancyVDBMapping.hpp:173): a leaf full of points becomes an active
  // tile, a leaf of the background fill an inactive tile. The live grid reads
  // them as tiles; a scan that changes a voxel of a tile turns it into a leaf;
  // saveMap writes the tiles, loadMap reads them back as leaves.
  char dir_tmpl[] = "/tmp/vdbgpu_prune_XXXXXX";
  const char* dir = mkdtemp(dir_tmpl);
  CHECK(dir != nullptr);
  if (!dir) return;
  OccupancyVDBMapping a(0.1), b(0.1);
  Config conf             = make_config(10);
  conf.map_directory_path = std::string(dir) + "/";
  a.setConfig(conf);
  b.setConfig(conf);
  a.addInputSource("test", conf.max_range, 0);
  OccupancyVDBMapping::PointCloudT::Ptr cloud(new OccupancyVDBMapping::PointCloudT);
  for (int x = 0; x < 8; ++x)
    for (int y = 0; y < 8; ++y)
      for (int z = 0; z < 8; ++z) cloud->points.emplace_back(0.1f * x + 0.05f, 0.1f * y + 0.05f, 0.1f * z + 0.05f);
  cloud->points.emplace_back(2.35f, 0.05f, 0.05f);
  a.createMapFromPointCloud(cloud, true, true);
  auto g            = a.getGrid();
  const float c_max = logodds(0.99), c_min = logodds(0.01);
  CHECK(g->leafCount() == 1 && g->tileCount() == 2);
  CHECK(g->activeVoxelCount() == 513);
  auto acc = g->getAccessor();
  CHECK(acc.isValueOn(openvdb::Coord(3, 4, 5)) && acc.getValue(openvdb::Coord(3, 4, 5)) == c_max);
  CHECK(!acc.isValueOn(openvdb::Coord(12, 1, 1)) && acc.getValue(openvdb::Coord(12, 1, 1)) == c_min);
  CHECK(acc.isValueOn(openvdb::Coord(23, 0, 0)));
  // a hit at voxel (14, 3, 3), misses at 12 and 13 (clamped: no change)
  OccupancyVDBMapping::PointCloudT::Ptr one(new OccupancyVDBMapping::PointCloudT);
  one->points.emplace_back(1.425f, 0.325f, 0.325f);
  a.insertPointCloud(one, Eigen::Matrix<double, 3, 1>(1.225, 0.325, 0.325), "test");
  CHECK(g->leafCount() == 2 && g->tileCount() == 1);
  CHECK(acc.getValue(openvdb::Coord(14, 3, 3)) > c_min && acc.getValue(openvdb::Coord(13, 3, 3)) == c_min);
  CHECK(acc.getValue(openvdb::Coord(3, 4, 5)) == c_max);
  CHECK(a.saveMap());
  const std::string file = find_file(dir, "_map.vdb");
  CHECK(!file.empty() && b.loadMap(file));
  auto bacc  = b.getGrid()->getAccessor();
  int same   = 0;
  int total  = 0;
  for (int x = -1; x <= 24; ++x)
    for (int y = -1; y <= 8; ++y)
      for (int z = -1; z <= 8; ++z)
      {
        const openvdb::Coord q(x, y, z);
        ++total;
        same += bacc.getValue(q) == acc.getValue(q) && bacc.isValueOn(q) == acc.isValueOn(q);
      }
  CHECK(same == total);
  CHECK(b.getGrid()->leafCount() == 3 && b.getGrid()->tileCount() == 0);
  std::remove(file.c_str());
  rmdir(dir);
}

int main()
{
  struct T
  {
    const char* name;
    void (*fn)();
  } tests[] = {
    {"SetConfig", set_config_test},
    {"InsertPositivePoint", [] { insert_point_test(1); }},
    {"InsertNegativePoint", [] { insert_point_test(-1); }},
    {"InsertMaxRangePoint", insert_max_range_test},
    {"ResetMap", reset_map_test},
    {"RaycastPointCloud", raycast_point_cloud_test},
    {"Raytrace", raytrace_test},
    {"AddDataToAccumulate", add_data_to_accumulate_test},
    {"PointEdits", point_edits_test},
    {"MapSection", map_section_test},
    {"Morphology", morphology_test},
    {"ArtificialAreas", artificial_areas_test},
    {"PcdPersistence", pcd_persistence_test},
    {"VdbPersistence", vdb_persistence_test},
    {"PrunedMap", pruned_map_test},
  };
  for (const T& t : tests)
  {
    const int before = g_failures;
    t.fn();
    std::printf("[%s] %s\n", g_failures == before ? "  OK  " : "FAILED", t.name);
    std::fflush(stdout);
  }
  std::printf("%d failure(s)\n", g_failures);
  return g_failures == 0 ? 0 : 1;
}
