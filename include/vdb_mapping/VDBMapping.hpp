// Drop-in replacement of vdb_mapping::VDBMapping (include/vdb_mapping/VDBMapping.hpp
// of the reference) for the occupancy-integration path on one MI355X.
//
// Same class template, member names, argument order, return types and
// console messages as the reference for: the constructor (:115-134),
// createVDBMap (:163-169), resetMap (:174-186), accumulateUpdate (:316-346),
// integrateUpdate (:375-387), insertPointCloud (:399-406), raycastPointCloud
// (:466-539), worldToIndex (:612-631), raytrace (:643-721), getGrid (:799),
// getMapMutex (:1343), addInputSource (:1352-1375) and setConfig (:1456-1469).
// Every method on the integration and query path calls the C ABI of
// include/vdbgpu.h (no voxel arithmetic on the host); worldToIndex, a
// standalone helper in the reference, is restated on the host.
//
// Fast mode (:520-527, :577-602) is provided by the library: accumulates cast
// through the topology snapshot of the last fast-mode integrate.
//
// addDataToAccumulate with its accumulation and integration threads (:133,
// :355-370, :1383-1430) runs in the library (vdbgpu_async_*).
//
// Direct edits (addPointsToGrid / removePointsFromGrid :413-447), artificial
// areas (restoreMapIntegrity, addArtificialAreas / Polygon / Wall
// :1152-1236) and map sections (createWorldBoundingBox /
// createIndexBoundingBox, getMapSectionGrid / UpdateGrid, applyMapSectionGrid
// / UpdateGrid :811-1085) run in the library too (vdbgpu_set_points,
// vdbgpu_*artificial*, vdbgpu_*section*), and so do the morphology members
// (morphologicalClose / Open / Dilate / ErodeMap :1095-1148, vdbgpu_morphology)
// that smooth a section before it is applied.
//
// PCD persistence (saveMapToPCD, loadMapFromPCD, createMapFromPointCloud
// :218-256, :295-307, OccupancyVDBMapping.hpp:136-174) exports and loads through
// the library (vdbgpu_export_active_points, vdbgpu_create_from_points).
//
// .vdb persistence (saveMap / loadMap :193-211, :263-284) and the byte-array
// wire format (gridToByteArray / byteArrayToGrid, compressString /
// decompressByteArray :1246-1338) write and read OpenVDB's archive layout on
// the host (vdb_io.hpp, zstd); link zlib and zstd.
//
// Not provided: castRayIntoGrid (:550-566, a per-ray host call) and updateMap
// (:731-792, reached through integrateUpdate).
#ifndef VDB_MAPPING_VDB_MAPPING_H_INCLUDED
#define VDB_MAPPING_VDB_MAPPING_H_INCLUDED

#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <chrono>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iterator>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <map>
#include <memory>
#include <mutex>
#include <optional>
#include <shared_mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "../vdbgpu.h"
#include "vdb_mapping/compat.hpp"
#include "vdb_mapping/grid.hpp"
#include "vdb_mapping/pcd.hpp"
#include "vdb_mapping/vdb_io.hpp"

#if __has_include(<zstd.h>)
#include <zstd.h>
#else
// zstd's C API (libzstd), for the byte arrays of gridToByteArray
extern "C" {
size_t ZSTD_compressBound(size_t srcSize);
size_t ZSTD_compress(void* dst, size_t dstCapacity, const void* src, size_t srcSize, int compressionLevel);
unsigned long long ZSTD_getDecompressedSize(const void* src, size_t srcSize);
size_t ZSTD_decompress(void* dst, size_t dstCapacity, const void* src, size_t compressedSize);
unsigned ZSTD_isError(size_t code);
const char* ZSTD_getErrorName(size_t code);
}
#endif

namespace vdb_mapping {

/*!
 * \brief Accumulation of configuration parameters (reference :64-71)
 */
struct BaseConfig
{
  double max_range;
  bool fast_mode;
  double accumulation_period;
  std::string map_directory_path;
};

template <typename TData, typename TConfig = BaseConfig>
class VDBMapping
{
  static_assert(std::is_same<TData, float>::value,
                "the gfx950 path implements the float log-odds map (OccupancyVDBMapping)");

public:
  using PointT      = pcl::PointXYZ;
  using PointCloudT = pcl::PointCloud<PointT>;
  using Vec3T       = openvdb::Vec3d;
  using GridT       = FloatGrid;
  using UpdateGridT = BoolGrid;

  VDBMapping()                  = delete;
  VDBMapping(const VDBMapping&) = delete;
  VDBMapping& operator=(const VDBMapping&) = delete;

  /*!
   * \brief Creates the device map (background 0, voxel size `resolution`)
   * on the current HIP device. Throws std::runtime_error without a GPU.
   */
  VDBMapping(const double resolution)
    : m_resolution(resolution)
    , m_config_set(false)
  {
    m_map_mutex = std::make_shared<std::shared_mutex>();
    const int32_t rc = vdbgpu_create(resolution, &m_map);
    if (rc != VDBGPU_OK)
    {
      throw std::runtime_error(std::string("vdbgpu_create failed: ") + vdbgpu_status_string(rc));
    }
    m_vdb_grid = createVDBMap(m_resolution);
  }

  virtual ~VDBMapping()
  {
    if (m_vdb_grid)
    {
      m_vdb_grid->detach();
    }
    vdbgpu_destroy(m_map);
  }

  /*!
   * \brief A grid object bound to this map: reads pull the leaves changed
   * since the last read, so the pointer stays live like the reference's.
   */
  typename GridT::Ptr createVDBMap(double resolution)
  {
    auto g = std::make_shared<GridT>(resolution);
    g->attach(m_map);
    return g;
  }

  void resetMap()
  {
    std::unique_lock map_lock(*m_map_mutex);
    m_vdb_grid->clear();
    m_vdb_grid->detach();
    vdbgpu_reset(m_map);
    m_vdb_grid = createVDBMap(m_resolution);
  }

  void accumulateUpdate(const PointCloudT::ConstPtr& cloud,
                        const Eigen::Matrix<double, 3, 1>& origin,
                        const std::string source_id)
  {
    auto source = m_sources.find(source_id);
    if (source == m_sources.end())
    {
      std::cout << "Tried to accumulate update for " << source_id << ". Source not available"
                << std::endl;
      return;
    }
    std::shared_lock map_lock(*m_map_mutex);
    if (source->second > 0)
    {
      if (!m_config_set)
      {
        std::cerr << "Map not properly configured. Did you call setConfig method?" << std::endl;
        return;
      }
      report(accumulate(source_id, *cloud, origin));
    }
  }

  void integrateUpdate()
  {
    std::unique_lock map_lock(*m_map_mutex);
    report(vdbgpu_integrate(m_map));
    m_vdb_grid->markStale();
  }

  bool insertPointCloud(const PointCloudT::ConstPtr& cloud,
                        const Eigen::Matrix<double, 3, 1>& origin,
                        const std::string source_id)
  {
    accumulateUpdate(cloud, origin, source_id);
    integrateUpdate();
    return true;
  }

  /*!
   * \brief Casts the cloud into `update_grid_acc` (visited voxels active,
   * endpoints inside the range valued true) without touching the map.
   */
  bool raycastPointCloud(const PointCloudT::ConstPtr& cloud,
                         const Eigen::Matrix<double, 3, 1>& origin,
                         const double raycast_range,
                         UpdateGridT::Accessor& update_grid_acc)
  {
    if (!m_config_set)
    {
      std::cerr << "Map not properly configured. Did you call setConfig method?" << std::endl;
      return false;
    }
    static const char* kScratch = "\x01vdbgpu-raycast";
    if (vdbgpu_add_source(m_map, kScratch, raycast_range > 0 ? raycast_range : -1.0, 0) != VDBGPU_OK)
    {
      return false;
    }
    bool ok = true;
    if (raycast_range > 0.0)
    {
      ok = report(accumulate(kScratch, *cloud, origin)) == VDBGPU_OK;
      const int64_t n = vdbgpu_update_leaf_count(m_map, kScratch);
      if (ok && n > 0)
      {
        std::vector<int32_t> o(3 * n);
        std::vector<uint64_t> vis(8 * n), hit(8 * n);
        const int64_t got = vdbgpu_export_update(m_map, kScratch, o.data(), vis.data(), hit.data(), n);
        for (int64_t i = 0; i < got; ++i)
        {
          update_grid_acc.mergeLeaf(o.data() + 3 * i, vis.data() + 8 * i, hit.data() + 8 * i);
        }
      }
    }
    vdbgpu_remove_source(m_map, kScratch);
    return ok;
  }

  /*!
   * \brief Rounded world->index (reference :612-631), host arithmetic only.
   */
  openvdb::Coord worldToIndex(const openvdb::Vec3d& world_coordinate) const
  {
    double c[3] = {world_coordinate[0], world_coordinate[1], world_coordinate[2]};
    const double inv = 1.0 / m_resolution;
    int32_t out[3];
    for (int a = 0; a < 3; ++a)
    {
      if (std::fmod(c[a], m_resolution))
      {
        c[a] = c[a] + (m_resolution / 2.0);
      }
      out[a] = static_cast<int32_t>(std::floor(c[a] * inv));
    }
    return openvdb::Coord(out[0], out[1], out[2]);
  }

  void raytrace(const openvdb::Vec3d& ray_origin_world,
                const openvdb::Vec3d& ray_direction,
                const double max_ray_lengths,
                bool& success,
                openvdb::Vec3d& end_point)
  {
    std::vector<openvdb::Vec3d> ray_origins_world = {ray_origin_world};
    std::vector<openvdb::Vec3d> ray_directions    = {ray_direction};
    std::vector<double> max_ray_length            = {max_ray_lengths};
    std::vector<bool> successes;
    std::vector<openvdb::Vec3d> end_points;
    raytrace(ray_origins_world, ray_directions, max_ray_length, successes, end_points);
    success   = successes[0];
    end_point = end_points[0];
  }

  /*!
   * \brief Batch raytrace (reference :675-721). Without an intersector (no
   * integrate in fast mode yet) the reference dereferences a null pointer;
   * here the call prints the reason and reports every ray as unsuccessful.
   */
  void raytrace(const std::vector<openvdb::Vec3d>& ray_origins_world,
                const std::vector<openvdb::Vec3d>& ray_directions,
                const std::vector<double>& max_ray_lengths,
                std::vector<bool>& successes,
                std::vector<openvdb::Vec3d>& end_points)
  {
    std::shared_lock map_lock(*m_map_mutex);
    const size_t n = ray_origins_world.size();
    successes.assign(n, false);
    end_points.assign(n, openvdb::Vec3d(0, 0, 0));
    std::vector<double> o(3 * n), d(3 * n), l(max_ray_lengths.begin(), max_ray_lengths.end());
    l.resize(n, 0.0);
    for (size_t i = 0; i < n; ++i)
    {
      for (int a = 0; a < 3; ++a)
      {
        o[3 * i + a] = ray_origins_world[i][a];
        d[3 * i + a] = ray_directions[i][a];
      }
    }
    std::vector<uint8_t> ok(n);
    std::vector<double> e(3 * n);
    if (report(vdbgpu_raytrace(m_map, (int64_t)n, o.data(), d.data(), l.data(), ok.data(), e.data())) !=
        VDBGPU_OK)
    {
      return;
    }
    for (size_t i = 0; i < n; ++i)
    {
      successes[i]  = ok[i] != 0;
      end_points[i] = openvdb::Vec3d(e[3 * i], e[3 * i + 1], e[3 * i + 2]);
    }
  }

  typename GridT::Ptr getGrid() const { return m_vdb_grid; }

  std::shared_ptr<std::shared_mutex> getMapMutex() { return m_map_mutex; }

  void addInputSource(std::string source_id, double max_range, double max_rate)
  {
    vdbgpu_add_source(m_map, source_id.c_str(), max_range, max_rate);
    m_sources[source_id] = (max_range == 0) ? m_max_range : max_range;
    m_rates[source_id]   = max_rate;
  }

  /*!
   * \brief Hands a cloud to the source's accumulation thread (reference
   * :355-370): latest only, accumulated by the source's worker and integrated
   * by the integration thread every accumulation_period. The threads run in
   * the library (vdbgpu_async_*) and start with the first call.
   */
  void addDataToAccumulate(const PointCloudT::ConstPtr& cloud,
                           const Eigen::Matrix<double, 3, 1>& origin,
                           const std::string source_id)
  {
    auto it = m_rates.find(source_id);
    if (it == m_rates.end())
    {
      std::cout << "Tried to add data for accumulation of " << source_id
                << ". Source not available" << std::endl;
      return;
    }
    vdbgpu_async_start(m_map, m_accumulation_period);
    const double o[3] = {origin.x(), origin.y(), origin.z()};
    const float* xyz  = cloud->points.empty() ? nullptr : &cloud->points[0].x;
    report(vdbgpu_async_submit(m_map,
                               source_id.c_str(),
                               xyz,
                               static_cast<int64_t>(cloud->points.size()),
                               static_cast<int64_t>(sizeof(PointT)),
                               o,
                               it->second));
  }

  /*!
   * \brief Waits until the accumulation threads have taken and accumulated
   * every handed cloud, then integrates (no reference counterpart: the
   * reference offers no way to wait for its threads).
   */
  void flushAccumulation() { report(vdbgpu_async_flush(m_map)); }

  /*!
   * \brief Writes the map to <map_directory_path><date>_map.vdb (reference
   * :193-211) in OpenVDB's archive layout (vdb_io.hpp; compatibility with
   * OpenVDB itself is not checked by this build).
   */
  bool saveMap() const
  {
    auto timestamp     = std::chrono::system_clock::now();
    std::time_t now_tt = std::chrono::system_clock::to_time_t(timestamp);
    std::tm tm         = *std::localtime(&now_tt);
    std::stringstream sstime;
    sstime << std::put_time(&tm, "%Y-%m-%d_%H-%M-%S");
    std::string map_name = m_map_directory_path + sstime.str() + "_map.vdb";
    std::cout << map_name << std::endl;
    std::shared_lock map_lock(*m_map_mutex);
    const std::string bytes = vdbio::write_archive(toArchiveGrid(*m_vdb_grid), true);
    map_lock.unlock();
    std::ofstream out(map_name, std::ios::binary);
    out.write(bytes.data(), (std::streamsize)bytes.size());
    return (bool)out;
  }

  /*!
   * \brief Loads a map from a .vdb file (reference :263-284): the last grid of
   * the file becomes the map. Float 5-4-3 trees at the map's resolution only;
   * like resetMap, pending update sets are dropped.
   */
  bool loadMap(const std::string& file_path)
  {
    std::ifstream in(file_path, std::ios::binary);
    if (!in)
    {
      std::cerr << "Could not open " << file_path << std::endl;
      return false;
    }
    const std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    std::vector<vdbio::Grid> grids;
    std::string err;
    if (!vdbio::read_archive(bytes, grids, err) || !grids.back().is_float ||
        std::fabs(grids.back().voxel_size - m_resolution) > 1e-12 * m_resolution)
    {
      std::cerr << "loadMap: " << (err.empty() ? "not a float grid at the map's resolution" : err) << std::endl;
      return false;
    }
    const vdbio::Grid& g = grids.back();
    std::vector<int32_t> o;
    std::vector<float> v;
    std::vector<uint64_t> a;
    for (const auto& l : g.leaves)
    {
      o.insert(o.end(), l.origin, l.origin + 3);
      v.insert(v.end(), l.values, l.values + 512);
      a.insert(a.end(), l.active, l.active + 8);
    }
    std::unique_lock map_lock(*m_map_mutex);
    int32_t rc = report(vdbgpu_reset(m_map));
    if (rc == VDBGPU_OK && !o.empty())
    {
      rc = report(vdbgpu_apply_section(m_map, (int64_t)(o.size() / 3), o.data(), v.data(), a.data()));
    }
    m_vdb_grid->markStale();
    return rc == VDBGPU_OK;
  }

  /*!
   * \brief zstd compression of a string at m_compression_level (reference
   * :1246-1268); the uncompressed bytes when compression fails.
   */
  std::vector<uint8_t> compressString(const std::string& string) const
  {
    auto uncompressed = std::vector<uint8_t>(string.begin(), string.end());
    size_t len        = ZSTD_compressBound(uncompressed.size());
    std::vector<uint8_t> compressed(len);
    const size_t ret =
      ZSTD_compress(compressed.data(), len, uncompressed.data(), uncompressed.size(), (int)m_compression_level);
    if (ZSTD_isError(ret))
    {
      std::cerr << "Compression using ZSTD failed: " << ZSTD_getErrorName(ret)
                << " , sending uncompressed byte array" << std::endl;
      return uncompressed;
    }
    compressed.resize(ret);
    return compressed;
  }

  /*!
   * \brief Decompresses a zstd byte array into a string (reference :1277-1297).
   */
  std::string decompressByteArray(const std::vector<uint8_t>& byte_array) const
  {
    const unsigned long long len = ZSTD_getDecompressedSize(byte_array.data(), byte_array.size());
    std::vector<uint8_t> uncompressed((size_t)len);
    const size_t size = ZSTD_decompress(uncompressed.data(), (size_t)len, byte_array.data(), byte_array.size());
    if (ZSTD_isError(size))
    {
      std::cerr << "Could not decompress map using ZSTD failed: " << ZSTD_getErrorName(size) << std::endl;
      return std::string();
    }
    return std::string(uncompressed.begin(), uncompressed.begin() + size);
  }

  /*!
   * \brief A grid as a compressed byte array (reference :1310-1317): OpenVDB's
   * stream layout (vdb_io.hpp), compressed with zstd.
   */
  template <typename TGrid>
  std::vector<uint8_t> gridToByteArray(typename TGrid::Ptr grid)
  {
    return compressString(vdbio::write_archive(toArchiveGrid(*grid), false));
  }

  /*!
   * \brief The grid of a byte array (reference :1327-1338); null when the
   * array holds no grid of type TGrid.
   */
  template <typename TGrid>
  typename TGrid::Ptr byteArrayToGrid(std::vector<uint8_t> byte_array)
  {
    std::vector<vdbio::Grid> grids;
    std::string err;
    if (!vdbio::read_archive(decompressByteArray(byte_array), grids, err))
    {
      std::cerr << "byteArrayToGrid: " << err << std::endl;
      return nullptr;
    }
    return fromArchiveGrid<TGrid>(grids.front());
  }

  /*!
   * \brief Writes the active voxel centres to
   * <map_directory_path><date>_active_values_map.pcd (reference :218-256),
   * in PCL's ASCII PCD layout.
   */
  bool saveMapToPCD()
  {
    auto timestamp     = std::chrono::system_clock::now();
    std::time_t now_tt = std::chrono::system_clock::to_time_t(timestamp);
    std::tm tm         = *std::localtime(&now_tt);
    std::stringstream sstime;
    sstime << std::put_time(&tm, "%Y-%m-%d_%H-%M-%S");
    std::string pcd_path = m_map_directory_path + sstime.str() + "_active_values_map.pcd";
    std::vector<float> xyz;
    {
      std::shared_lock map_lock(*m_map_mutex);
      const int64_t n = vdbgpu_export_active_points(m_map, nullptr, 0);
      if (n > 0)
      {
        xyz.resize(3 * n);
        if (vdbgpu_export_active_points(m_map, xyz.data(), n) != n)
        {
          xyz.clear();
        }
      }
    }
    if (!pcd::writeAscii(pcd_path, xyz))
    {
      std::cerr << "Could not write PCD file." << std::endl;
      return false;
    }
    std::cout << "Wrote pcd to: " << pcd_path << std::endl;
    return true;
  }

  /*!
   * \brief Loads a PCD file into the map (reference :295-307): ASCII or binary
   * (uncompressed) files with float x, y, z fields.
   */
  bool loadMapFromPCD(const std::string& file_path, const bool set_background, const bool clear_map)
  {
    PointCloudT::Ptr cloud(new PointCloudT);
    std::vector<float> xyz;
    if (!pcd::read(file_path, xyz))
    {
      std::cerr << "Could not open PCD file" << std::endl;
      return false;
    }
    for (size_t i = 0; i < xyz.size() / 3; ++i)
    {
      cloud->points.emplace_back(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
    }
    std::unique_lock map_lock(*m_map_mutex);
    createMapFromPointCloud(cloud, set_background, clear_map);
    map_lock.unlock();
    return true;
  }

  /*!
   * \brief Builds the map from points (reference OccupancyVDBMapping.hpp:136-174):
   * every point's voxel occupied (max log-odds, active), optionally the
   * inactive voxels of the active bounding box set to min log-odds, then
   * pruneGrid: uniform leaves become 8^3 tiles (FloatGrid::forEachTile).
   */
  virtual void createMapFromPointCloud(const PointCloudT::Ptr& cloud,
                                       const bool set_background,
                                       const bool clear_map)
  {
    const float* xyz = cloud->points.empty() ? nullptr : &cloud->points[0].x;
    report(vdbgpu_create_from_points(m_map,
                                     xyz,
                                     static_cast<int64_t>(cloud->points.size()),
                                     static_cast<int64_t>(sizeof(PointT)),
                                     set_background ? 1 : 0,
                                     clear_map ? 1 : 0));
    m_vdb_grid->markRebuild();  // leaves may have become tiles
  }

  /*!
   * \brief Sets the voxel of every point free (min log-odds, inactive),
   * reference :413-427.
   */
  bool removePointsFromGrid(const PointCloudT::ConstPtr& cloud) { return editPoints(*cloud, 0); }

  /*!
   * \brief Sets the voxel of every point occupied (max log-odds, active),
   * reference :434-447.
   */
  bool addPointsToGrid(const PointCloudT::ConstPtr& cloud) { return editPoints(*cloud, 1); }

  /*!
   * \brief World box around the 8 transformed corners of [min, max]
   * (reference :811-850).
   */
  openvdb::BBoxd createWorldBoundingBox(const Eigen::Matrix<double, 3, 1>& min_boundary,
                                        const Eigen::Matrix<double, 3, 1>& max_boundary,
                                        const Eigen::Matrix<double, 4, 4>& map_to_reference_tf) const
  {
    float lo[3] = {FLT_MAX, FLT_MAX, FLT_MAX}, hi[3] = {-FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (int c = 0; c < 8; ++c)
    {
      const double p[3] = {(double)(float)((c & 4) ? max_boundary.x() : min_boundary.x()),
                           (double)(float)((c & 2) ? max_boundary.y() : min_boundary.y()),
                           (double)(float)((c & 1) ? max_boundary.z() : min_boundary.z())};
      for (int r = 0; r < 3; ++r)
      {
        const float q = (float)(map_to_reference_tf(r, 0) * p[0] + map_to_reference_tf(r, 1) * p[1] +
                                map_to_reference_tf(r, 2) * p[2] + map_to_reference_tf(r, 3));
        lo[r] = std::min(lo[r], q);
        hi[r] = std::max(hi[r], q);
      }
    }
    return openvdb::BBoxd(openvdb::Vec3d(lo[0], lo[1], lo[2]), openvdb::Vec3d(hi[0], hi[1], hi[2]));
  }

  /*!
   * \brief Index box of createWorldBoundingBox (reference :858-874).
   */
  openvdb::CoordBBox createIndexBoundingBox(const Eigen::Matrix<double, 3, 1>& min_boundary,
                                            const Eigen::Matrix<double, 3, 1>& max_boundary,
                                            const Eigen::Matrix<double, 4, 4>& map_to_reference_tf) const
  {
    const int32_t* b = indexBox(min_boundary, max_boundary, map_to_reference_tf).data();
    return openvdb::CoordBBox(openvdb::Coord(b[0], b[1], b[2]), openvdb::Coord(b[3], b[4], b[5]));
  }

  /*!
   * \brief Map section as an update grid (reference :884-894, :922-1014).
   */
  typename UpdateGridT::Ptr getMapSectionUpdateGrid(const Eigen::Matrix<double, 3, 1>& min_boundary,
                                                    const Eigen::Matrix<double, 3, 1>& max_boundary,
                                                    const Eigen::Matrix<double, 4, 4>& map_to_reference_tf,
                                                    const bool full_grid = false) const
  {
    const std::array<int32_t, 6> box = indexBox(min_boundary, max_boundary, map_to_reference_tf);
    std::vector<int32_t> o;
    std::vector<float> v;
    std::vector<uint64_t> a;
    exportSection(box, full_grid, o, v, a);
    typename UpdateGridT::Ptr grid = UpdateGridT::create(false);
    grid->setVoxelSize(m_resolution);  // the section's linear transform (reference :928)
    auto acc                       = grid->getAccessor();
    for (size_t i = 0; i < o.size() / 3; ++i)
    {
      uint64_t val[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // true where the float value is non-zero
      for (int n = 0; n < 512; ++n)
      {
        if (v[512 * i + n] != 0.0f)
        {
          val[n >> 6] |= 1ull << (n & 63);
        }
      }
      acc.mergeLeaf(&o[3 * i], &a[8 * i], val);
    }
    grid->insertMeta("bb_min", openvdb::Vec3d(box[0], box[1], box[2]));
    grid->insertMeta("bb_max", openvdb::Vec3d(box[3], box[4], box[5]));
    return grid;
  }

  /*!
   * \brief Map section as a grid (reference :902-912, :922-1014).
   */
  typename GridT::Ptr getMapSectionGrid(const Eigen::Matrix<double, 3, 1>& min_boundary,
                                        const Eigen::Matrix<double, 3, 1>& max_boundary,
                                        const Eigen::Matrix<double, 4, 4>& map_to_reference_tf,
                                        const bool full_grid = false) const
  {
    const std::array<int32_t, 6> box = indexBox(min_boundary, max_boundary, map_to_reference_tf);
    std::vector<int32_t> o;
    std::vector<float> v;
    std::vector<uint64_t> a;
    exportSection(box, full_grid, o, v, a);
    typename GridT::Ptr grid = GridT::create(m_resolution);
    for (size_t i = 0; i < o.size() / 3; ++i)
    {
      HostLeaf l;
      std::memcpy(l.origin, &o[3 * i], sizeof l.origin);
      std::memcpy(l.values, &v[512 * i], sizeof l.values);
      std::memcpy(l.mask, &a[8 * i], sizeof l.mask);
      grid->insertLeaf(l);
    }
    grid->insertMeta("bb_min", openvdb::Vec3d(box[0], box[1], box[2]));
    grid->insertMeta("bb_max", openvdb::Vec3d(box[3], box[4], box[5]));
    return grid;
  }

  /*!
   * \brief Writes a section grid into the map (reference :1022-1050). With
   * smooth_map the section is closed first (morphologicalCloseMap, which
   * changes the caller's section grid as the reference's does).
   */
  void applyMapSectionGrid(const typename GridT::Ptr section,
                           bool smooth_map          = false,
                           int smoothing_iterations = 2)
  {
    if (smooth_map)
    {
      morphologicalCloseMap<GridT>(section, smoothing_iterations);
    }
    std::vector<int32_t> o;
    std::vector<float> v;
    std::vector<uint64_t> a;
    section->forEachLeaf([&](const HostLeaf& l) {
      o.insert(o.end(), l.origin, l.origin + 3);
      v.insert(v.end(), l.values, l.values + 512);
      a.insert(a.end(), l.mask, l.mask + 8);
    });
    std::unique_lock map_lock(*m_map_mutex);
    report(vdbgpu_apply_section(m_map, (int64_t)(o.size() / 3), o.data(), v.data(), a.data()));
    m_vdb_grid->markStale();
  }

  /*!
   * \brief Replaces the active states inside the section's box by the
   * section's (reference :1058-1085), closed first with smooth_map.
   */
  void applyMapSectionUpdateGrid(const typename UpdateGridT::Ptr section,
                                 bool smooth_map          = false,
                                 int smoothing_iterations = 2)
  {
    if (smooth_map)
    {
      morphologicalCloseMap<UpdateGridT>(section, smoothing_iterations);
    }
    const openvdb::Vec3d mn = section->template metaValue<openvdb::Vec3d>("bb_min");
    const openvdb::Vec3d mx = section->template metaValue<openvdb::Vec3d>("bb_max");
    const int32_t box[6]    = {(int32_t)std::floor(mn.x()), (int32_t)std::floor(mn.y()),
                               (int32_t)std::floor(mn.z()), (int32_t)std::floor(mx.x()),
                               (int32_t)std::floor(mx.y()), (int32_t)std::floor(mx.z())};
    std::vector<int32_t> o;
    std::vector<uint64_t> a;
    section->forEachLeaf([&](const typename UpdateGridT::Leaf& l) {
      o.insert(o.end(), l.origin, l.origin + 3);
      a.insert(a.end(), l.active, l.active + 8);
    });
    std::unique_lock map_lock(*m_map_mutex);
    report(vdbgpu_apply_section_update(m_map, box, (int64_t)(o.size() / 3), o.data(), a.data()));
    m_vdb_grid->markStale();
  }

  /*!
   * \brief Morphological closing (reference :1095-1099): dilation, then
   * erosion, `iterations` 26-neighbour steps each, on the GPU.
   */
  template <typename TGrid>
  void morphologicalCloseMap(typename TGrid::Ptr grid, int iterations)
  {
    morphology(*grid, 2, iterations);
  }

  /*!
   * \brief Morphological opening (reference :1109-1113): erosion, then dilation.
   */
  template <typename TGrid>
  void morphologicalOpenMap(typename TGrid::Ptr grid, int iterations)
  {
    morphology(*grid, 3, iterations);
  }

  /*!
   * \brief dilateActiveValues with NN_FACE_EDGE_VERTEX (reference :1123-1130):
   * each step activates every voxel with an active 26-neighbour; a missing leaf
   * is created with the background value.
   */
  template <typename TGrid>
  void morphologicalDilateMap(typename TGrid::Ptr grid, int iterations)
  {
    morphology(*grid, 0, iterations);
  }

  /*!
   * \brief erodeActiveValues with NN_FACE_EDGE_VERTEX (reference :1139-1146):
   * each step keeps an active voxel only if all 26 neighbours are active;
   * leaves stay, emptied or not.
   */
  template <typename TGrid>
  void morphologicalErodeMap(typename TGrid::Ptr grid, int iterations)
  {
    morphology(*grid, 1, iterations);
  }

  /*!
   * \brief Restores the state of every artificial voxel (reference :1152-1166).
   */
  void restoreMapIntegrity()
  {
    std::unique_lock map_lock(*m_map_mutex);
    report(vdbgpu_restore_map_integrity(m_map));
    m_vdb_grid->markStale();
  }

  /*!
   * \brief Replaces the artificial areas (reference :1175-1190).
   */
  void addArtificialAreas(const std::vector<std::vector<Eigen::Matrix<double, 4, 1> > >& artificial_areas,
                          const double negative_height,
                          const double positive_height)
  {
    std::vector<int32_t> counts;
    std::vector<double> xyz;
    for (const auto& area : artificial_areas)
    {
      counts.push_back((int32_t)area.size());
      for (const auto& p : area)
      {
        xyz.insert(xyz.end(), {p.x(), p.y(), p.z()});
      }
    }
    std::unique_lock map_lock(*m_map_mutex);
    report(vdbgpu_add_artificial_areas(
      m_map, (int32_t)counts.size(), counts.data(), xyz.data(), negative_height, positive_height));
    m_vdb_grid->markStale();
  }

  /*!
   * \brief Adds the walls of one polygon (reference :1198-1207).
   */
  void addArtificialPolygon(const std::vector<Eigen::Matrix<double, 4, 1> >& polygon,
                            const double negative_height,
                            const double positive_height)
  {
    std::vector<double> walls;
    for (size_t i = 0; i < polygon.size(); i++)
    {
      const auto& a = polygon[i];
      const auto& b = polygon[(i + 1) % polygon.size()];
      walls.insert(walls.end(), {a.x(), a.y(), a.z(), b.x(), b.y(), b.z()});
    }
    report(vdbgpu_add_artificial_walls(
      m_map, (int32_t)(walls.size() / 6), walls.data(), negative_height, positive_height));
  }

  /*!
   * \brief Adds one wall (reference :1217-1236).
   */
  void addArtificialWall(const Eigen::Matrix<double, 4, 1>& start,
                         const Eigen::Matrix<double, 4, 1>& end,
                         const double negative_height,
                         const double positive_height)
  {
    const double w[6] = {start.x(), start.y(), start.z(), end.x(), end.y(), end.z()};
    report(vdbgpu_add_artificial_walls(m_map, 1, w, negative_height, positive_height));
  }

  virtual void setConfig(const TConfig& config)
  {
    if (config.max_range < 0.0)
    {
      std::cerr << "Max range of " << config.max_range << " invalid. Range cannot be negative."
                << std::endl;
      return;
    }
    m_max_range           = config.max_range;
    m_map_directory_path  = config.map_directory_path;
    m_fast_mode           = config.fast_mode;
    m_accumulation_period = config.accumulation_period;
    m_config_set          = true;
  }

  /*!
   * \brief The C-ABI handle (for callers that want the raw interface).
   */
  vdbgpu_map* handle() const { return m_map; }

protected:
  bool editPoints(const PointCloudT& cloud, int32_t occupied)
  {
    const float* xyz = cloud.points.empty() ? nullptr : &cloud.points[0].x;
    const int32_t rc = vdbgpu_set_points(
      m_map, xyz, static_cast<int64_t>(cloud.points.size()), static_cast<int64_t>(sizeof(PointT)), occupied);
    m_vdb_grid->markStale();
    return report(rc) == VDBGPU_OK;
  }

  std::array<int32_t, 6> indexBox(const Eigen::Matrix<double, 3, 1>& min_boundary,
                                  const Eigen::Matrix<double, 3, 1>& max_boundary,
                                  const Eigen::Matrix<double, 4, 4>& tf) const
  {
    const double mn[3] = {min_boundary.x(), min_boundary.y(), min_boundary.z()};
    const double mx[3] = {max_boundary.x(), max_boundary.y(), max_boundary.z()};
    double t[16];
    for (int r = 0; r < 4; ++r)
    {
      for (int c = 0; c < 4; ++c)
      {
        t[4 * r + c] = tf(r, c);
      }
    }
    std::array<int32_t, 6> out{};
    report(vdbgpu_index_bbox(m_resolution, mn, mx, t, out.data()));
    return out;
  }

  static vdbio::Grid toArchiveGrid(const FloatGrid& grid)
  {
    vdbio::Grid g;
    g.is_float   = true;
    g.grid_class = "level set";
    g.voxel_size = grid.voxelSize();
    grid.forEachLeaf([&](const HostLeaf& l) {
      vdbio::Leaf a;
      std::memcpy(a.origin, l.origin, sizeof a.origin);
      std::memcpy(a.active, l.mask, sizeof a.active);
      std::memcpy(a.values, l.values, sizeof a.values);
      std::memset(a.bits, 0, sizeof a.bits);
      g.leaves.push_back(a);
    });
    grid.forEachTile([&](const HostTile& t) {
      vdbio::Tile a;
      std::memcpy(a.origin, t.origin, sizeof a.origin);
      a.value  = t.value;
      a.active = t.active;
      g.tiles.push_back(a);
    });
    for (const char* k : {"bb_min", "bb_max"})
    {
      if (grid.hasMeta(k))
      {
        const openvdb::Vec3d v = grid.template metaValue<openvdb::Vec3d>(k);
        g.vec3d_meta[k]        = {v.x(), v.y(), v.z()};
      }
    }
    return g;
  }
  static vdbio::Grid toArchiveGrid(const BoolGrid& grid)
  {
    vdbio::Grid g;
    g.is_float   = false;
    g.voxel_size = grid.voxelSize();
    grid.forEachLeaf([&](const BoolGrid::Leaf& l) {
      vdbio::Leaf a;
      std::memcpy(a.origin, l.origin, sizeof a.origin);
      std::memcpy(a.active, l.active, sizeof a.active);
      std::memcpy(a.bits, l.value, sizeof a.bits);
      std::memset(a.values, 0, sizeof a.values);
      g.leaves.push_back(a);
    });
    for (const char* k : {"bb_min", "bb_max"})
    {
      if (grid.hasMeta(k))
      {
        const openvdb::Vec3d v = grid.template metaValue<openvdb::Vec3d>(k);
        g.vec3d_meta[k]        = {v.x(), v.y(), v.z()};
      }
    }
    return g;
  }
  template <typename TGrid>
  static typename TGrid::Ptr fromArchiveGrid(const vdbio::Grid& g)
  {
    if (g.is_float != std::is_same<TGrid, FloatGrid>::value)
    {
      return nullptr;  // the reference's gridPtrCast gives null on a type mismatch
    }
    typename TGrid::Ptr out;
    if constexpr (std::is_same<TGrid, FloatGrid>::value)
    {
      out = TGrid::create(g.voxel_size);
    }
    else
    {
      out = TGrid::create(false);
      out->setVoxelSize(g.voxel_size);
    }
    for (const auto& kv : g.vec3d_meta)
    {
      out->insertMeta(kv.first, openvdb::Vec3d(kv.second[0], kv.second[1], kv.second[2]));
    }
    if constexpr (std::is_same<TGrid, FloatGrid>::value)
    {
      for (const auto& l : g.leaves)
      {
        HostLeaf h;
        std::memcpy(h.origin, l.origin, sizeof h.origin);
        std::memcpy(h.values, l.values, sizeof h.values);
        std::memcpy(h.mask, l.active, sizeof h.mask);
        out->insertLeaf(h);
      }
    }
    else
    {
      auto acc = out->getAccessor();
      for (const auto& l : g.leaves)
      {
        acc.mergeLeaf(l.origin, l.active, l.bits);
      }
    }
    return out;
  }

  // vdbgpu_morphology over the leaves of a detached grid (a section); the
  // new masks are written back, created leaves take the background value.
  template <typename TGridObj>
  void morphology(TGridObj& grid, int32_t op, int iterations)
  {
    if (grid.attached())
    {
      std::cerr << "vdbgpu: morphology of the live map grid is not supported" << std::endl;
      return;
    }
    std::vector<int32_t> o;
    std::vector<uint64_t> a;
    grid.forEachLeaf([&](const auto& l) {
      o.insert(o.end(), l.origin, l.origin + 3);
      const uint64_t* w = leafActiveWords(l);
      a.insert(a.end(), w, w + 8);
    });
    const int64_t n = static_cast<int64_t>(o.size() / 3);
    const int64_t k = vdbgpu_morphology(m_map, op, iterations, n, o.data(), a.data(), nullptr, nullptr, 0);
    std::vector<int32_t> ro(3 * std::max<int64_t>(k, 0));
    std::vector<uint64_t> rm(8 * std::max<int64_t>(k, 0));
    if (k < 0 || vdbgpu_morphology(m_map, op, iterations, n, o.data(), a.data(), ro.data(), rm.data(), k) != k)
    {
      std::cerr << "vdbgpu: morphology failed: " << vdbgpu_last_error(m_map) << std::endl;
      return;
    }
    for (int64_t i = 0; i < k; ++i)
    {
      grid.setLeafMask(&ro[3 * i], &rm[8 * i]);
    }
  }

  void exportSection(const std::array<int32_t, 6>& box,
                     bool full_grid,
                     std::vector<int32_t>& o,
                     std::vector<float>& v,
                     std::vector<uint64_t>& a) const
  {
    std::shared_lock map_lock(*m_map_mutex);
    const int64_t n = vdbgpu_export_section(m_map, box.data(), full_grid ? 1 : 0, nullptr, nullptr, nullptr, 0);
    if (n <= 0)
    {
      return;
    }
    o.resize(3 * n);
    v.resize(512 * n);
    a.resize(8 * n);
    const int64_t got = vdbgpu_export_section(m_map, box.data(), full_grid ? 1 : 0, o.data(), v.data(), a.data(), n);
    if (got != n)
    {
      o.clear();
      v.clear();
      a.clear();
    }
  }

  int32_t accumulate(const std::string& source_id,
                     const PointCloudT& cloud,
                     const Eigen::Matrix<double, 3, 1>& origin)
  {
    const double o[3] = {origin.x(), origin.y(), origin.z()};
    const float* xyz  = cloud.points.empty() ? nullptr : &cloud.points[0].x;
    return vdbgpu_accumulate(m_map,
                             source_id.c_str(),
                             xyz,
                             static_cast<int64_t>(cloud.points.size()),
                             static_cast<int64_t>(sizeof(PointT)),
                             o,
                             0);
  }

  int32_t report(int32_t rc) const
  {
    if (rc != VDBGPU_OK && rc != VDBGPU_UNKNOWN_SOURCE && rc != VDBGPU_NOT_CONFIGURED)
    {
      std::cerr << "vdbgpu: " << vdbgpu_status_string(rc) << ": " << vdbgpu_last_error(m_map)
                << std::endl;
    }
    return rc;
  }

  vdbgpu_map* m_map = nullptr;
  typename GridT::Ptr m_vdb_grid;
  double m_max_range = 0;
  double m_resolution;
  bool m_fast_mode = false;
  std::string m_map_directory_path;
  unsigned int m_compression_level = 1;  // zstd level of gridToByteArray (reference :1524)
  bool m_config_set;
  std::map<std::string, double> m_sources;
  std::map<std::string, double> m_rates;
  double m_accumulation_period = 0;
  mutable std::shared_ptr<std::shared_mutex> m_map_mutex;
};

} // namespace vdb_mapping

#endif /* VDB_MAPPING_VDB_MAPPING_H_INCLUDED */
