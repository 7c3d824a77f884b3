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
// Not provided (each has no call on this path): persistence and PCD I/O (:193-307), map sections,
// byte arrays and morphology (:810-1338), direct point edits and artificial
// areas (:413-447, :1152-1236), castRayIntoGrid (:550-566, per-ray host call)
// and updateMap (:731-792, reached through integrateUpdate).
#ifndef VDB_MAPPING_VDB_MAPPING_H_INCLUDED
#define VDB_MAPPING_VDB_MAPPING_H_INCLUDED

#include <cmath>
#include <cstring>
#include <iostream>
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
  bool m_config_set;
  std::map<std::string, double> m_sources;
  std::map<std::string, double> m_rates;
  double m_accumulation_period = 0;
  mutable std::shared_ptr<std::shared_mutex> m_map_mutex;
};

} // namespace vdb_mapping

#endif /* VDB_MAPPING_VDB_MAPPING_H_INCLUDED */
