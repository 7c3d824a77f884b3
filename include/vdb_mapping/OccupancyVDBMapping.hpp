// Drop-in replacement of vdb_mapping::OccupancyVDBMapping (reference
// include/vdb_mapping/OccupancyVDBMapping.hpp): the log-odds specialisation.
// setConfig keeps the reference's checks, messages and partial application
// (OccupancyVDBMapping.hpp:59-89); the per-voxel update rule
// (updateFreeNode / updateOccupiedNode, :92-117) runs in the gfx950 apply
// kernel with the constants computed here on the host in double and rounded
// to float, exactly as the reference does.
#ifndef VDB_MAPPING_OCCUPANCY_VDB_MAPPING_H_INCLUDED
#define VDB_MAPPING_OCCUPANCY_VDB_MAPPING_H_INCLUDED

#include <array>

#include "vdb_mapping/VDBMapping.hpp"

namespace vdb_mapping {

/*!
 * \brief Accumulation of configuration parameters (reference :38-44)
 */
struct Config : BaseConfig
{
  double prob_hit;
  double prob_miss;
  double prob_thres_min;
  double prob_thres_max;
};

class OccupancyVDBMapping : public VDBMapping<float, Config>
{
public:
  OccupancyVDBMapping(const double resolution)
    : VDBMapping<float, Config>(resolution)
  {
  }

  inline void setConfig(const Config& config) override
  {
    VDBMapping::setConfig(config);

    vdbgpu_config c;
    std::memset(&c, 0, sizeof c);
    c.max_range           = config.max_range;
    c.fast_mode           = config.fast_mode ? 1 : 0;
    c.accumulation_period = config.accumulation_period;
    c.prob_hit            = config.prob_hit;
    c.prob_miss           = config.prob_miss;
    c.prob_thres_min      = config.prob_thres_min;
    c.prob_thres_max      = config.prob_thres_max;
    vdbgpu_set_config(m_map, &c);

    if (config.prob_miss > 0.5)
    {
      std::cerr << "Probability for a miss should be below 0.5 but is " << config.prob_miss
                << std::endl;
      return;
    }
    if (config.prob_hit < 0.5)
    {
      std::cerr << "Probability for a hit should be above 0.5 but is " << config.prob_hit
                << std::endl;
      return;
    }
    m_config_set = true;
  }

  /*!
   * \brief Log-odds constants in force: hit, miss, thres_min, thres_max,
   * clamp max, clamp min.
   */
  std::array<float, 6> logodds() const
  {
    std::array<float, 6> out{};
    vdbgpu_logodds(m_map, out.data());
    return out;
  }
};

} // namespace vdb_mapping

#endif /* VDB_MAPPING_OCCUPANCY_VDB_MAPPING_H_INCLUDED */
