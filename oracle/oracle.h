/*
 * CPU restatement of vdb_mapping's occupancy-integration path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity checker: tests/, the
 * smoke check and the CPU-baseline leg of bench.py may load it. The product
 * (libvdbgpu.so, the vdb_mapping_amd package, the drop-in headers under include/vdb_mapping)
 * never links, loads or calls it.
 *
 * Each function cites the reference code it restates
 * (include/vdb_mapping/VDBMapping.hpp = VM, OccupancyVDBMapping.hpp = OCC).
 * The voxel arithmetic of the reference lives in OpenVDB, which is not part
 * of the reference tree; the restated OpenVDB semantics are listed in
 * oracle.cpp next to the code that uses them.
 *
 * Pinning (tests/test_oracle.py): the reference's known-answer tests, closed
 * forms, an exact-rational DDA walk and brute-force set checks. raytrace and
 * the fast-mode cast are checked against an exact-rational flat leaf-cell
 * march (tests/test_oracle_query_exact.py) on random maps with inactive leaves
 * inside and outside the active box, which also catches planted mistakes
 * (clip skipped, first voxel tested, box pad off by one either way, a later
 * run, spans dropped). Parity unpinned (no check here can compare them with
 * OpenVDB itself): the rounding forms U1-U3, the double tie-break between axes
 * of different extent (U5), the OpenVDB semantics restated for raytrace (U6,
 * U7, U9, U10) and the fast-mode hits walk (U11), which the brute force
 * restates in the same terms, leaf creation by no-op updates (U8), the section-apply visit of
 * internal-node tiles (U12), the bounding-box corner transform (U13) and the
 * leaf set of the morphology operators (U14; their voxel states are pinned
 * against scipy.ndimage's binary dilation and erosion).
 */
#ifndef VDB_ORACLE_H
#define VDB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors BaseConfig (VM:64-71) + Config (OCC:38-44). */
typedef struct orc_config {
  double max_range;
  int32_t fast_mode;
  int32_t _pad;
  double accumulation_period;
  double prob_hit;
  double prob_miss;
  double prob_thres_min;
  double prob_thres_max;
} orc_config;

/* Status codes shared with the GPU library (include/vdbgpu.h). */
enum {
  ORC_OK = 0,
  ORC_NOT_CONFIGURED = 1,
  ORC_UNKNOWN_SOURCE = 2,
  ORC_BAD_MAX_RANGE = 3,
  ORC_BAD_PROB_MISS = 4,
  ORC_BAD_PROB_HIT = 5,
  ORC_OUT_OF_RANGE = 6,
  ORC_UNSUPPORTED = 7,
  ORC_BAD_ARGUMENT = 9,
  ORC_NO_INTERSECTOR = 10
};

void* orc_create(double resolution);
void orc_destroy(void* h);
int orc_set_config(void* h, const orc_config* cfg);
int orc_add_source(void* h, const char* source_id, double max_range, double max_rate);
int orc_accumulate(void* h, const char* source_id, const float* xyz, int64_t n,
                   int64_t stride_bytes, const double origin[3]);
int orc_integrate(void* h);
int orc_insert(void* h, const char* source_id, const float* xyz, int64_t n,
               int64_t stride_bytes, const double origin[3]);
void orc_reset(void* h);
int orc_empty(void* h);
/* Log-odds constants as set by setConfig: hit, miss, thres_min, thres_max, max, min. */
void orc_logodds(void* h, float out[6]);

/* Map export: leaves sorted by origin (x, then y, then z). */
int64_t orc_leaf_count(void* h);
int64_t orc_export_leaves(void* h, int32_t* origins, float* values, uint64_t* active_masks,
                          int64_t cap);
/* The leaf-level tiles pruneGrid made (createMapFromPointCloud, OCC:173), sorted by
 * origin: origin, value, state. Tiles are not leaves: orc_leaf_count and
 * orc_export_leaves leave them out. */
int64_t orc_export_tiles(void* h, int32_t* origins, float* values, uint8_t* active, int64_t cap);
/* Pending update set of one source (VM:89 UpdateGridT): visited and hit masks per leaf. */
int64_t orc_update_leaf_count(void* h, const char* source_id);
int64_t orc_export_update(void* h, const char* source_id, int32_t* origins, uint64_t* visited,
                          uint64_t* hit, int64_t cap);

/* Single-voxel read, like an OpenVDB accessor: returns value, writes active flag. */
float orc_get_value(void* h, const int32_t ijk[3], int32_t* active);

/* Primitives, exposed so tests can pin them on their own. */
void orc_world_to_index(double resolution, const double world[3], int32_t out[3]);
int64_t orc_cast_ray(const int32_t s[3], const int32_t e[3], int32_t* out_xyz, int64_t cap);

/* castRayIntoGridFast (VM:577-602): the map-active voxels a fast-mode cast from
 * voxel s to voxel e marks, in visiting order; needs the intersector (else -1). */
int64_t orc_cast_ray_fast(void* h, const int32_t s[3], const int32_t e[3], int32_t* out_xyz,
                          int64_t cap);

/* Batch raytrace (VM:675-721). Requires the intersector state the reference
 * builds in integrateUpdate when fast_mode is set (VM:1436-1449). */
int orc_raytrace(void* h, int64_t n, const double* origins, const double* dirs,
                 const double* max_lengths, uint8_t* success, double* end_points);

/* addPointsToGrid (occupied = 1) / removePointsFromGrid (occupied = 0), VM:413-447. */
int orc_set_points(void* h, const float* xyz, int64_t n, int64_t stride_bytes, int32_t occupied);

/* addArtificialAreas (VM:1175-1236): polygons of counts[p] (x, y, z) vertices. */
int orc_add_artificial_areas(void* h, int32_t n_polygons, const int32_t* counts,
                             const double* xyz, double negative_height, double positive_height);
/* addArtificialWall (VM:1217-1236) for n walls (start xyz, end xyz each), no restore. */
int orc_add_artificial_walls(void* h, int32_t n_walls, const double* xyz, double negative_height,
                             double positive_height);
/* restoreMapIntegrity (VM:1152-1166). */
int orc_restore_map_integrity(void* h);
/* The artificial grid: origins and voxel masks per leaf, sorted by origin. */
int64_t orc_export_artificial(void* h, int32_t* origins, uint64_t* masks, int64_t cap);

/* getMapSection (VM:922-1014) over the inclusive index box bbox = (min xyz, max xyz). */
int64_t orc_export_section(void* h, const int32_t bbox[6], int32_t full_grid, int32_t* origins,
                           float* values, uint64_t* masks, int64_t cap);
/* applyMapSectionGrid without smoothing (VM:1022-1050). */
int orc_apply_section(void* h, int64_t n, const int32_t* origins, const float* values,
                      const uint64_t* masks);
/* applyMapSectionUpdateGrid without smoothing (VM:1058-1085). */
int orc_apply_section_update(void* h, const int32_t bbox[6], int64_t n, const int32_t* origins,
                             const uint64_t* masks);
/* createIndexBoundingBox (VM:811-874); tf is a row-major 4x4 matrix. */
void orc_index_bbox(double resolution, const double min_boundary[3],
                    const double max_boundary[3], const double tf[16], int32_t out[6]);

/* saveMapToPCD's point list (VM:218-256): active voxel centres in tree order, as float xyz. */
int64_t orc_export_active_points(void* h, float* xyz, int64_t cap);
/* createMapFromPointCloud (OCC:136-174), pruneGrid at the leaf level included. */
int orc_create_from_points(void* h, const float* xyz, int64_t n, int64_t stride_bytes,
                           int32_t set_background, int32_t clear_map);

/* The morphology members of VM:1095-1148 on a leaf set (U14): op 0 dilate,
 * 1 erode, 2 close (dilate, then erode), 3 open (erode, then dilate), each by
 * `iterations` 26-neighbour steps. Output: the n input leaves in input order
 * with their new masks, then the leaves dilation created, sorted by origin.
 * Null outputs return the count, a short cap -count, bad input -1. */
int64_t orc_morphology(int32_t op, int32_t iterations, int64_t n, const int32_t* origins,
                       const uint64_t* masks, int32_t* out_origins, uint64_t* out_masks,
                       int64_t cap);

#ifdef __cplusplus
}
#endif

#endif
