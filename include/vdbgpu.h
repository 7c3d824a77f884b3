/*
 * vdbgpu: C ABI of the MI355X (gfx950) occupancy-integration path of
 * vdb_mapping. Plain pointers and sizes only; no C++ or HIP types cross this
 * boundary. Every call on one map handle is serialised by the handle's own
 * mutex and runs on the handle's own HIP stream; different handles are
 * independent and may be used from different threads.
 *
 * Reference interface each entry point replaces (file:line in the
 * vdb_mapping repository; VM = include/vdb_mapping/VDBMapping.hpp,
 * OCC = include/vdb_mapping/OccupancyVDBMapping.hpp) is cited per function.
 * The drop-in C++ classes in include/vdb_mapping/ call only these functions.
 */
#ifndef VDBGPU_H
#define VDBGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 0 = success. */
enum {
  VDBGPU_OK = 0,
  VDBGPU_NOT_CONFIGURED = 1, /* VM:478-482: raycast before setConfig */
  VDBGPU_UNKNOWN_SOURCE = 2, /* VM:320-326: source id never added */
  VDBGPU_BAD_MAX_RANGE = 3,  /* VM:1458-1463: max_range < 0 (rest of setConfig still applied) */
  VDBGPU_BAD_PROB_MISS = 4,  /* OCC:64-69: prob_miss > 0.5 */
  VDBGPU_BAD_PROB_HIT = 5,   /* OCC:70-75: prob_hit < 0.5 */
  VDBGPU_OUT_OF_RANGE = 6,   /* a voxel index outside [-2^23, 2^23) or non-finite: scan rejected */
  VDBGPU_UNSUPPORTED = 7,    /* a ray longer than 8192 voxels along one axis */
  VDBGPU_DEVICE_ERROR = 8,   /* HIP runtime error or out of device memory */
  VDBGPU_BAD_ARGUMENT = 9,
  VDBGPU_NO_INTERSECTOR = 10, /* raytrace, or a fast-mode cast into a non-empty map, before the
                               * source has an intersector (VM:590, :698, :1438); the reference
                               * dereferences a null pointer there */
  VDBGPU_OVER_BUDGET = 11     /* serving: the maps' device footprint exceeds the HBM budget */
};

/* BaseConfig (VM:64-71) + Config (OCC:38-44), field for field.
 * map_directory_path is host-only state kept by the C++ wrapper. */
typedef struct vdbgpu_config {
  double max_range;
  int32_t fast_mode;
  int32_t _pad;
  double accumulation_period;
  double prob_hit;
  double prob_miss;
  double prob_thres_min;
  double prob_thres_max;
} vdbgpu_config;

/* Cumulative counters of one map handle. Kernel times are measured with HIP
 * events on the handle's stream when profiling is enabled. */
typedef struct vdbgpu_stats {
  uint64_t scans;          /* accumulate calls that cast rays */
  uint64_t rays;           /* points cast (NaN points excluded) */
  uint64_t visits;         /* DDA voxel visits, sum of 1 + |dx|+|dy|+|dz| */
  uint64_t voxel_updates;  /* unique update-set voxels applied to the map */
  uint64_t update_leaves;  /* update-set leaves applied */
  uint64_t leaf_count;     /* leaves in the map now */
  uint64_t device_bytes;   /* device memory held by the handle */
  /* summed kernel times: prep (points -> rays), traverse (hash traversal, or the brick cut),
   * compact (update-leaf listing and map-leaf lookup/creation), apply (log-odds update),
   * raytrace; ms_walk below: the brick walk */
  double ms_prep, ms_traverse, ms_compact, ms_apply, ms_raytrace;
  uint64_t n_prep, n_traverse, n_compact, n_apply, n_raytrace;      /* launches timed */
  double ms_walk;          /* brick path: the brick walk kernel (ms_traverse: the brick cut) */
  uint64_t n_walk;
  uint64_t scratch_bytes;  /* part of device_bytes that vdbgpu_release_scratch would free */
} vdbgpu_stats;

/* The leaf pool of a map grows in place in address ranges reserved at its first leaf (HIP
 * virtual memory); with the environment variable VDBGPU_PLAIN_POOL=1 set when a map gets its
 * first leaf, or where the driver refuses a mapping, it uses plain allocations that grow by copy.
 * Results are the same either way. */

/* Launch parameters of the integration kernels (see vdbgpu_set_tuning). */
typedef struct vdbgpu_tuning {
  int32_t block;          /* threads per workgroup of the traversal kernel (64, 128, 192, 256) */
  int32_t steps_per_lane; /* DDA visits per lane work item (chunk length) */
  int32_t waves_per_cu;   /* resident waves per CU requested for the traversal grid */
  double load_factor;     /* update-set hash target load (0.3..0.9) */
  int32_t traversal;      /* plain inserts: 0 = brick walk (default), 1 = update-set hash traversal */
  int32_t defer;          /* brick-walk inserts: the map update of up to this many consecutive scans (1..8,
                           * default 8) runs as one pass that reads and writes each leaf once and applies
                           * the scans in insert order; any other call on the map (reads, edits, syncs)
                           * applies the waiting scans first. 0 or 1: every scan is applied on its own. */
} vdbgpu_tuning;

typedef struct vdbgpu_map vdbgpu_map;

/* Library version (major*10000 + minor*100 + patch). */
int32_t vdbgpu_version(void);
/* Human-readable name of a status code. */
const char* vdbgpu_status_string(int32_t status);
/* Number of gfx950 devices visible; 0 when no GPU is present. */
int32_t vdbgpu_device_count(void);

/* OccupancyVDBMapping(resolution) OCC:49-52 -> VDBMapping(resolution)
 * VM:115-134 (createVDBMap VM:163-169): empty map of 8^3 float leaves,
 * background 0, voxel size `resolution`. Uses the current HIP device. */
int32_t vdbgpu_create(double resolution, vdbgpu_map** out);
/* ~VDBMapping VM:139-154: frees every device allocation of the handle. */
void vdbgpu_destroy(vdbgpu_map* map);

/* OccupancyVDBMapping::setConfig OCC:59-89 calling VDBMapping::setConfig
 * VM:1456-1469, including the reference's partial-application behaviour. */
int32_t vdbgpu_set_config(vdbgpu_map* map, const vdbgpu_config* config);
/* Log-odds constants now in force: hit, miss, thres_min, thres_max, clamp max,
 * clamp min (OCC:79-87). */
int32_t vdbgpu_logodds(vdbgpu_map* map, float out[6]);

/* addInputSource VM:1352-1375: max_range 0 selects the config's max_range.
 * max_rate only matters to the reference's background threads. */
int32_t vdbgpu_add_source(vdbgpu_map* map, const char* source_id, double max_range,
                          double max_rate);

/* Remove a source and its pending update set. No reference counterpart (the
 * reference never removes sources); used by the drop-in raycastPointCloud to
 * cast into a private scratch source. */
int32_t vdbgpu_remove_source(vdbgpu_map* map, const char* source_id);

/* The source lookup of addDataToAccumulate (VM:359-364): VDBGPU_OK if the
 * source was added, VDBGPU_UNKNOWN_SOURCE if not. Has no other effect. */
int32_t vdbgpu_has_source(vdbgpu_map* map, const char* source_id);

/* accumulateUpdate VM:316-346 -> raycastPointCloud VM:466-539: cast every
 * point of the cloud from `origin` into the source's update set. `xyz` holds
 * n points of 3 float32 at `stride_bytes` (16 for pcl::PointXYZ); it is a host
 * pointer, or a device pointer when `points_on_device` is non-zero. */
int32_t vdbgpu_accumulate(vdbgpu_map* map, const char* source_id, const float* xyz, int64_t n,
                          int64_t stride_bytes, const double origin[3], int32_t points_on_device);
/* integrateUpdate VM:375-387 -> updateMap VM:731-792 with updateOccupiedNode /
 * updateFreeNode OCC:92-117: apply and clear every source's update set, in
 * source-id order. */
int32_t vdbgpu_integrate(vdbgpu_map* map);
/* insertPointCloud VM:399-406: accumulate then integrate. */
int32_t vdbgpu_insert(vdbgpu_map* map, const char* source_id, const float* xyz, int64_t n,
                      int64_t stride_bytes, const double origin[3], int32_t points_on_device);

/* Batched insertPointCloud: nscans clouds of one source, concatenated in
 * `xyz` (counts[i] points each, same stride), integrated with a result
 * identical to vdbgpu_insert on each scan in order (VM:399-406 per scan).
 * origins: nscans x 3 doubles. max_ranges: per-scan max range, or NULL for
 * the source's range (a range <= 0 casts nothing for that scan). With
 * tuning.traversal = 0 (the default) the scans stream through the brick path,
 * each scan's cast enqueued before the previous scan's status is read (with
 * the environment variable VDBGPU_GRAPH=1 the casts are replayed from captured
 * HIP graphs, one per set of buffers); with traversal = 1 up to 128 scans
 * share one (leaf, scan)-keyed update set and one scan-ordered fold. No
 * reference counterpart for the per-scan ranges. */
int32_t vdbgpu_insert_batch(vdbgpu_map* map, const char* source_id, int32_t nscans,
                            const float* xyz, const int64_t* counts, int64_t stride_bytes,
                            const double* origins, const double* max_ranges,
                            int32_t points_on_device);

/* resetMap VM:174-186: empty map and empty update sets. */
int32_t vdbgpu_reset(vdbgpu_map* map);
/* GridT::empty() as used at VM:494 and VM:1438: 1 if no voxel was ever modified, 0 if one
 * was, a status code (> 1) on error. */
int32_t vdbgpu_empty(vdbgpu_map* map);

/* getGrid VM:799, as a leaf dump in openvdb::FloatTree leaf layout: for leaf
 * i, origins[3i..3i+2] (multiples of 8), values[512i..] float32 indexed by
 * ((x&7)<<6)|((y&7)<<3)|(z&7), active_masks[8i..] (bit n&63 of word n>>6).
 * Leaves are sorted by origin (x, then y, then z). Returns the leaf count, or
 * minus the needed count when cap is too small (nothing written). Any pointer
 * may be NULL to query the count only. */
int64_t vdbgpu_leaf_count(vdbgpu_map* map);
int64_t vdbgpu_export_leaves(vdbgpu_map* map, int32_t* origins, float* values,
                             uint64_t* active_masks, int64_t cap);
/* Leaves changed by the integrations since the previous call of this
 * function (same layout; used to keep a host grid live without a full dump). */
int64_t vdbgpu_export_dirty_leaves(vdbgpu_map* map, int32_t* origins, float* values,
                                   uint64_t* active_masks, int64_t cap);

/* The source's pending update set (UpdateGridT VM:89, filled by
 * raycastPointCloud VM:466-539): per leaf, visited (= active) and hit
 * (= value true) masks in the same bit layout, sorted by origin. */
int64_t vdbgpu_update_leaf_count(vdbgpu_map* map, const char* source_id);
int64_t vdbgpu_export_update(vdbgpu_map* map, const char* source_id, int32_t* origins,
                             uint64_t* visited, uint64_t* hit, int64_t cap);

/* raytrace VM:675-721 (batch; the single-ray form VM:643-666 is n = 1):
 * origins, directions (n x 3 double, world frame), max lengths (n); writes
 * success[n] and end_points (n x 3 double). Uses the topology snapshot taken
 * by the last integrate in fast mode (updateVolumeRayIntersectors
 * VM:1436-1449); VDBGPU_NO_INTERSECTOR when there is none. Host pointers. */
int32_t vdbgpu_raytrace(vdbgpu_map* map, int64_t n, const double* origins,
                        const double* directions, const double* max_lengths, uint8_t* success,
                        double* end_points);

/* Device buffers for callers that keep point clouds resident in HBM (the
 * points_on_device form of vdbgpu_accumulate / vdbgpu_insert). No reference
 * counterpart: the reference takes host pcl clouds only (VM:399). Allocation
 * is on the current HIP device; copies are synchronous. */
int32_t vdbgpu_device_alloc(int64_t bytes, void** out);
int32_t vdbgpu_device_free(void* ptr);
int32_t vdbgpu_device_upload(void* dst, const void* src, int64_t bytes);
int32_t vdbgpu_device_download(void* dst, const void* src, int64_t bytes);
/* Device-to-device copy (the plain copy kernel the traffic counters are
 * calibrated against); returns after the copy has completed. */
int32_t vdbgpu_device_copy(void* dst, const void* src, int64_t bytes);

/* raytrace with every array in device memory (same layout). */
int32_t vdbgpu_raytrace_device(vdbgpu_map* map, int64_t n, const double* origins,
                               const double* directions, const double* max_lengths,
                               uint8_t* success, double* end_points);

/* raytrace (same arguments and results as vdbgpu_raytrace, host pointers) that
 * allocates nothing: no hipMalloc / hipFree, no growth of the handle's buffers.
 * No counterpart in the reference; for callers that cannot allocate (code under
 * HIP graph capture, tight serving or planning loops, harnesses that forbid
 * allocation). Queries, results and the rejected-query counter are staged in a
 * static 32 MiB device array of the library, one per GPU, held until the
 * process ends. Limits: at most 393216 queries per call (VDBGPU_BAD_ARGUMENT
 * beyond); map updates still waiting to be applied (deferred brick-path scans)
 * are refused with VDBGPU_BAD_ARGUMENT instead of applied, since applying them
 * may allocate: call vdbgpu_synchronize first. Calls on one GPU share its
 * array: the library serialises them within a process; calls from different
 * processes on one GPU use different arrays. */
int32_t vdbgpu_raytrace_noalloc(vdbgpu_map* map, int64_t n, const double* origins,
                                const double* directions, const double* max_lengths,
                                uint8_t* success, double* end_points);

/* Free the handle's per-scan working memory (point staging, per-ray arrays,
 * brick tables, segments, records, record sets, fold and group tables, query
 * staging, idle update sets) after applying any waiting scans; the map, the
 * sources' pending updates, the raytrace snapshot and the artificial grid are
 * kept, and the next call that needs a buffer allocates it again. No
 * reference counterpart: the serving loop (vdbgpu_server_*) calls it to hand
 * device memory to other maps while requests wait for HBM. */
int32_t vdbgpu_release_scratch(vdbgpu_map* map);

/* Wait for all work of the handle. */
int32_t vdbgpu_synchronize(vdbgpu_map* map);
/* Counters; profiling adds HIP-event timing of every kernel launch. */
int32_t vdbgpu_get_stats(vdbgpu_map* map, vdbgpu_stats* out);
int32_t vdbgpu_set_profiling(vdbgpu_map* map, int32_t enabled);
int32_t vdbgpu_set_tuning(vdbgpu_map* map, const vdbgpu_tuning* tuning);
int32_t vdbgpu_get_tuning(vdbgpu_map* map, vdbgpu_tuning* out);
/* Text of the last error on this handle ("" when none). */
const char* vdbgpu_last_error(vdbgpu_map* map);

/* ---- Serving: many robot streams, one device map each ------------------
 * No reference counterpart (the reference serves one map per process through
 * its accumulation threads, VM:355-370, :1383-1430, which merge and drop
 * scans). Here every submitted scan is integrated exactly once, in arrival
 * order per map: a worker per map takes all requests queued while its map was
 * busy and integrates them in one vdbgpu_insert_batch (same result as one by
 * one); different maps run on their own streams and overlap on the device.
 * HBM admission (hbm_budget_bytes, 0 = no budget): a batch runs when the maps'
 * device memory, the reservations of the batches in flight and its own
 * projected need fit the budget (per-scan working memory per point of its
 * largest scan plus map growth per point, from what finished batches needed;
 * 2 KiB and 1 KiB to start); otherwise it waits, and lanes give back their
 * per-scan working memory (vdbgpu_release_scratch) as batches finish. When
 * nothing is in flight and nothing is left to free, it runs alone if the maps
 * are within the budget; its requests fail with VDBGPU_OVER_BUDGET only when
 * the maps alone fill the budget (maps never shrink; reset or destroy maps and
 * resubmit). */
typedef struct vdbgpu_server vdbgpu_server;

typedef struct vdbgpu_request_record {
  int64_t request_id;
  int64_t points;
  int32_t map_index;
  int32_t batch_size; /* requests integrated together with this one */
  int32_t status;
  int32_t mem_waits;  /* times its batch waited for device memory before admission */
  double latency_ms;  /* submit -> device work finished */
} vdbgpu_request_record;

int32_t vdbgpu_server_create(int32_t n_maps, double resolution, const vdbgpu_config* config,
                             double max_range, int64_t hbm_budget_bytes, int32_t max_batch,
                             vdbgpu_server** out);
/* Copies the cloud (host memory) and queues it for map `map_index`. */
int32_t vdbgpu_server_submit(vdbgpu_server* server, int32_t map_index, const float* xyz, int64_t n,
                             int64_t stride_bytes, const double origin[3], int64_t* request_id);
/* Wait until every submitted request has finished; returns the first error. */
int32_t vdbgpu_server_drain(vdbgpu_server* server);
/* Finished requests in completion order (NULL out = count only). */
int64_t vdbgpu_server_records(vdbgpu_server* server, vdbgpu_request_record* out, int64_t cap);
/* The map of one stream (owned by the server; valid until destroy). */
vdbgpu_map* vdbgpu_server_map(vdbgpu_server* server, int32_t map_index);
int64_t vdbgpu_server_device_bytes(vdbgpu_server* server);
void vdbgpu_server_destroy(vdbgpu_server* server);

/* ---- asynchronous ingestion (opt-in) ----------------------------------------
 * The reference's addDataToAccumulate (VM:355-370), per-source accumulation
 * threads and integration thread (VM:133, :1383-1430) as an engine per map:
 * each source has a latest-only slot (a new cloud replaces one not yet taken)
 * and a worker that accumulates it and then waits out 1/max_rate; one thread
 * integrates every accumulation_period seconds. Scans accumulated within a
 * period OR-merge and replaced scans are dropped, exactly the reference's
 * semantics; for per-scan exactness use vdbgpu_insert or the server. Not
 * started until vdbgpu_async_start (the reference starts its threads in the
 * constructor and addInputSource). Calls must not race vdbgpu_async_stop. */
int32_t vdbgpu_async_start(vdbgpu_map* map, double accumulation_period_s);
/* Copies the cloud into the source's slot (xyz, stride as vdbgpu_insert). */
int32_t vdbgpu_async_submit(vdbgpu_map* map, const char* source_id, const float* xyz, int64_t n,
                            int64_t stride_bytes, const double origin[3], double max_rate);
/* Waits until every slot is empty and its accumulate done, then integrates;
 * returns the first error any engine call met. */
int32_t vdbgpu_async_flush(vdbgpu_map* map);
/* out: submitted, replaced (dropped), accumulated, integrations. */
int32_t vdbgpu_async_counters(vdbgpu_map* map, int64_t out[4]);
/* Stops and joins the threads (vdbgpu_destroy does it too). */
int32_t vdbgpu_async_stop(vdbgpu_map* map);

/* ---- direct map edits --------------------------------------------------------
 * addPointsToGrid (occupied = 1) / removePointsFromGrid (occupied = 0),
 * VM:413-447: the voxel floor(p / resolution) of every point (no half-voxel
 * shift) becomes (max log-odds, active) or (min log-odds, inactive)
 * (setNodeToOccupied / setNodeToFree, OCC:118-129). Pending update sets are
 * not touched. A NaN or out-of-range point rejects the call
 * (VDBGPU_OUT_OF_RANGE) before any change. */
int32_t vdbgpu_set_points(vdbgpu_map* map, const float* xyz, int64_t n, int64_t stride_bytes,
                          int32_t occupied);

/* addArtificialAreas (VM:1175-1236): restores the previous areas
 * (restoreMapIntegrity), then walls between consecutive vertices (and last
 * to first) of each polygon, rounded like sensor origins (VM:612-631), cast
 * for every height layer k in [(int)(negative_height / res),
 * (int)(positive_height / res)) into the artificial grid. Polygon p has
 * counts[p] vertices, taken in order from xyz (x, y, z per vertex). The
 * artificial voxels are set active at the end of every later non-empty
 * map update (VM:785-789). Walls longer than 8192 voxels along one axis give
 * VDBGPU_UNSUPPORTED; a rejected call changes nothing. */
int32_t vdbgpu_add_artificial_areas(vdbgpu_map* map, int32_t n_polygons, const int32_t* counts,
                                    const double* xyz, double negative_height,
                                    double positive_height);
/* addArtificialWall (VM:1217-1236) for n_walls walls, xyz = {start x, y, z,
 * end x, y, z} per wall, each cast for the height layers as above, added to
 * the artificial grid without a restore (addArtificialPolygon is the walls
 * i -> (i + 1) mod size of a polygon). A rejected call changes nothing. */
int32_t vdbgpu_add_artificial_walls(vdbgpu_map* map, int32_t n_walls, const double* xyz,
                                    double negative_height, double positive_height);
/* restoreMapIntegrity (VM:1152-1166): every artificial voxel gets
 * active = value > thres_max (setNodeState, OCC:130-134); the grid is cleared. */
int32_t vdbgpu_restore_map_integrity(vdbgpu_map* map);
/* The artificial grid: leaf origins and voxel masks, sorted by origin (the
 * m_artificial_area_grid the reference keeps); null outputs return the count,
 * a short cap returns -count. */
int64_t vdbgpu_export_artificial(vdbgpu_map* map, int32_t* origins, uint64_t* masks, int64_t cap);

/* ---- map sections -------------------------------------------------------------
 * getMapSectionGrid / getMapSectionUpdateGrid (VM:884-1014) over the inclusive
 * index box bbox = {min x, y, z, max x, y, z} (createIndexBoundingBox, see
 * vdbgpu_index_bbox). full_grid = 1: the inside voxels of every overlapping
 * leaf with value and state (extractFullLeaf); 0: the inside active voxels
 * with value 1.0 (extractSparseLeaf; true in an update grid). Voxels outside
 * the box are (0, inactive). Leaves sorted by origin as in
 * vdbgpu_export_leaves; null outputs return the count, a short cap returns
 * -count, an error -1. */
int64_t vdbgpu_export_section(vdbgpu_map* map, const int32_t bbox[6], int32_t full_grid,
                              int32_t* origins, float* values, uint64_t* masks, int64_t cap);
/* applyMapSectionGrid (VM:1022-1050): every voxel of every
 * section leaf is written to the map with its value and state, and the
 * inactive tiles of the section grid's internal nodes write (0, inactive) at
 * their origins, as the reference's cbeginValueAll loop does. Origins must be
 * distinct multiples of 8. For smooth_map = true the caller closes the
 * section first (vdbgpu_morphology, op 2). */
int32_t vdbgpu_apply_section(vdbgpu_map* map, int64_t n, const int32_t* origins,
                             const float* values, const uint64_t* masks);
/* applyMapSectionUpdateGrid (VM:1058-1085), smoothing as above: active map
 * voxels inside bbox (the section's bb_min / bb_max) are deactivated, then
 * every active section voxel is activated (a missing leaf is created with
 * value 0). */
int32_t vdbgpu_apply_section_update(vdbgpu_map* map, const int32_t bbox[6], int64_t n,
                                    const int32_t* origins, const uint64_t* masks);
/* The public morphology members (VM:1095-1148) on a leaf set, e.g. a section
 * grid's leaves: op 0 morphologicalDilateMap, 1 morphologicalErodeMap,
 * 2 morphologicalCloseMap (dilate, then erode), 3 morphologicalOpenMap
 * (erode, then dilate), each `iterations` steps of tools::dilateActiveValues /
 * erodeActiveValues with 26-neighbour connectivity (NN_FACE_EDGE_VERTEX): a
 * dilation step activates every voxel with an active neighbour, an erosion
 * step keeps an active voxel only if all 26 neighbours are active (a voxel in
 * no leaf is inactive). n leaves (origins distinct multiples of 8, 8 active-mask
 * words each) in; out: the n input leaves in input order with their new masks,
 * then the leaves dilation created (sorted by origin; a created leaf takes the
 * grid's background value, an emptied leaf stays). Values are not touched.
 * Null outputs return the count, a short cap -count, an error -1
 * (vdbgpu_last_error); iterations <= 0 return the input unchanged. Runs on
 * the map's device and stream; the map itself is not read or changed. This
 * is what applyMapSectionGrid / UpdateGrid run first when smooth_map is set
 * (morphologicalCloseMap on the section, VM:1026-1029, :1062-1065). */
int64_t vdbgpu_morphology(vdbgpu_map* map, int32_t op, int32_t iterations, int64_t n,
                          const int32_t* origins, const uint64_t* masks, int32_t* out_origins,
                          uint64_t* out_masks, int64_t cap);
/* ---- persistence ---------------------------------------------------------------
 * The point list of saveMapToPCD (VM:218-256): every active voxel at
 * indexToWorld(ijk) + 0.5 resolution, as float xyz, in OpenVDB's cbeginValueOn
 * order (root children by origin, 128^3 nodes and leaves by offset, voxels by
 * offset). A null output returns the count, a short cap -count, an error -1. */
int64_t vdbgpu_export_active_points(vdbgpu_map* map, float* xyz, int64_t cap);
/* createMapFromPointCloud (OCC:136-174), the map half of loadMapFromPCD: with
 * clear_map the grid is emptied (update sets kept); every point's voxel
 * floor(p / resolution) becomes (max log-odds, active); with set_background
 * every inactive voxel of the active-voxel bounding box gets min log-odds;
 * then pruneGrid (OCC:173) at the leaf level: every leaf of 512 equal values,
 * all active or all inactive, becomes a tile. Tiles are not leaves: leaf
 * counts, leaf and section exports skip them, vdbgpu_export_tiles lists them,
 * the raytrace topology skips inactive ones, saveMapToPCD's point list holds
 * an active tile once (at its origin). Voxel reads and writes see the tile's
 * value; the first write that changes a voxel of a tile makes it a leaf again.
 * A NaN or out-of-range point rejects the call before any change; a
 * background box of more than 2^27 leaves returns VDBGPU_UNSUPPORTED. */
int32_t vdbgpu_create_from_points(vdbgpu_map* map, const float* xyz, int64_t n, int64_t stride_bytes,
                                  int32_t set_background, int32_t clear_map);
/* The leaf-level tiles of the last createMapFromPointCloud that are still
 * tiles, sorted by origin: origins (multiples of 8), tile values, active
 * flags (1 = active). A null output returns the count, a short cap -count,
 * an error -1. Replaces reading the tiles of the 16^3 internal nodes of the
 * reference's pruned grid (tree::InternalNode tile table, OCC:173). */
int64_t vdbgpu_export_tiles(vdbgpu_map* map, int32_t* origins, float* values, uint8_t* active,
                            int64_t cap);

/* createIndexBoundingBox (VM:811-874): tf is the row-major 4x4 map-to-reference
 * transform; out = {min x, y, z, max x, y, z}. */
int32_t vdbgpu_index_bbox(double resolution, const double min_boundary[3],
                          const double max_boundary[3], const double tf[16], int32_t out[6]);

#ifdef __cplusplus
}
#endif

#endif /* VDBGPU_H */
