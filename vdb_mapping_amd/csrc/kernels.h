// Host-side launchers of the gfx950 kernels (defined in *_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"

namespace vdbgpu {

hipError_t launch_build_table(hipStream_t s, double* tbl, uint32_t n_lo, uint32_t n_hi);
hipError_t launch_prep(hipStream_t s, int grid, const float* pts, int64_t stride_bytes, bool vec4,
                       ScanInfo* scans, const BatchCtl* ctl, double res, double inv,
                       uint32_t tbl_n, int4* rayinfo, uint32_t* pre, uint32_t* bsum, Status* st);
hipError_t launch_scan_blocks(hipStream_t s, const uint32_t* bsum, uint64_t* boff,
                              const BatchCtl* ctl, Status* st);
hipError_t launch_ray_offsets(hipStream_t s, int grid, const int4* rayinfo, const uint32_t* pre,
                              const uint64_t* boff, const BatchCtl* ctl, uint32_t chunk,
                              uint64_t* roff, uint32_t* first_ray, uint64_t max_chunks,
                              Status* st);
hipError_t launch_traverse(hipStream_t s, int grid, int block, const int4* rayinfo,
                           const uint64_t* roff, const uint32_t* first_ray, const ScanInfo* scans,
                           const BatchCtl* ctl, const double* tbl, UpdView u, Status* st,
                           uint32_t chunk);
hipError_t launch_fast_cast(hipStream_t s, int grid, const int4* rayinfo, const ScanInfo* scans,
                            const BatchCtl* ctl, RtView rt, MapView mp, UpdView u, bool cast,
                            Status* st);
hipError_t launch_compact(hipStream_t s, int grid, UpdView u, uint32_t cap, uint32_t* list,
                          unsigned int* counter, const Status* st);
hipError_t launch_group_build(hipStream_t s, int grid, UpdView u, const uint32_t* list, GroupView gv,
                              bool create_hit, bool create_miss, const Status* st);
hipError_t launch_map_check(hipStream_t s, Status* st, const Status* stmap, uint32_t lcap,
                            uint32_t mcap);
hipError_t launch_group_insert(hipStream_t s, int grid, GroupView gv, const uint32_t* glist, MapView mp,
                               uint32_t* gidx, unsigned long long* noop_set, uint32_t noop_mask,
                               const Status* st, Status* stmap);
hipError_t launch_fold(hipStream_t s, int grid, UpdView u, GroupView gv, const uint32_t* glist,
                       const uint32_t* gidx, MapView mp, const float lo[6], const Status* st,
                       Status* stcount);
hipError_t launch_clear_keys(hipStream_t s, int grid, UpdView u, const uint32_t* list,
                             const Status* st);
hipError_t launch_zero_words(hipStream_t s, unsigned long long* p, uint32_t n);
hipError_t launch_map_insert(hipStream_t s, int grid, UpdView u, const uint32_t* list,
                             uint32_t count, MapView mp, uint32_t* eidx, bool create_hit,
                             bool create_miss, unsigned long long* noop_set, uint32_t noop_mask,
                             Status* st, Status* stc);
hipError_t launch_rec_gather(hipStream_t s, int grid, const uint32_t* list, UpdView sparse, const RecDest* dest,
                             const Status* st);
hipError_t launch_put_dest(hipStream_t s, RecDest* dst, RecDest v);
hipError_t launch_iota_u32(hipStream_t s, uint32_t* p, uint32_t n);
hipError_t launch_snap_count(hipStream_t s, const Status* st, uint32_t* host_dst);
// Deferred brick-path scans applied as one group pass (integrate_kernels.hip).
constexpr int kDeferMax = 8;                       // scans per group pass
constexpr uint32_t kGroupScanShift = 29;           // gslot code: scan << 29 | record
constexpr uint32_t kGroupRecMask = (1u << kGroupScanShift) - 1;
struct GroupScans {
  const unsigned long long* masks[kDeferMax];  // record masks of scan t (16 words per position)
  int nscan;
};
hipError_t launch_group_add(hipStream_t s, const uint32_t* eidx, const uint32_t* list, uint32_t count, uint32_t t,
                            uint32_t* gslot, uint32_t* rgrp, uint32_t rgrp_stride, uint32_t* glist, uint32_t* grec,
                            unsigned int* ngroups);
hipError_t launch_apply_groups(hipStream_t s, int grid, const GroupScans& gs, const uint32_t* glist,
                               const uint32_t* grec, const unsigned int* ngroups, uint32_t* gslot, MapView mp,
                               const float lo[6], Status* st);
hipError_t launch_apply(hipStream_t s, int grid, UpdView u, const uint32_t* list,
                        const uint32_t* eidx, uint32_t count, MapView mp, const float lo[6],
                        Status* st, bool clear);
hipError_t launch_rehash_update(hipStream_t s, int grid, UpdView from, uint32_t from_cap,
                                UpdView to, Status* st);
// stc non-null: the leaf count is read from stc->leaf_count on the device.
hipError_t launch_rehash_map(hipStream_t s, int grid, MapView mp, uint32_t leaf_count, const Status* stc);
hipError_t launch_fill_u64(hipStream_t s, unsigned long long* p, uint64_t n, uint64_t v);
hipError_t launch_stage_scan(hipStream_t s, const ScanInfo& sc, const BatchCtl& ctl, ScanInfo* d_scans,
                             BatchCtl* d_ctl, Status* st, unsigned long long* bkeys, unsigned long long* bcnt,
                             uint64_t nb);
hipError_t launch_dirty_list(hipStream_t s, int grid, uint8_t* dirty, uint32_t leaf_count,
                             uint32_t* list, Status* st);
hipError_t launch_gather_leaves(hipStream_t s, int grid, const uint32_t* list, uint32_t count,
                                MapView mp, unsigned long long* keys_out, float* vals_out,
                                unsigned long long* act_out);

// Entries of a workgroup's brick table in k_seg_count / k_seg_fill (LDS); the
// count stores each workgroup's table (keys; brick slot << 32 | segments) at
// workgroup * kSegTab of wkey / wsc, and the fill rereads it.
constexpr int kSegTab = 1024;
hipError_t launch_seg_count(hipStream_t s, int grid, const int4* rayinfo, const ScanInfo* scans,
                            const BatchCtl* ctl, const double* tbl, BrickView bv, unsigned long long* wkey,
                            unsigned long long* wsc, Status* st);
hipError_t launch_brick_offsets(hipStream_t s, BrickView bv, uint64_t* bsum, uint64_t* boff,
                                BrickItem* items, uint32_t seg_cap, uint32_t item_cap,
                                uint32_t* split_slot, uint32_t split_cap, Status* st);
hipError_t launch_seg_fill(hipStream_t s, int grid, const int4* rayinfo, const ScanInfo* scans,
                           const BatchCtl* ctl, const double* tbl, BrickView bv,
                           const unsigned long long* wkey, const unsigned long long* wsc, uint4* segs,
                           uint32_t seg_cap, Status* st);
hipError_t launch_brick_walk(hipStream_t s, int grid, const BrickItem* items, const uint4* segs,
                             const int4* rayinfo, const ScanInfo* scans, const double* tbl,
                             BrickView bv, unsigned long long* shared_bits, UpdView rec,
                             unsigned long long* leaf_mask, uint32_t rec_cap, Status* st);
hipError_t launch_rec_list(hipStream_t s, const unsigned long long* leaf_mask, uint64_t* bsum, uint64_t* boff,
                           uint32_t nb_cap, uint32_t* list, Status* st);
hipError_t launch_split_emit(hipStream_t s, int grid, const uint32_t* split_slot,
                             unsigned long long* shared_bits, BrickView bv, UpdView rec,
                             unsigned long long* leaf_mask, uint32_t rec_cap, Status* st);

hipError_t launch_snapshot(hipStream_t s, MapView mp, uint32_t leaf_count, RtView rt, int* bbox, TileView tv);
hipError_t launch_copy_set(hipStream_t s, const unsigned long long* src, uint32_t src_cap,
                           unsigned long long* dst, uint32_t dst_mask);
hipError_t launch_raytrace(hipStream_t s, int grid, RtView rt, MapView mp, double inv, double res,
                           int64_t n, const double* origins, const double* dirs,
                           const double* lens, uint8_t* success, double* ends, unsigned int* err);

hipError_t launch_mark_points(hipStream_t s, const float* pts, int64_t stride_bytes, uint32_t n, double inv,
                              UpdView u, Status* st);
hipError_t launch_mark_walls(hipStream_t s, const double* verts, const int4* walls, const uint64_t* prefix,
                             uint32_t n_walls, uint64_t total, double res, double inv, const double* tbl,
                             uint32_t tbl_n, UpdView u, Status* st);
hipError_t launch_set_voxels(hipStream_t s, UpdView u, const uint32_t* list, const uint32_t* eidx,
                             uint32_t count, MapView mp, int op, float val, float t_max, bool clear);
hipError_t launch_section_pick(hipStream_t s, MapView mp, uint32_t leaf_count, const int* bbox, bool full,
                               uint32_t* out, unsigned int* counter, TileView tv);
hipError_t launch_section_gather(hipStream_t s, MapView mp, const uint32_t* list, uint32_t count,
                                 const int* bbox, bool full, unsigned long long* keys_out, float* vals_out,
                                 unsigned long long* act_out);
hipError_t launch_upsert_leaves(hipStream_t s, const unsigned long long* keys, const uint8_t* need, uint32_t n,
                                MapView mp, uint32_t* eidx, uint16_t* first, Status* st);
hipError_t launch_section_write(hipStream_t s, const float* vals, const unsigned long long* act,
                                const uint32_t* eidx, const uint16_t* first, uint32_t n, MapView mp);
hipError_t launch_section_tiles(hipStream_t s, const unsigned long long* n128, uint32_t c128,
                                const unsigned long long* n4096, uint32_t c4096,
                                const unsigned long long* sleaf, uint32_t sleaf_mask,
                                const unsigned long long* s128, uint32_t s128_mask, MapView mp);
hipError_t launch_section_clear(hipStream_t s, MapView mp, uint32_t leaf_count, const int* bbox, TileView tv);
// pruneGrid at the leaf level over leaves [0, leaf_count): tiles set in tv (bits and on
// zeroed before), their indices appended to list at *counter.
hipError_t launch_prune(hipStream_t s, MapView mp, uint32_t leaf_count, TileView tv, uint32_t* list,
                        unsigned int* counter);
// Tiles of `list` whose voxels are no longer all (val, state) become leaves again.
hipError_t launch_tile_check(hipStream_t s, MapView mp, const uint32_t* list, uint32_t n, TileView tv);
hipError_t launch_section_prep(hipStream_t s, const int* origins, const float* vals, const unsigned long long* act,
                               uint32_t n, unsigned long long* keys, uint8_t* need, uint16_t* first,
                               unsigned long long* sleaf, uint32_t sleaf_mask, unsigned long long* s128,
                               uint32_t s128_mask, unsigned long long* n128, unsigned long long* s4096,
                               uint32_t s4096_mask, unsigned long long* n4096, unsigned int* counters);
hipError_t launch_section_or(hipStream_t s, const unsigned long long* act, const uint32_t* eidx, uint32_t n,
                             MapView mp);

hipError_t launch_active_bbox(hipStream_t s, MapView mp, uint32_t leaf_count, int* box);
hipError_t launch_box_keys(hipStream_t s, const int lo[3], const uint32_t dims[3], unsigned long long* keys,
                           uint8_t* need, bool need_v);
hipError_t launch_box_fill(hipStream_t s, const unsigned long long* keys, const uint32_t* eidx, uint32_t n,
                           const int* box, MapView mp, float v);

// Leaf table of the morphology operators (morph_kernels.hip): per leaf index
// its origin (3 ints) and 27 neighbour indices; a hash of leaf key -> index.
struct MorphView {
  int* org;
  unsigned long long* keys;  // hash slots, kEmpty = free
  uint32_t* hidx;            // hash slot -> leaf index
  uint32_t hmask;
  uint32_t* nbr;
};
hipError_t launch_morph_seed(hipStream_t s, MorphView mv, uint32_t count);
hipError_t launch_morph_grow(hipStream_t s, MorphView mv, unsigned long long* masks, uint8_t* keep, uint32_t count,
                             uint32_t cap, unsigned int* added);
hipError_t launch_morph_nbrs(hipStream_t s, MorphView mv, uint32_t count);
hipError_t launch_morph_step(hipStream_t s, const unsigned long long* in, unsigned long long* out,
                             const uint32_t* nbr, uint32_t count, bool erode);
hipError_t launch_morph_keep(hipStream_t s, const unsigned long long* masks, uint8_t* keep, uint32_t count);

}  // namespace vdbgpu
