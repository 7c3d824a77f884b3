"""MI355X (gfx950) occupancy-integration path of vdb_mapping.

The product is libvdbgpu.so (C ABI in include/vdbgpu.h, HIP kernels in
csrc/); this module is a thin ctypes binding of that ABI with the method
names of the reference's OccupancyVDBMapping (insertPointCloud ->
insert_point_cloud, ...). There is no CPU fallback: every method calls the
library, and loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvdbgpu.so")
INCLUDE_DIR = os.path.join(os.path.dirname(HERE), "include")

OK, NOT_CONFIGURED, UNKNOWN_SOURCE, BAD_MAX_RANGE, BAD_PROB_MISS, BAD_PROB_HIT = 0, 1, 2, 3, 4, 5
OUT_OF_RANGE, UNSUPPORTED, DEVICE_ERROR, BAD_ARGUMENT, NO_INTERSECTOR = 6, 7, 8, 9, 10
OVER_BUDGET = 11
NOALLOC_MAX_QUERIES = 393216  # vdbgpu_raytrace_noalloc: queries per call its static scratch holds

# Every symbol include/vdbgpu.h declares.
ABI_SYMBOLS = [
    "vdbgpu_version", "vdbgpu_status_string", "vdbgpu_device_count", "vdbgpu_create",
    "vdbgpu_remove_source", "vdbgpu_has_source",
    "vdbgpu_destroy", "vdbgpu_set_config", "vdbgpu_logodds", "vdbgpu_add_source",
    "vdbgpu_accumulate", "vdbgpu_integrate", "vdbgpu_insert", "vdbgpu_insert_batch", "vdbgpu_reset",
    "vdbgpu_empty",
    "vdbgpu_leaf_count", "vdbgpu_export_leaves", "vdbgpu_export_dirty_leaves",
    "vdbgpu_update_leaf_count", "vdbgpu_export_update", "vdbgpu_raytrace", "vdbgpu_raytrace_device",
    "vdbgpu_raytrace_noalloc", "vdbgpu_release_scratch", "vdbgpu_synchronize",
    "vdbgpu_get_stats", "vdbgpu_set_profiling", "vdbgpu_set_tuning", "vdbgpu_get_tuning",
    "vdbgpu_last_error", "vdbgpu_device_alloc", "vdbgpu_device_free", "vdbgpu_device_upload",
    "vdbgpu_device_download", "vdbgpu_device_copy", "vdbgpu_async_start", "vdbgpu_async_submit",
    "vdbgpu_async_flush", "vdbgpu_async_counters", "vdbgpu_async_stop", "vdbgpu_server_create", "vdbgpu_server_submit", "vdbgpu_server_drain",
    "vdbgpu_server_records", "vdbgpu_server_map", "vdbgpu_server_device_bytes",
    "vdbgpu_server_destroy", "vdbgpu_set_points", "vdbgpu_add_artificial_areas",
    "vdbgpu_restore_map_integrity", "vdbgpu_export_artificial", "vdbgpu_add_artificial_walls", "vdbgpu_export_section",
    "vdbgpu_apply_section", "vdbgpu_apply_section_update", "vdbgpu_index_bbox",
    "vdbgpu_export_active_points", "vdbgpu_create_from_points", "vdbgpu_export_tiles", "vdbgpu_morphology",
]


class Config(ctypes.Structure):
    """vdbgpu_config: BaseConfig (VDBMapping.hpp:64-71) + Config (OccupancyVDBMapping.hpp:38-44)."""

    _fields_ = [
        ("max_range", ctypes.c_double),
        ("fast_mode", ctypes.c_int32),
        ("_pad", ctypes.c_int32),
        ("accumulation_period", ctypes.c_double),
        ("prob_hit", ctypes.c_double),
        ("prob_miss", ctypes.c_double),
        ("prob_thres_min", ctypes.c_double),
        ("prob_thres_max", ctypes.c_double),
    ]


def make_config(max_range=10.0, fast_mode=False, prob_hit=0.9, prob_miss=0.1,
                prob_thres_min=0.49, prob_thres_max=0.51, accumulation_period=0.0) -> Config:
    return Config(max_range, int(fast_mode), 0, accumulation_period, prob_hit, prob_miss,
                  prob_thres_min, prob_thres_max)


class Stats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in
                ("scans", "rays", "visits", "voxel_updates", "update_leaves", "leaf_count",
                 "device_bytes")] + \
               [(n, ctypes.c_double) for n in ("ms_prep", "ms_traverse", "ms_compact", "ms_apply",
                                               "ms_raytrace")] + \
               [(n, ctypes.c_uint64) for n in ("n_prep", "n_traverse", "n_compact", "n_apply",
                                               "n_raytrace")] + \
               [("ms_walk", ctypes.c_double), ("n_walk", ctypes.c_uint64), ("scratch_bytes", ctypes.c_uint64)]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Tuning(ctypes.Structure):
    _fields_ = [("block", ctypes.c_int32), ("steps_per_lane", ctypes.c_int32),
                ("waves_per_cu", ctypes.c_int32), ("load_factor", ctypes.c_double),
                ("traversal", ctypes.c_int32), ("defer", ctypes.c_int32)]


class RequestRecord(ctypes.Structure):
    _fields_ = [("request_id", ctypes.c_int64), ("points", ctypes.c_int64), ("map_index", ctypes.c_int32),
                ("batch_size", ctypes.c_int32), ("status", ctypes.c_int32), ("mem_waits", ctypes.c_int32),
                ("latency_ms", ctypes.c_double)]


class VdbGpuError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libvdbgpu.so (raises when it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise VdbGpuError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        P, D, I32, I64 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int64
        L.vdbgpu_version.restype = I32
        L.vdbgpu_status_string.restype = ctypes.c_char_p
        L.vdbgpu_status_string.argtypes = [I32]
        L.vdbgpu_device_count.restype = I32
        L.vdbgpu_create.restype = I32
        L.vdbgpu_create.argtypes = [D, ctypes.POINTER(P)]
        L.vdbgpu_destroy.argtypes = [P]
        L.vdbgpu_set_config.restype = I32
        L.vdbgpu_set_config.argtypes = [P, ctypes.POINTER(Config)]
        L.vdbgpu_logodds.restype = I32
        L.vdbgpu_logodds.argtypes = [P, P]
        L.vdbgpu_add_source.restype = I32
        L.vdbgpu_add_source.argtypes = [P, ctypes.c_char_p, D, D]
        L.vdbgpu_remove_source.restype = I32
        L.vdbgpu_remove_source.argtypes = [P, ctypes.c_char_p]
        L.vdbgpu_has_source.restype = I32
        L.vdbgpu_has_source.argtypes = [P, ctypes.c_char_p]
        for f in (L.vdbgpu_accumulate, L.vdbgpu_insert):
            f.restype = I32
            f.argtypes = [P, ctypes.c_char_p, P, I64, I64, P, I32]
        L.vdbgpu_insert_batch.restype = I32
        L.vdbgpu_insert_batch.argtypes = [P, ctypes.c_char_p, I32, P, P, I64, P, P, I32]
        for f in (L.vdbgpu_integrate, L.vdbgpu_reset, L.vdbgpu_empty, L.vdbgpu_synchronize):
            f.restype = I32
            f.argtypes = [P]
        L.vdbgpu_leaf_count.restype = I64
        L.vdbgpu_leaf_count.argtypes = [P]
        for f in (L.vdbgpu_export_leaves, L.vdbgpu_export_dirty_leaves):
            f.restype = I64
            f.argtypes = [P, P, P, P, I64]
        L.vdbgpu_update_leaf_count.restype = I64
        L.vdbgpu_update_leaf_count.argtypes = [P, ctypes.c_char_p]
        L.vdbgpu_export_update.restype = I64
        L.vdbgpu_export_update.argtypes = [P, ctypes.c_char_p, P, P, P, I64]
        for f in (L.vdbgpu_raytrace, L.vdbgpu_raytrace_device, L.vdbgpu_raytrace_noalloc):
            f.restype = I32
            f.argtypes = [P, I64, P, P, P, P, P]
        L.vdbgpu_get_stats.restype = I32
        L.vdbgpu_get_stats.argtypes = [P, ctypes.POINTER(Stats)]
        L.vdbgpu_set_profiling.restype = I32
        L.vdbgpu_set_profiling.argtypes = [P, I32]
        L.vdbgpu_set_tuning.restype = I32
        L.vdbgpu_set_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
        L.vdbgpu_get_tuning.restype = I32
        L.vdbgpu_get_tuning.argtypes = [P, ctypes.POINTER(Tuning)]
        L.vdbgpu_last_error.restype = ctypes.c_char_p
        L.vdbgpu_last_error.argtypes = [P]
        L.vdbgpu_device_alloc.restype = I32
        L.vdbgpu_device_alloc.argtypes = [I64, ctypes.POINTER(P)]
        L.vdbgpu_device_free.restype = I32
        L.vdbgpu_device_free.argtypes = [P]
        L.vdbgpu_async_start.restype = I32
        L.vdbgpu_async_start.argtypes = [P, D]
        L.vdbgpu_async_submit.restype = I32
        L.vdbgpu_async_submit.argtypes = [P, ctypes.c_char_p, P, I64, I64, P, D]
        for f in (L.vdbgpu_async_flush, L.vdbgpu_async_stop):
            f.restype = I32
            f.argtypes = [P]
        L.vdbgpu_async_counters.restype = I32
        L.vdbgpu_async_counters.argtypes = [P, P]
        for f in (L.vdbgpu_device_upload, L.vdbgpu_device_download, L.vdbgpu_device_copy):
            f.restype = I32
            f.argtypes = [P, P, I64]
        L.vdbgpu_server_create.restype = I32
        L.vdbgpu_server_create.argtypes = [I32, D, ctypes.POINTER(Config), D, I64, I32, ctypes.POINTER(P)]
        L.vdbgpu_server_submit.restype = I32
        L.vdbgpu_server_submit.argtypes = [P, I32, P, I64, I64, P, ctypes.POINTER(I64)]
        L.vdbgpu_server_drain.restype = I32
        L.vdbgpu_server_drain.argtypes = [P]
        L.vdbgpu_server_records.restype = I64
        L.vdbgpu_server_records.argtypes = [P, P, I64]
        L.vdbgpu_server_map.restype = P
        L.vdbgpu_server_map.argtypes = [P, I32]
        L.vdbgpu_server_device_bytes.restype = I64
        L.vdbgpu_server_device_bytes.argtypes = [P]
        L.vdbgpu_server_destroy.argtypes = [P]
        L.vdbgpu_set_points.restype = I32
        L.vdbgpu_set_points.argtypes = [P, P, I64, I64, I32]
        L.vdbgpu_add_artificial_areas.restype = I32
        L.vdbgpu_add_artificial_areas.argtypes = [P, I32, P, P, D, D]
        L.vdbgpu_add_artificial_walls.restype = I32
        L.vdbgpu_add_artificial_walls.argtypes = [P, I32, P, D, D]
        L.vdbgpu_restore_map_integrity.restype = I32
        L.vdbgpu_restore_map_integrity.argtypes = [P]
        L.vdbgpu_export_artificial.restype = I64
        L.vdbgpu_export_artificial.argtypes = [P, P, P, I64]
        L.vdbgpu_export_section.restype = I64
        L.vdbgpu_export_section.argtypes = [P, P, I32, P, P, P, I64]
        L.vdbgpu_apply_section.restype = I32
        L.vdbgpu_apply_section.argtypes = [P, I64, P, P, P]
        L.vdbgpu_apply_section_update.restype = I32
        L.vdbgpu_apply_section_update.argtypes = [P, P, I64, P, P]
        L.vdbgpu_export_active_points.restype = I64
        L.vdbgpu_export_active_points.argtypes = [P, P, I64]
        L.vdbgpu_create_from_points.restype = I32
        L.vdbgpu_create_from_points.argtypes = [P, P, I64, I64, I32, I32]
        L.vdbgpu_export_tiles.restype = I64
        L.vdbgpu_export_tiles.argtypes = [P, P, P, P, I64]
        L.vdbgpu_morphology.restype = I64
        L.vdbgpu_morphology.argtypes = [P, I32, I32, I64, P, P, P, P, I64]
        L.vdbgpu_index_bbox.restype = I32
        L.vdbgpu_index_bbox.argtypes = [D, P, P, P, P]
        _lib = L
    return _lib


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, int):
        return ctypes.c_void_p(a)
    return a.ctypes.data_as(ctypes.c_void_p)


def index_bbox(resolution: float, min_boundary, max_boundary, tf=None) -> np.ndarray:
    """createIndexBoundingBox: inclusive index box (min xyz, max xyz) of the
    world box [min, max] seen through the 4x4 map-to-reference transform."""
    mn = np.ascontiguousarray(min_boundary, dtype=np.float64)
    mx = np.ascontiguousarray(max_boundary, dtype=np.float64)
    t = np.ascontiguousarray(np.eye(4) if tf is None else tf, dtype=np.float64)
    out = np.zeros(6, np.int32)
    rc = lib().vdbgpu_index_bbox(resolution, _ptr(mn), _ptr(mx), _ptr(t), _ptr(out))
    if rc:
        raise VdbGpuError(f"index_bbox: {status_string(rc)}")
    return out


def polygons_arg(polygons):
    """Flatten a list of (k, 3+) vertex arrays into (counts int32, xyz float64)."""
    counts = np.array([len(p) for p in polygons], np.int32)
    rows = [np.asarray(p, np.float64).reshape(len(p), -1)[:, :3] for p in polygons if len(p)]
    xyz = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros((0, 3)), dtype=np.float64)
    return counts, xyz


def status_string(code: int) -> str:
    return lib().vdbgpu_status_string(code).decode()


class GpuMap:
    """OccupancyVDBMapping on one MI355X: method names follow the reference
    (OccupancyVDBMapping.hpp / VDBMapping.hpp), arguments are numpy arrays.

    Methods return the library's status code (0 = ok) where the reference
    returns bool/void; `check=True` raises VdbGpuError on non-zero codes.
    """

    def __init__(self, resolution: float, _borrowed_handle=None):
        self.h = ctypes.c_void_p()
        self._owned = _borrowed_handle is None
        if _borrowed_handle is not None:
            self.h = ctypes.c_void_p(_borrowed_handle)
        else:
            rc = lib().vdbgpu_create(resolution, ctypes.byref(self.h))
            if rc != OK:
                raise VdbGpuError(f"vdbgpu_create failed: {status_string(rc)}")
        self.resolution = resolution

    def close(self):
        if getattr(self, "h", None) and self.h.value and getattr(self, "_owned", True):
            lib().vdbgpu_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        self.close()

    def last_error(self) -> str:
        return lib().vdbgpu_last_error(self.h).decode()

    def _check(self, rc: int, check: bool) -> int:
        if check and rc != OK:
            raise VdbGpuError(f"{status_string(rc)}: {self.last_error()}")
        return rc

    def set_config(self, cfg: Config) -> int:
        return lib().vdbgpu_set_config(self.h, ctypes.byref(cfg))

    def logodds(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        lib().vdbgpu_logodds(self.h, _ptr(out))
        return out

    def add_input_source(self, source_id: str, max_range: float, max_rate: float = 0.0) -> int:
        return lib().vdbgpu_add_source(self.h, source_id.encode(), max_range, max_rate)

    def remove_input_source(self, source_id: str) -> int:
        return lib().vdbgpu_remove_source(self.h, source_id.encode())

    def has_source(self, source_id: str) -> bool:
        return lib().vdbgpu_has_source(self.h, source_id.encode()) == OK

    @staticmethod
    def _pts(points):
        pts = np.ascontiguousarray(points, dtype=np.float32)
        if pts.ndim != 2 or pts.shape[1] not in (3, 4):
            raise ValueError("points must be (n,3) or (n,4) float32")
        return pts, pts.shape[1] * 4

    def accumulate(self, source_id: str, points, origin, check=False) -> int:
        pts, stride = self._pts(points)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        rc = lib().vdbgpu_accumulate(self.h, source_id.encode(), _ptr(pts), pts.shape[0], stride,
                                     _ptr(o), 0)
        return self._check(rc, check)

    def accumulate_device(self, source_id: str, dev_ptr: int, n: int, stride_bytes: int, origin,
                          check=False) -> int:
        o = np.ascontiguousarray(origin, dtype=np.float64)
        rc = lib().vdbgpu_accumulate(self.h, source_id.encode(), ctypes.c_void_p(dev_ptr), n,
                                     stride_bytes, _ptr(o), 1)
        return self._check(rc, check)

    def integrate(self, check=False) -> int:
        return self._check(lib().vdbgpu_integrate(self.h), check)

    def insert(self, source_id: str, points, origin, check=False) -> int:
        pts, stride = self._pts(points)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        rc = lib().vdbgpu_insert(self.h, source_id.encode(), _ptr(pts), pts.shape[0], stride,
                                 _ptr(o), 0)
        return self._check(rc, check)

    insert_point_cloud = insert

    def insert_batch(self, source_id: str, clouds, origins, max_ranges=None, check=False) -> int:
        """Batched insertPointCloud of several host clouds ((n_i, 3|4) float32 each)."""
        clouds = [np.ascontiguousarray(c, dtype=np.float32) for c in clouds]
        width = clouds[0].shape[1] if clouds else 4
        if any(c.ndim != 2 or c.shape[1] != width for c in clouds):
            raise ValueError("all clouds must be (n, 3) or all (n, 4) float32")
        cat = np.concatenate(clouds) if clouds else np.zeros((0, width), np.float32)
        counts = np.array([c.shape[0] for c in clouds], np.int64)
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        r = None if max_ranges is None else np.ascontiguousarray(max_ranges, dtype=np.float64)
        rc = lib().vdbgpu_insert_batch(self.h, source_id.encode(), len(clouds), _ptr(cat), _ptr(counts),
                                       width * 4, _ptr(o), _ptr(r), 0)
        return self._check(rc, check)

    def insert_batch_device(self, source_id: str, dev_ptr: int, counts, stride_bytes: int, origins,
                            max_ranges=None, check=False) -> int:
        counts = np.ascontiguousarray(counts, dtype=np.int64)
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        r = None if max_ranges is None else np.ascontiguousarray(max_ranges, dtype=np.float64)
        rc = lib().vdbgpu_insert_batch(self.h, source_id.encode(), len(counts), ctypes.c_void_p(dev_ptr),
                                       _ptr(counts), stride_bytes, _ptr(o), _ptr(r), 1)
        return self._check(rc, check)

    def insert_device(self, source_id: str, dev_ptr: int, n: int, stride_bytes: int, origin,
                      check=False) -> int:
        o = np.ascontiguousarray(origin, dtype=np.float64)
        rc = lib().vdbgpu_insert(self.h, source_id.encode(), ctypes.c_void_p(dev_ptr), n,
                                 stride_bytes, _ptr(o), 1)
        return self._check(rc, check)

    def reset(self) -> int:
        return lib().vdbgpu_reset(self.h)

    reset_map = reset

    def empty(self) -> bool:
        return bool(lib().vdbgpu_empty(self.h))

    def synchronize(self) -> int:
        return lib().vdbgpu_synchronize(self.h)

    def leaf_count(self) -> int:
        return int(lib().vdbgpu_leaf_count(self.h))

    def export_leaves(self):
        """Leaf dump of the map (getGrid): origins (n,3) i32, values (n,512) f32, masks (n,8) u64."""
        n = self.leaf_count()
        o = np.zeros((n, 3), np.int32)
        v = np.zeros((n, 512), np.float32)
        m = np.zeros((n, 8), np.uint64)
        got = lib().vdbgpu_export_leaves(self.h, _ptr(o), _ptr(v), _ptr(m), n)
        if got != n:
            raise VdbGpuError(f"export_leaves returned {got}, expected {n}: {self.last_error()}")
        return o, v, m

    def export_tiles(self):
        """Leaf-level tiles of the last create_from_points (pruneGrid), sorted by
        origin: origins (n,3) i32, values (n,) f32, active (n,) bool."""
        n = lib().vdbgpu_export_tiles(self.h, None, None, None, 0)
        if n < 0:
            raise VdbGpuError(f"export_tiles failed: {self.last_error()}")
        o = np.zeros((n, 3), np.int32)
        v = np.zeros(n, np.float32)
        a = np.zeros(n, np.uint8)
        got = lib().vdbgpu_export_tiles(self.h, _ptr(o), _ptr(v), _ptr(a), n)
        if got != n:
            raise VdbGpuError(f"export_tiles returned {got}, expected {n}: {self.last_error()}")
        return o, v, a.astype(bool)

    def export_dirty_leaves(self):
        cap = max(self.leaf_count(), 1)
        o = np.zeros((cap, 3), np.int32)
        v = np.zeros((cap, 512), np.float32)
        m = np.zeros((cap, 8), np.uint64)
        got = lib().vdbgpu_export_dirty_leaves(self.h, _ptr(o), _ptr(v), _ptr(m), cap)
        if got < 0:
            raise VdbGpuError(f"export_dirty_leaves failed ({got}): {self.last_error()}")
        return o[:got], v[:got], m[:got]

    def export_update(self, source_id: str):
        n = lib().vdbgpu_update_leaf_count(self.h, source_id.encode())
        if n < 0:
            raise KeyError(source_id)
        o = np.zeros((n, 3), np.int32)
        vis = np.zeros((n, 8), np.uint64)
        hit = np.zeros((n, 8), np.uint64)
        got = lib().vdbgpu_export_update(self.h, source_id.encode(), _ptr(o), _ptr(vis), _ptr(hit), n)
        if got != n:
            raise VdbGpuError(f"export_update returned {got}, expected {n}: {self.last_error()}")
        return o, vis, hit

    # direct edits, artificial areas and map sections
    def set_points(self, points, occupied: bool, check=False) -> int:
        pts, stride = self._pts(points)
        rc = lib().vdbgpu_set_points(self.h, _ptr(pts), pts.shape[0], stride, int(bool(occupied)))
        return self._check(rc, check)

    def add_points_to_grid(self, points, check=False) -> int:
        return self.set_points(points, True, check)

    def remove_points_from_grid(self, points, check=False) -> int:
        return self.set_points(points, False, check)

    def add_artificial_areas(self, polygons, negative_height: float, positive_height: float,
                             check=False) -> int:
        counts, xyz = polygons_arg(polygons)
        rc = lib().vdbgpu_add_artificial_areas(self.h, len(counts), _ptr(counts), _ptr(xyz),
                                               negative_height, positive_height)
        return self._check(rc, check)

    def add_artificial_walls(self, walls, negative_height: float, positive_height: float,
                             check=False) -> int:
        """addArtificialWall for each (start, end) pair, without a restore."""
        w = np.ascontiguousarray([[*np.asarray(a, float)[:3], *np.asarray(b, float)[:3]] for a, b in walls],
                                 dtype=np.float64).reshape(-1, 6)
        rc = lib().vdbgpu_add_artificial_walls(self.h, w.shape[0], _ptr(w), negative_height,
                                               positive_height)
        return self._check(rc, check)

    def add_artificial_wall(self, start, end, negative_height: float, positive_height: float,
                            check=False) -> int:
        return self.add_artificial_walls([(start, end)], negative_height, positive_height, check)

    def add_artificial_polygon(self, polygon, negative_height: float, positive_height: float,
                               check=False) -> int:
        """addArtificialPolygon: walls i -> (i + 1) mod size, without a restore."""
        p = [np.asarray(v, float)[:3] for v in polygon]
        walls = [(p[i], p[(i + 1) % len(p)]) for i in range(len(p))]
        return self.add_artificial_walls(walls, negative_height, positive_height, check)

    def restore_map_integrity(self, check=False) -> int:
        return self._check(lib().vdbgpu_restore_map_integrity(self.h), check)

    def export_artificial(self):
        n = lib().vdbgpu_export_artificial(self.h, None, None, 0)
        if n < 0:
            raise VdbGpuError(f"export_artificial failed: {self.last_error()}")
        o = np.zeros((n, 3), np.int32)
        m = np.zeros((n, 8), np.uint64)
        if lib().vdbgpu_export_artificial(self.h, _ptr(o), _ptr(m), n) != n:
            raise VdbGpuError(f"export_artificial failed: {self.last_error()}")
        return o, m

    def export_section(self, bbox, full_grid: bool):
        """getMapSectionGrid (full_grid) / getMapSectionUpdateGrid over an
        inclusive index box: origins (n,3), values (n,512), masks (n,8)."""
        b = np.ascontiguousarray(bbox, dtype=np.int32)
        n = lib().vdbgpu_export_section(self.h, _ptr(b), int(bool(full_grid)), None, None, None, 0)
        if n < 0:
            raise VdbGpuError(f"export_section failed: {self.last_error()}")
        o = np.zeros((n, 3), np.int32)
        v = np.zeros((n, 512), np.float32)
        m = np.zeros((n, 8), np.uint64)
        got = lib().vdbgpu_export_section(self.h, _ptr(b), int(bool(full_grid)), _ptr(o), _ptr(v),
                                          _ptr(m), n)
        if got != n:
            raise VdbGpuError(f"export_section returned {got}, expected {n}: {self.last_error()}")
        return o, v, m

    def apply_section(self, origins, values, masks, check=False) -> int:
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 512)
        m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, 8)
        return self._check(lib().vdbgpu_apply_section(self.h, o.shape[0], _ptr(o), _ptr(v), _ptr(m)),
                           check)

    def apply_section_update(self, bbox, origins, masks, check=False) -> int:
        b = np.ascontiguousarray(bbox, dtype=np.int32)
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, 8)
        return self._check(lib().vdbgpu_apply_section_update(self.h, _ptr(b), o.shape[0], _ptr(o),
                                                             _ptr(m)), check)

    # persistence (the map side of saveMapToPCD / loadMapFromPCD)
    def export_active_points(self) -> np.ndarray:
        """Active voxel centres (n, 3) float32 in OpenVDB's value-on order."""
        n = lib().vdbgpu_export_active_points(self.h, None, 0)
        if n < 0:
            raise VdbGpuError(f"export_active_points failed: {self.last_error()}")
        out = np.zeros((n, 3), np.float32)
        if lib().vdbgpu_export_active_points(self.h, _ptr(out), n) != n:
            raise VdbGpuError(f"export_active_points failed: {self.last_error()}")
        return out

    def create_map_from_points(self, points, set_background: bool, clear_map: bool, check=False) -> int:
        pts, stride = self._pts(points)
        rc = lib().vdbgpu_create_from_points(self.h, _ptr(pts), pts.shape[0], stride,
                                             int(bool(set_background)), int(bool(clear_map)))
        return self._check(rc, check)

    # morphology of a leaf set (VM:1095-1148), on this map's device
    MORPH_OPS = {"dilate": 0, "erode": 1, "close": 2, "open": 3}

    def morphology(self, op: str, iterations: int, origins, masks):
        """(origins (k, 3) int32, masks (k, 8) uint64) after op: the input
        leaves in input order, then the leaves dilation created, by origin."""
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, 8)
        code = self.MORPH_OPS[op]
        k = lib().vdbgpu_morphology(self.h, code, iterations, o.shape[0], _ptr(o), _ptr(m), None, None, 0)
        if k < 0:
            raise VdbGpuError(f"morphology failed: {self.last_error()}")
        oo = np.zeros((max(k, 1), 3), np.int32)
        om = np.zeros((max(k, 1), 8), np.uint64)
        if lib().vdbgpu_morphology(self.h, code, iterations, o.shape[0], _ptr(o), _ptr(m), _ptr(oo), _ptr(om),
                                   k) != k:
            raise VdbGpuError(f"morphology failed: {self.last_error()}")
        return oo[:k], om[:k]

    def raytrace(self, origins, dirs, lengths):
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        ln = np.ascontiguousarray(lengths, dtype=np.float64).reshape(-1)
        n = o.shape[0]
        succ = np.zeros(n, np.uint8)
        ends = np.zeros((n, 3), np.float64)
        st = lib().vdbgpu_raytrace(self.h, n, _ptr(o), _ptr(d), _ptr(ln), _ptr(succ), _ptr(ends))
        return st, succ.astype(bool), ends

    def raytrace_noalloc(self, origins, dirs, lengths):
        """raytrace through the allocation-free entry point (static device scratch,
        at most NOALLOC_MAX_QUERIES queries, no map updates waiting)."""
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        ln = np.ascontiguousarray(lengths, dtype=np.float64).reshape(-1)
        n = o.shape[0]
        succ = np.zeros(n, np.uint8)
        ends = np.zeros((n, 3), np.float64)
        st = lib().vdbgpu_raytrace_noalloc(self.h, n, _ptr(o), _ptr(d), _ptr(ln), _ptr(succ), _ptr(ends))
        return st, succ.astype(bool), ends

    def raytrace_device(self, n: int, d_origins: int, d_dirs: int, d_lens: int, d_success: int,
                        d_ends: int) -> int:
        return lib().vdbgpu_raytrace_device(self.h, n, ctypes.c_void_p(d_origins),
                                            ctypes.c_void_p(d_dirs), ctypes.c_void_p(d_lens),
                                            ctypes.c_void_p(d_success), ctypes.c_void_p(d_ends))

    # asynchronous ingestion (addDataToAccumulate + worker/integration threads)
    def async_start(self, accumulation_period_s: float) -> int:
        return lib().vdbgpu_async_start(self.h, accumulation_period_s)

    def add_data_to_accumulate(self, source_id: str, points, origin, max_rate: float = 0.0) -> int:
        pts, stride = self._pts(points)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        return lib().vdbgpu_async_submit(self.h, source_id.encode(), _ptr(pts), pts.shape[0],
                                         stride, _ptr(o), max_rate)

    def async_flush(self) -> int:
        return lib().vdbgpu_async_flush(self.h)

    def async_counters(self) -> dict:
        out = np.zeros(4, np.int64)
        lib().vdbgpu_async_counters(self.h, _ptr(out))
        return dict(zip(("submitted", "replaced", "accumulated", "integrations"), out.tolist()))

    def async_stop(self) -> int:
        return lib().vdbgpu_async_stop(self.h)

    def stats(self) -> dict:
        s = Stats()
        lib().vdbgpu_get_stats(self.h, ctypes.byref(s))
        return s.as_dict()

    def set_profiling(self, on: bool) -> None:
        lib().vdbgpu_set_profiling(self.h, int(on))

    def set_tuning(self, block=256, steps_per_lane=32, waves_per_cu=16, load_factor=0.5,
                   traversal=0, defer=8) -> int:
        """traversal: 0 = brick walk for plain inserts, 1 = update-set hash traversal;
        defer: brick-walk scans whose map update runs as one pass (1 = every scan alone)."""
        t = Tuning(block, steps_per_lane, waves_per_cu, load_factor, traversal, defer)
        return lib().vdbgpu_set_tuning(self.h, ctypes.byref(t))

    def get_tuning(self) -> dict:
        t = Tuning()
        lib().vdbgpu_get_tuning(self.h, ctypes.byref(t))
        return {n: getattr(t, n) for n, _ in t._fields_ if n != "_pad"}


class DeviceBuffer:
    """Device memory from libvdbgpu (same HIP runtime as the kernels)."""

    def __init__(self, nbytes: int):
        self.ptr = ctypes.c_void_p()
        self.nbytes = int(nbytes)
        rc = lib().vdbgpu_device_alloc(self.nbytes, ctypes.byref(self.ptr))
        if rc != OK:
            raise VdbGpuError(f"device alloc of {nbytes} bytes failed: {status_string(rc)}")

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(max(a.nbytes, 1))
        b.upload(a)
        return b

    def upload(self, a: np.ndarray, offset: int = 0) -> None:
        a = np.ascontiguousarray(a)
        assert offset + a.nbytes <= self.nbytes
        rc = lib().vdbgpu_device_upload(ctypes.c_void_p(self.ptr.value + offset), _ptr(a), a.nbytes)
        if rc != OK:
            raise VdbGpuError("upload failed")

    def download(self, a: np.ndarray, offset: int = 0) -> np.ndarray:
        assert a.flags.c_contiguous and offset + a.nbytes <= self.nbytes
        rc = lib().vdbgpu_device_download(_ptr(a), ctypes.c_void_p(self.ptr.value + offset), a.nbytes)
        if rc != OK:
            raise VdbGpuError("download failed")
        return a

    def copy_from(self, other: "DeviceBuffer", nbytes: int) -> None:
        assert nbytes <= self.nbytes and nbytes <= other.nbytes
        if lib().vdbgpu_device_copy(self.ptr, other.ptr, nbytes) != OK:
            raise VdbGpuError("device copy failed")

    @property
    def address(self) -> int:
        return int(self.ptr.value)

    def free(self) -> None:
        if self.ptr.value:
            lib().vdbgpu_device_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        self.free()


class Server:
    """vdbgpu_server: one device map per robot stream, continuous batching per map."""

    def __init__(self, n_maps: int, resolution: float, cfg: Config, max_range: float,
                 hbm_budget_bytes: int = 0, max_batch: int = 64):
        self.h = ctypes.c_void_p()
        rc = lib().vdbgpu_server_create(n_maps, resolution, ctypes.byref(cfg), max_range, hbm_budget_bytes,
                                        max_batch, ctypes.byref(self.h))
        if rc != OK:
            raise VdbGpuError(f"vdbgpu_server_create failed: {status_string(rc)}")
        self.n_maps, self.resolution = n_maps, resolution

    def submit(self, map_index: int, points, origin) -> int:
        pts = np.ascontiguousarray(points, dtype=np.float32)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        rid = ctypes.c_int64()
        rc = lib().vdbgpu_server_submit(self.h, map_index, _ptr(pts), pts.shape[0], pts.shape[1] * 4, _ptr(o),
                                        ctypes.byref(rid))
        if rc != OK:
            raise VdbGpuError(f"submit failed: {status_string(rc)}")
        return rid.value

    def drain(self) -> int:
        return lib().vdbgpu_server_drain(self.h)

    def records(self):
        n = lib().vdbgpu_server_records(self.h, None, 0)
        arr = (RequestRecord * max(n, 1))()
        got = lib().vdbgpu_server_records(self.h, arr, n)
        return [{k: getattr(arr[i], k) for k, _ in RequestRecord._fields_} for i in range(got)]

    def map(self, i: int) -> "GpuMap":
        return GpuMap(self.resolution, _borrowed_handle=lib().vdbgpu_server_map(self.h, i))

    def device_bytes(self) -> int:
        return int(lib().vdbgpu_server_device_bytes(self.h))

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            lib().vdbgpu_server_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        self.close()


def device_count() -> int:
    return int(lib().vdbgpu_device_count())
