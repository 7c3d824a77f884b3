"""ctypes wrapper of the CPU oracle (oracle/oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, always as the checker or the baseline,
never as the measured or shipped path.  The product package
(vdb_mapping_amd) does not import it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")

OK, NOT_CONFIGURED, UNKNOWN_SOURCE, BAD_MAX_RANGE, BAD_PROB_MISS, BAD_PROB_HIT = 0, 1, 2, 3, 4, 5
OUT_OF_RANGE, UNSUPPORTED, BAD_ARGUMENT, NO_INTERSECTOR = 6, 7, 9, 10

BUILD_CMD = [
    "g++", "-O2", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared",
    os.path.join(HERE, "oracle.cpp"), "-o", LIB_PATH,
]


def build() -> None:
    """Compile liborc.so from oracle.cpp (g++, no FMA contraction)."""
    subprocess.run(BUILD_CMD, check=True)


class Config(ctypes.Structure):
    """BaseConfig (VDBMapping.hpp:64-71) + Config (OccupancyVDBMapping.hpp:38-44)."""

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


_lib = None


def _bind(L):
    """Declare the argument and result types of every exported function on L."""
    P, D, I64, I32 = ctypes.c_void_p, ctypes.c_double, ctypes.c_int64, ctypes.c_int32
    L.orc_create.restype = P
    L.orc_create.argtypes = [D]
    L.orc_destroy.argtypes = [P]
    L.orc_set_config.argtypes = [P, ctypes.POINTER(Config)]
    L.orc_add_source.argtypes = [P, ctypes.c_char_p, D, D]
    for f in (L.orc_accumulate, L.orc_insert):
        f.argtypes = [P, ctypes.c_char_p, P, I64, I64, P]
    L.orc_integrate.argtypes = [P]
    L.orc_reset.argtypes = [P]
    L.orc_empty.argtypes = [P]
    L.orc_logodds.argtypes = [P, P]
    L.orc_leaf_count.restype = I64
    L.orc_leaf_count.argtypes = [P]
    L.orc_export_leaves.restype = I64
    L.orc_export_leaves.argtypes = [P, P, P, P, I64]
    L.orc_update_leaf_count.restype = I64
    L.orc_update_leaf_count.argtypes = [P, ctypes.c_char_p]
    L.orc_export_update.restype = I64
    L.orc_export_update.argtypes = [P, ctypes.c_char_p, P, P, P, I64]
    L.orc_get_value.restype = ctypes.c_float
    L.orc_get_value.argtypes = [P, P, P]
    L.orc_world_to_index.argtypes = [D, P, P]
    L.orc_cast_ray.restype = I64
    L.orc_cast_ray.argtypes = [P, P, P, I64]
    L.orc_cast_ray_fast.restype = I64
    L.orc_cast_ray_fast.argtypes = [P, P, P, P, I64]
    L.orc_raytrace.argtypes = [P, I64, P, P, P, P, P]
    L.orc_set_points.argtypes = [P, P, I64, I64, I32]
    L.orc_add_artificial_areas.argtypes = [P, I32, P, P, D, D]
    L.orc_restore_map_integrity.argtypes = [P]
    L.orc_add_artificial_walls.argtypes = [P, I32, P, D, D]
    L.orc_export_artificial.restype = I64
    L.orc_export_artificial.argtypes = [P, P, P, I64]
    L.orc_export_section.restype = I64
    L.orc_export_section.argtypes = [P, P, I32, P, P, P, I64]
    L.orc_apply_section.argtypes = [P, I64, P, P, P]
    L.orc_apply_section_update.argtypes = [P, P, I64, P, P]
    L.orc_index_bbox.argtypes = [D, P, P, P, P]
    L.orc_export_active_points.restype = I64
    L.orc_export_active_points.argtypes = [P, P, I64]
    L.orc_create_from_points.argtypes = [P, P, I64, I64, I32, I32]
    L.orc_export_tiles.restype = I64
    L.orc_export_tiles.argtypes = [P, P, P, P, I64]
    L.orc_morphology.restype = I64
    L.orc_morphology.argtypes = [I32, I32, I64, P, P, P, P, I64]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = _bind(ctypes.CDLL(LIB_PATH))
    return _lib


def mutant(replacements, defines=(), workdir=None):
    """A copy of the oracle with planted mistakes, for tests that check that a
    brute-force reference notices them: each (old, new) pair of `replacements`
    must occur exactly once in oracle.cpp; `defines` are extra -D flags.
    Compiled with the oracle's own flags into `workdir` (a temporary directory
    by default); returns the bound library, to pass as OracleMap(..., L=...)."""
    import hashlib
    import tempfile
    src = open(os.path.join(HERE, "oracle.cpp")).read()
    for old, new in replacements:
        if src.count(old) != 1:
            raise ValueError(f"mutation target must occur once in oracle.cpp: {old!r}")
        src = src.replace(old, new)
    tag = hashlib.sha1((src + repr(tuple(defines))).encode()).hexdigest()[:16]
    d = workdir or os.path.join(tempfile.gettempdir(), "orc_mutants")
    os.makedirs(d, exist_ok=True)
    cpp = os.path.join(d, f"m_{tag}.cpp")
    so = os.path.join(d, f"m_{tag}.so")
    if not os.path.exists(so):
        with open(cpp, "w") as f:
            f.write(src)
        cmd = [c for c in BUILD_CMD[:-3]] + ["-I" + HERE, *["-D" + x for x in defines], cpp, "-o", so + ".tmp"]
        subprocess.run(cmd, check=True)
        os.replace(so + ".tmp", so)
    return _bind(ctypes.CDLL(so))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def world_to_index(res: float, w) -> np.ndarray:
    w = np.ascontiguousarray(w, dtype=np.float64)
    out = np.zeros(3, np.int32)
    lib().orc_world_to_index(res, _ptr(w), _ptr(out))
    return out


def cast_ray(s, e) -> np.ndarray:
    s = np.ascontiguousarray(s, dtype=np.int32)
    e = np.ascontiguousarray(e, dtype=np.int32)
    n = lib().orc_cast_ray(_ptr(s), _ptr(e), None, 0)
    out = np.zeros((max(n, 1), 3), np.int32)
    lib().orc_cast_ray(_ptr(s), _ptr(e), _ptr(out), n)
    return out[:n]


def index_bbox(res: float, min_boundary, max_boundary, tf=None) -> np.ndarray:
    """createIndexBoundingBox: inclusive index box (min xyz, max xyz)."""
    mn = np.ascontiguousarray(min_boundary, dtype=np.float64)
    mx = np.ascontiguousarray(max_boundary, dtype=np.float64)
    t = np.ascontiguousarray(np.eye(4) if tf is None else tf, dtype=np.float64)
    out = np.zeros(6, np.int32)
    lib().orc_index_bbox(res, _ptr(mn), _ptr(mx), _ptr(t), _ptr(out))
    return out


MORPH_OPS = {"dilate": 0, "erode": 1, "close": 2, "open": 3}


def morphology(op: str, iterations: int, origins, masks):
    """Morphology of a leaf set (VM:1095-1148): op in MORPH_OPS, 26-neighbour
    steps. Returns (origins (k, 3) int32, masks (k, 8) uint64): the input
    leaves in input order, then the leaves dilation created, sorted by origin."""
    o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
    m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, 8)
    n = len(o)
    k = lib().orc_morphology(MORPH_OPS[op], iterations, n, _ptr(o), _ptr(m), None, None, 0)
    if k < 0:
        raise ValueError("bad morphology input")
    oo = np.zeros((max(k, 1), 3), np.int32)
    om = np.zeros((max(k, 1), 8), np.uint64)
    lib().orc_morphology(MORPH_OPS[op], iterations, n, _ptr(o), _ptr(m), _ptr(oo), _ptr(om), k)
    return oo[:k], om[:k]


def polygons_arg(polygons):
    """Flatten a list of (k, 3+) vertex arrays into (counts int32, xyz float64)."""
    counts = np.array([len(p) for p in polygons], np.int32)
    rows = [np.asarray(p, np.float64).reshape(len(p), -1)[:, :3] for p in polygons if len(p)]
    xyz = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros((0, 3)), dtype=np.float64)
    return counts, xyz


class OracleMap:
    """OccupancyVDBMapping restated on the CPU (same method meaning as the reference)."""

    def __init__(self, resolution: float, L=None):
        self.L = L or lib()
        self.h = self.L.orc_create(resolution)
        self.resolution = resolution

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_config(self, cfg: Config) -> int:
        return self.L.orc_set_config(self.h, ctypes.byref(cfg))

    def add_input_source(self, source_id: str, max_range: float, max_rate: float = 0.0) -> int:
        return self.L.orc_add_source(self.h, source_id.encode(), max_range, max_rate)

    @staticmethod
    def _pts(points):
        pts = np.ascontiguousarray(points, dtype=np.float32)
        if pts.ndim != 2 or pts.shape[1] not in (3, 4):
            raise ValueError("points must be (n,3) or (n,4) float32")
        return pts, pts.shape[1] * 4

    def accumulate(self, source_id: str, points, origin) -> int:
        pts, stride = self._pts(points)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        return self.L.orc_accumulate(self.h, source_id.encode(), _ptr(pts), pts.shape[0], stride,
                                    _ptr(o))

    def integrate(self) -> int:
        return self.L.orc_integrate(self.h)

    def insert(self, source_id: str, points, origin) -> int:
        pts, stride = self._pts(points)
        o = np.ascontiguousarray(origin, dtype=np.float64)
        return self.L.orc_insert(self.h, source_id.encode(), _ptr(pts), pts.shape[0], stride,
                                _ptr(o))

    def reset(self) -> None:
        self.L.orc_reset(self.h)

    def empty(self) -> bool:
        return bool(self.L.orc_empty(self.h))

    def logodds(self) -> np.ndarray:
        out = np.zeros(6, np.float32)
        self.L.orc_logodds(self.h, _ptr(out))
        return out

    def export_leaves(self):
        n = self.L.orc_leaf_count(self.h)
        o = np.zeros((n, 3), np.int32)
        v = np.zeros((n, 512), np.float32)
        m = np.zeros((n, 8), np.uint64)
        got = self.L.orc_export_leaves(self.h, _ptr(o), _ptr(v), _ptr(m), n)
        assert got == n
        return o, v, m

    def export_tiles(self):
        """Leaf-level tiles of pruneGrid: (origins (k, 3) int32, values (k,) float32, active (k,) bool)."""
        n = self.L.orc_export_tiles(self.h, None, None, None, 0)
        o = np.zeros((n, 3), np.int32)
        v = np.zeros(n, np.float32)
        a = np.zeros(n, np.uint8)
        assert self.L.orc_export_tiles(self.h, _ptr(o), _ptr(v), _ptr(a), n) == n
        return o, v, a.astype(bool)

    def export_update(self, source_id: str):
        n = self.L.orc_update_leaf_count(self.h, source_id.encode())
        if n < 0:
            raise KeyError(source_id)
        o = np.zeros((n, 3), np.int32)
        vis = np.zeros((n, 8), np.uint64)
        hit = np.zeros((n, 8), np.uint64)
        got = self.L.orc_export_update(self.h, source_id.encode(), _ptr(o), _ptr(vis), _ptr(hit), n)
        assert got == n
        return o, vis, hit

    def get_value(self, ijk):
        c = np.ascontiguousarray(ijk, dtype=np.int32)
        act = np.zeros(1, np.int32)
        v = self.L.orc_get_value(self.h, _ptr(c), _ptr(act))
        return np.float32(v), bool(act[0])

    # map edits and sections
    def set_points(self, points, occupied: bool) -> int:
        pts, stride = self._pts(points)
        return self.L.orc_set_points(self.h, _ptr(pts), pts.shape[0], stride, int(bool(occupied)))

    def add_points_to_grid(self, points) -> int:
        return self.set_points(points, True)

    def remove_points_from_grid(self, points) -> int:
        return self.set_points(points, False)

    def add_artificial_areas(self, polygons, negative_height: float, positive_height: float) -> int:
        counts, xyz = polygons_arg(polygons)
        return self.L.orc_add_artificial_areas(self.h, len(counts), _ptr(counts), _ptr(xyz),
                                              negative_height, positive_height)

    def add_artificial_wall(self, start, end, negative_height: float, positive_height: float) -> int:
        return self.add_artificial_walls([(start, end)], negative_height, positive_height)

    def add_artificial_walls(self, walls, negative_height: float, positive_height: float) -> int:
        w = np.ascontiguousarray([[*np.asarray(a, float)[:3], *np.asarray(b, float)[:3]] for a, b in walls],
                                 dtype=np.float64).reshape(-1, 6)
        return self.L.orc_add_artificial_walls(self.h, w.shape[0], _ptr(w), negative_height, positive_height)

    def add_artificial_polygon(self, polygon, negative_height: float, positive_height: float) -> int:
        p = [np.asarray(v, float)[:3] for v in polygon]
        walls = [(p[i], p[(i + 1) % len(p)]) for i in range(len(p))]
        return self.add_artificial_walls(walls, negative_height, positive_height)

    def restore_map_integrity(self) -> int:
        return self.L.orc_restore_map_integrity(self.h)

    def export_artificial(self):
        n = self.L.orc_export_artificial(self.h, None, None, 0)
        o = np.zeros((n, 3), np.int32)
        m = np.zeros((n, 8), np.uint64)
        assert self.L.orc_export_artificial(self.h, _ptr(o), _ptr(m), n) == n
        return o, m

    def export_section(self, bbox, full_grid: bool):
        b = np.ascontiguousarray(bbox, dtype=np.int32)
        n = self.L.orc_export_section(self.h, _ptr(b), int(bool(full_grid)), None, None, None, 0)
        o = np.zeros((n, 3), np.int32)
        v = np.zeros((n, 512), np.float32)
        m = np.zeros((n, 8), np.uint64)
        got = self.L.orc_export_section(self.h, _ptr(b), int(bool(full_grid)), _ptr(o), _ptr(v),
                                       _ptr(m), n)
        assert got == n
        return o, v, m

    def apply_section(self, origins, values, masks) -> int:
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1, 512)
        m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, 8)
        return self.L.orc_apply_section(self.h, o.shape[0], _ptr(o), _ptr(v), _ptr(m))

    def apply_section_update(self, bbox, origins, masks) -> int:
        b = np.ascontiguousarray(bbox, dtype=np.int32)
        o = np.ascontiguousarray(origins, dtype=np.int32).reshape(-1, 3)
        m = np.ascontiguousarray(masks, dtype=np.uint64).reshape(-1, 8)
        return self.L.orc_apply_section_update(self.h, _ptr(b), o.shape[0], _ptr(o), _ptr(m))

    def export_active_points(self) -> np.ndarray:
        n = self.L.orc_export_active_points(self.h, None, 0)
        out = np.zeros((n, 3), np.float32)
        assert self.L.orc_export_active_points(self.h, _ptr(out), n) == n
        return out

    def create_map_from_points(self, points, set_background: bool, clear_map: bool) -> int:
        pts, stride = self._pts(points)
        return self.L.orc_create_from_points(self.h, _ptr(pts), pts.shape[0], stride, int(bool(set_background)),
                                            int(bool(clear_map)))

    def raytrace(self, origins, dirs, lengths):
        o = np.ascontiguousarray(origins, dtype=np.float64).reshape(-1, 3)
        d = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        ln = np.ascontiguousarray(lengths, dtype=np.float64).reshape(-1)
        n = o.shape[0]
        succ = np.zeros(n, np.uint8)
        ends = np.zeros((n, 3), np.float64)
        st = self.L.orc_raytrace(self.h, n, _ptr(o), _ptr(d), _ptr(ln), _ptr(succ), _ptr(ends))
        return st, succ.astype(bool), ends

    def cast_ray_fast(self, s, e) -> np.ndarray:
        """castRayIntoGridFast (VM:577-602) from voxel s to voxel e: the marked map
        voxels in visiting order, (k, 3) int32. Needs the fast-mode intersector."""
        s = np.ascontiguousarray(s, dtype=np.int32)
        e = np.ascontiguousarray(e, dtype=np.int32)
        n = self.L.orc_cast_ray_fast(self.h, _ptr(s), _ptr(e), None, 0)
        if n < 0:
            raise RuntimeError("no intersector: fast mode needs an integrate on a non-empty map")
        out = np.zeros((max(n, 1), 3), np.int32)
        self.L.orc_cast_ray_fast(self.h, _ptr(s), _ptr(e), _ptr(out), n)
        return out[:n]
