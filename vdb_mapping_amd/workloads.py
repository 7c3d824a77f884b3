"""Named synthetic workloads (scenes, trajectories and scans) over workloads.c.

Every workload is seeded: workload k uses seed 1000 + k and its scan i uses
(1000 + k) * 1_000_003 + i, so tests, the CPU baseline and the GPU bench see
identical bytes. Points are float32 (x, y, z, 1.0) rows in the map frame
(16-byte stride, the pcl::PointXYZ layout).
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libworkloads.so")
BUILD_CMD = ["gcc", "-O2", "-fopenmp", "-fPIC", "-shared", os.path.join(HERE, "workloads.c"),
             "-o", LIB_PATH, "-lm"]

# Mapping parameters shared by the named workloads (prob_thres_* is the
# "clamp 0.12/0.97" of the workload description; the reference's own clamp is
# fixed at logit(0.99)).
PROBS = dict(prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12, prob_thres_max=0.97)

_lib = None


def build() -> None:
    subprocess.run(BUILD_CMD, check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P, D, I32, I64, U64 = (ctypes.c_void_p, ctypes.c_double, ctypes.c_int32, ctypes.c_int64,
                               ctypes.c_uint64)
        L.wl_scene_bytes.restype = I32
        L.wl_uniform_u64.restype = U64
        L.wl_uniform_u64.argtypes = [U64]
        L.wl_scene_room.argtypes = [U64, I32, D, D, P]
        L.wl_trajectory.argtypes = [U64, I32, D, D, D, P]
        L.wl_scene_outdoor.argtypes = [U64, I32, I32, D, P, I32, D, D, P]
        L.wl_scene_indoor.argtypes = [U64, I32, I32, P, I32, D, P]
        L.wl_lidar_scan.argtypes = [P, P, I32, I32, I64, D, D, D, U64, I32, P]
        L.wl_depth_frame.argtypes = [P, P, I32, I32, D, D, D, D, D, D, D, U64, I32, P]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if isinstance(a, np.ndarray) else a


def _threads() -> int:
    return max(1, min(16, os.cpu_count() or 1))


def scan_seed(k: int, i: int) -> int:
    return (1000 + k) * 1_000_003 + i


class Scene:
    def __init__(self):
        self.buf = ctypes.create_string_buffer(lib().wl_scene_bytes())

    @property
    def ptr(self):
        return ctypes.cast(self.buf, ctypes.c_void_p)


def room_scene(seed: int, nbox: int = 8, side_min: float = 0.3, side_max: float = 2.0) -> Scene:
    s = Scene()
    lib().wl_scene_room(seed, nbox, side_min, side_max, s.ptr)
    return s


def trajectory(seed: int, n: int, step: float = 1.0, max_yaw_deg: float = 5.0,
               z: float = 1.8) -> np.ndarray:
    poses = np.zeros((n, 6), np.float64)
    lib().wl_trajectory(seed, n, step, max_yaw_deg, z, _p(poses))
    return poses


def outdoor_scene(seed: int, poses: np.ndarray, nbox: int = 100, ncyl: int = 100,
                  half: float = 100.0, clear: float = 3.0, sky_range: float = 100.0) -> Scene:
    s = Scene()
    poses = np.ascontiguousarray(poses, np.float64)
    lib().wl_scene_outdoor(seed, nbox, ncyl, half, _p(poses), poses.shape[0], clear, sky_range, s.ptr)
    return s


def indoor_scene(seed: int, poses: np.ndarray, nshelf: int = 10, nbox: int = 12,
                 clear: float = 0.6) -> Scene:
    s = Scene()
    poses = np.ascontiguousarray(poses, np.float64)
    lib().wl_scene_indoor(seed, nshelf, nbox, _p(poses), poses.shape[0], clear, s.ptr)
    return s


def lidar_scan(scene: Scene, pose, rings: int, cols: int, npts: int, elev_min: float,
               elev_max: float, sigma: float, seed: int) -> np.ndarray:
    pose = np.ascontiguousarray(pose, np.float64)
    out = np.empty((npts, 4), np.float32)
    lib().wl_lidar_scan(scene.ptr, _p(pose), rings, cols, npts, elev_min, elev_max, sigma, seed,
                        _threads(), _p(out))
    return out


def depth_frame(scene: Scene, pose, width=640, height=480, fx=525.0, fy=525.0, cx=320.0,
                cy=240.0, dmin=0.3, dmax=5.0, k_noise=0.002, seed=0) -> np.ndarray:
    pose = np.ascontiguousarray(pose, np.float64)
    out = np.empty((width * height, 4), np.float32)
    lib().wl_depth_frame(scene.ptr, _p(pose), width, height, fx, fy, cx, cy, dmin, dmax, k_noise,
                         seed, _threads(), _p(out))
    return out


@dataclass
class Scan:
    points: np.ndarray           # (n, 4) float32
    origin: np.ndarray           # (3,) float64
    max_range: float


@dataclass
class Workload:
    name: str
    resolution: float
    max_range: float
    scans: list = field(default_factory=list)   # list[Scan] or a lazy generator
    n_scans: int = 0
    probs: dict = field(default_factory=lambda: dict(PROBS))

    def scan(self, i: int) -> Scan:
        return self._make(i)


def room_scan_workload(n_scans: int | None = None) -> Workload:
    """One 32,768-point scan (32 rings x 1024 columns, elevation [-15, 15] deg)
    from the origin of a 20 x 20 x 5 m room with 8 random boxes; range noise
    0.02 m; resolution 0.1 m, max range 15 m."""
    k = 1
    scene = room_scene(1000 + k)
    wl = Workload("room_scan_32k_0.1m", 0.1, 15.0, n_scans=1 if n_scans is None else n_scans)

    def make(i):
        pts = lidar_scan(scene, np.zeros(6), 32, 1024, 32768, -15.0, 15.0, 0.02, scan_seed(k, i))
        return Scan(pts, np.zeros(3), 15.0)

    wl._make = make
    return wl


def outdoor_stream_workload(n_scans: int = 100) -> Workload:
    """n_scans scans of 262,144 points (128 rings x 2048 columns, elevation
    [-22.5, 22.5] deg) along a 1 m-step trajectory with <= 5 deg yaw drift,
    sensor 1.8 m above a ground plane with 100 boxes and 100 vertical
    cylinders in a 200 x 200 m area; range noise 0.03 m; rays that hit nothing
    report 100 m. Resolution 0.05 m, max range 50 m."""
    k = 2
    poses = trajectory(1000 + k, 100, 1.0, 5.0, 1.8)
    scene = outdoor_scene(1000 + k, poses)
    wl = Workload("outdoor_stream_262k_0.05m", 0.05, 50.0, n_scans=n_scans)
    wl.poses = poses

    def make(i):
        pts = lidar_scan(scene, poses[i], 128, 2048, 262144, -22.5, 22.5, 0.03, scan_seed(k, i))
        return Scan(pts, poses[i, :3].copy(), 50.0)

    wl._make = make
    return wl


def mixed_batch_workload(n_scans: int = 64) -> Workload:
    """n_scans scans in the room scene with point counts log-uniform in
    [10k, 300k] (64 rings, elevation [-22.5, 22.5] deg), random 6-DoF poses
    (yaw uniform, roll and pitch within 3 deg) and max range uniform in
    [20, 50] m; resolution 0.05 m."""
    k = 3
    scene = room_scene(1000 + k - 2)  # the room scan's scene (seed 1001)
    rng = np.random.Generator(np.random.PCG64(1000 + k))
    counts = np.exp(rng.uniform(math.log(10_000), math.log(300_000), n_scans)).astype(np.int64)
    ranges = rng.uniform(20.0, 50.0, n_scans)
    poses = np.zeros((n_scans, 6))
    poses[:, 0] = rng.uniform(-8.0, 8.0, n_scans)
    poses[:, 1] = rng.uniform(-8.0, 8.0, n_scans)
    poses[:, 2] = rng.uniform(-0.5, 2.0, n_scans)
    poses[:, 3] = np.radians(rng.uniform(-3.0, 3.0, n_scans))
    poses[:, 4] = np.radians(rng.uniform(-3.0, 3.0, n_scans))
    poses[:, 5] = rng.uniform(-math.pi, math.pi, n_scans)
    wl = Workload("mixed_batch_64_0.05m", 0.05, 50.0, n_scans=n_scans)
    wl.counts, wl.ranges, wl.poses = counts, ranges, poses

    def make(i):
        n = int(counts[i])
        cols = (n + 63) // 64
        pts = lidar_scan(scene, poses[i], 64, cols, n, -22.5, 22.5, 0.02, scan_seed(k, i))
        return Scan(pts, poses[i, :3].copy(), float(ranges[i]))

    wl._make = make
    return wl


def rgbd_workload(n_frames: int = 200) -> Workload:
    """n_frames 640 x 480 pinhole depth frames (fx = fy = 525, cx = 320,
    cy = 240), depth in [0.3, 5] m, noise 0.002 d^2, in an 8 x 8 x 3 m room
    with shelves and boxes; resolution 0.02 m, max range 5 m."""
    k = 4
    rng = np.random.Generator(np.random.PCG64(1000 + k))
    poses = np.zeros((n_frames, 6))
    t = np.linspace(0, 2 * math.pi, n_frames, endpoint=False)
    poses[:, 0] = 1.5 * np.cos(t) + rng.normal(0, 0.05, n_frames)
    poses[:, 1] = 1.5 * np.sin(t) + rng.normal(0, 0.05, n_frames)
    poses[:, 2] = 1.2 + rng.normal(0, 0.05, n_frames)
    poses[:, 4] = np.radians(rng.uniform(-10, 10, n_frames))
    poses[:, 5] = t + math.pi / 2 + rng.normal(0, 0.2, n_frames)
    scene = indoor_scene(1000 + k, poses)
    wl = Workload("rgbd_640x480_0.02m", 0.02, 5.0, n_scans=n_frames)
    wl.poses = poses

    def make(i):
        pts = depth_frame(scene, poses[i], seed=scan_seed(k, i))
        return Scan(pts, poses[i, :3].copy(), 5.0)

    wl._make = make
    return wl


def multi_robot_pool(n_scans: int = 48) -> Workload:
    """The scans of the multi-robot serving workload: n_scans scans in the room
    scene (seed 1001) with point counts log-uniform in [20k, 262,144] (64
    rings, elevation [-22.5, 22.5] deg), random poses (x, y in [-8, 8] m, z in
    [-0.5, 2] m, roll and pitch within 3 deg, yaw uniform); resolution 0.05 m,
    max range 50 m. Robot streams draw their scans from this pool."""
    k = 5
    scene = room_scene(1001)
    rng = np.random.default_rng(1000 + k)
    counts, poses = [], []
    for _ in range(n_scans):
        counts.append(int(math.exp(rng.uniform(math.log(20_000), math.log(262_144)))))
        poses.append([rng.uniform(-8, 8), rng.uniform(-8, 8), rng.uniform(-0.5, 2.0),
                      math.radians(rng.uniform(-3, 3)), math.radians(rng.uniform(-3, 3)),
                      rng.uniform(-math.pi, math.pi)])
    wl = Workload("multi_robot_room_0.05m", 0.05, 50.0, n_scans=n_scans)
    wl.counts, wl.poses = np.array(counts, np.int64), np.array(poses, np.float64)

    def make(i):
        n = int(wl.counts[i])
        pts = lidar_scan(scene, wl.poses[i], 64, (n + 63) // 64, n, -22.5, 22.5, 0.02, scan_seed(k, i))
        return Scan(pts, wl.poses[i, :3].copy(), 50.0)

    wl._make = make
    return wl


def multi_robot_outdoor(robot: int, n_scans: int) -> Workload:
    """Scans of robot `robot` in the large-map serving workload: its own outdoor
    world (ground plane, 100 boxes and 100 cylinders in 200 x 200 m around its
    path) and a 1 m-step trajectory with <= 5 deg yaw drift; point counts
    log-uniform in [20k, 262,144] (64 rings, elevation [-22.5, 22.5] deg),
    range noise 0.03 m; resolution 0.05 m, max range 50 m. Its map grows along
    the path, by several GB over a hundred scans."""
    k = 5
    seed = (1000 + k) * 1000 + robot
    poses = trajectory(seed, n_scans, 1.0, 5.0, 1.8)
    scene = outdoor_scene(seed, poses)
    rng = np.random.default_rng(seed)
    counts = np.exp(rng.uniform(math.log(20_000), math.log(262_144), n_scans)).astype(np.int64)
    wl = Workload("multi_robot_outdoor_0.05m", 0.05, 50.0, n_scans=n_scans)
    wl.counts, wl.poses = counts, poses

    def make(i):
        n = int(counts[i])
        pts = lidar_scan(scene, poses[i], 64, (n + 63) // 64, n, -22.5, 22.5, 0.03,
                         seed * 1_000_003 + i)
        return Scan(pts, poses[i, :3].copy(), 50.0)

    wl._make = make
    return wl


def multi_robot_bursts(n_maps: int = 16, repeats: int = 5):
    """Request order of the serving test: map i receives pool scan 3i `repeats`
    times (a robot standing still: voxels hit that often turn active at the
    workload's probabilities), then pool scans 3i + 1 and 3i + 2, all as one
    burst per map. Returns [(map, pool index)] in submission order."""
    return [(i, k) for i in range(n_maps) for k in [3 * i] * repeats + [3 * i + 1, 3 * i + 2]]


def raytrace_queries(n: int, seed: int = 1004 * 1_000_003 + 999_999, half=3.5, zlo=0.3, zhi=2.5,
                     length=5.0):
    """n query rays: origins uniform in the RGB-D room interior, unit
    directions uniform on the sphere, max length `length`."""
    rng = np.random.Generator(np.random.PCG64(seed))
    o = np.stack([rng.uniform(-half, half, n), rng.uniform(-half, half, n),
                  rng.uniform(zlo, zhi, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return o, d, np.full(n, length)
