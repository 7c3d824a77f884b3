"""GPU parity: the HIP path (through the C ABI) against the CPU oracle, bit for bit.

Visited sets, hit flags, leaf topology, log-odds values (compared as uint32
bit patterns, tolerance zero) and active masks must be identical.
"""
import math

import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from parity import both, pair, same_map, same_update
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def logit32(p):
    return np.float32(math.log(p) - math.log(1 - p))


# ---------------------------------------------- reference known answers --

def test_known_answer_set_config():
    """tests/mapping.cpp:6-30 through the GPU path."""
    g = V.GpuMap(1.0)
    pts = np.array([[0, 0, 1]], np.float32)
    assert g.insert("test", pts, [0, 0, 0]) == V.UNKNOWN_SOURCE
    assert g.leaf_count() == 0
    assert g.set_config(V.make_config()) == V.OK
    g.add_input_source("test", 10.0, 0)
    assert g.insert("test", pts, [0, 0, 0]) == V.OK
    o, v, m = g.export_leaves()
    assert o.tolist() == [[0, 0, 0]]
    assert v[0, 0] == logit32(0.1) and v[0, 1] == logit32(0.9)
    assert int(m[0, 0]) == 0b10


@pytest.mark.parametrize("z,R", [(0.5, 10.0), (-0.5, 10.0), (0.7, 0.5)])
def test_known_answer_insert(z, R):
    """tests/mapping.cpp:32-131 (positive, negative and max-range point)."""
    g, o = pair(0.1, max_range=R)
    pts = np.array([[0, 0, z]], np.float32)
    both(g, o, "insert", "test", pts, [0, 0, 0])
    same_map(g, o)


def test_known_answer_reset():
    g, o = pair(1.0)
    both(g, o, "insert", "test", np.array([[0, 0, 1]], np.float32), [0, 0, 0])
    same_map(g, o)
    g.reset()
    o.reset()
    assert g.leaf_count() == 0 and g.empty()
    both(g, o, "insert", "test", np.array([[0, 0, 2]], np.float32), [0, 0, 0])
    same_map(g, o)


# ------------------------------------------------------- random parity --

@pytest.mark.parametrize("seed", range(6))
def test_random_scans(seed):
    rng = np.random.default_rng(100 + seed)
    res = [0.1, 0.05, 0.2, 0.5, 0.02, 1.0][seed]
    R = [5.0, 3.0, 12.0, 30.0, 1.0, 40.0][seed]
    g, o = pair(res, max_range=R, prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12,
                prob_thres_max=0.97)
    for scan in range(4):
        origin = rng.uniform(-2, 2, 3)
        n = int(rng.integers(1, 3000))
        pts = (origin + rng.normal(0, R * 0.7, (n, 3))).astype(np.float32)
        pts[rng.random(n) < 0.02] = np.nan
        both(g, o, "accumulate", "test", pts, origin)
        same_update(g, o, "test")
        both(g, o, "integrate")
        same_map(g, o)


TIE_ENDS = [(a, a, 0) for a in (1, 2, 7, 64, 333)] + [(a, a, a) for a in (1, 3, 8, 100, 999)] + \
    [(a, 2 * a, 0) for a in (1, 5, 50)] + [(3, 1, 0), (5, 15, 0), (15, 5, 3), (30, 10, 6),
                                           (1000, 999, 1), (7, -7, 7), (-999, 333, -111)]


@pytest.mark.parametrize("steps", [1, 2, 3, 7, 64, 1000])
def test_mid_ray_entry_tie_rays(steps):
    """Every chunk length forces mid-ray entry at different visits; with
    steps_per_lane = 1 every visit of every ray is entered from the crossing
    table. Rays with exact crossing ties (diagonals, rational multiples)."""
    g, o = pair(1.0, max_range=5000.0)
    assert g.set_tuning(block=256, steps_per_lane=steps, waves_per_cu=2, load_factor=0.5) == V.OK
    ends = []
    for e in TIE_ENDS:
        for sx in (1, -1):
            for sy in (1, -1):
                ends.append((sx * e[0] + 0.25, sy * e[1] + 0.25, e[2] + 0.25))
    pts = np.array(ends, np.float32)
    both(g, o, "accumulate", "test", pts, [0.25, 0.25, 0.25])
    same_update(g, o, "test")
    both(g, o, "integrate")
    same_map(g, o)


@pytest.mark.parametrize("tuning", [(64, 16, 1, 0.7), (128, 32, 4, 0.5), (256, 256, 16, 0.5),
                                    (192, 4096, 8, 0.3)])
def test_tuning_does_not_change_results(tuning):
    w = W.room_scan_workload()
    s = w.scan(0)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    assert g.set_tuning(*tuning, traversal=1) == V.OK  # the launch shape of the hash traversal
    assert g.set_tuning(512, 128, 8, 0.5) == V.BAD_ARGUMENT  # above the kernel's launch bound
    both(g, o, "insert", "test", s.points[::7], s.origin)
    same_map(g, o)


def test_grid_aligned_and_half_voxel_points():
    """Coordinates on exact multiples of the resolution (fmod == 0, no shift)
    and on half-voxel boundaries exercise the rounding branch of worldToIndex."""
    res = 0.05
    g, o = pair(res, max_range=20.0)
    k = np.arange(-200, 201)
    pts = []
    for x in k:
        pts.append((x * res, 3 * res, -2 * res))
        pts.append(((x + 0.5) * res, (x % 7) * res, 0.5 * res))
        pts.append((np.float32(x * 0.1), np.float32(x * -0.3), np.float32(1.25)))
    pts = np.array(pts, np.float32)
    both(g, o, "accumulate", "test", pts, [0.0, 0.0, 0.0])
    same_update(g, o, "test")
    both(g, o, "integrate")
    same_map(g, o)


def test_multi_source_and_or_merge():
    g, o = pair(0.1, max_range=6.0, sources=("a", "b", "c"), source_ranges=(6.0, 0.0, 2.5))
    rng = np.random.default_rng(9)
    for src in ("b", "a", "c", "a"):
        origin = rng.uniform(-1, 1, 3)
        pts = (origin + rng.normal(0, 3, (500, 3))).astype(np.float32)
        both(g, o, "accumulate", src, pts, origin)
    for src in ("a", "b", "c"):
        same_update(g, o, src)
    both(g, o, "integrate")
    same_map(g, o)


def test_update_set_growth():
    """Few long rays: the initial update set is far too small, so it must grow
    (rehash + re-run of the traversal) and still match."""
    g, o = pair(0.02, max_range=60.0)
    rng = np.random.default_rng(4)
    d = rng.normal(size=(40, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = (d * 55.0).astype(np.float32)
    both(g, o, "insert", "test", pts, [0, 0, 0])
    same_map(g, o)


def test_out_of_range_rejects_scan():
    g, o = pair(1.0, max_range=0.0)
    g.add_input_source("far", 1e8)
    o.add_input_source("far", 1e8)
    pts = np.array([[1, 2, 3], [9.0e6, 0, 0]], np.float32)
    assert g.accumulate("far", pts, [0, 0, 0]) == V.OUT_OF_RANGE
    assert o.accumulate("far", pts, [0, 0, 0]) == O.OUT_OF_RANGE
    assert g.export_update("far")[0].shape[0] == 0
    both(g, o, "accumulate", "far", pts[:1], [0, 0, 0])
    same_update(g, o, "far")


def test_nan_origin_and_zero_range_source():
    g, o = pair(0.1, max_range=5.0, sources=("s", "z"), source_ranges=(5.0, -1.0))
    pts = np.array([[1, 1, 1]], np.float32)
    both(g, o, "insert", "s", pts, [np.nan, 0, 0])
    both(g, o, "insert", "z", pts, [0, 0, 0])  # max_range <= 0: nothing is cast
    assert g.leaf_count() == 0 and g.empty()


def test_device_resident_points():
    w = W.room_scan_workload()
    s = w.scan(0)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    dev = V.DeviceBuffer.from_array(s.points)
    back = dev.download(np.empty_like(s.points))
    assert back.tobytes() == s.points.tobytes()
    assert g.insert_device("test", dev.address, s.points.shape[0], 16, s.origin) == V.OK
    assert o.insert("test", s.points, s.origin) == O.OK
    same_map(g, o)


def test_dirty_leaf_export():
    g, o = pair(0.1, max_range=5.0)
    rng = np.random.default_rng(2)
    pts = rng.normal(0, 2, (400, 3)).astype(np.float32)
    both(g, o, "insert", "test", pts, [0, 0, 0])
    d_o, d_v, d_m = g.export_dirty_leaves()
    f_o, f_v, f_m = g.export_leaves()
    assert np.array_equal(d_o, f_o) and np.array_equal(d_v, f_v) and np.array_equal(d_m, f_m)
    assert g.export_dirty_leaves()[0].shape[0] == 0
    both(g, o, "insert", "test", np.array([[0, 0, 0.35]], np.float32), [0, 0, 0])
    d_o, _, _ = g.export_dirty_leaves()
    assert d_o.tolist() == [[0, 0, 0]]
    same_map(g, o)


# ------------------------------------------------------- named workloads --

def test_room_scan_workload_bit_exact():
    """32,768-point room scan at 0.1 m, max range 15 m."""
    w = W.room_scan_workload()
    s = w.scan(0)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    both(g, o, "accumulate", "test", s.points, s.origin)
    n_upd = same_update(g, o, "test")
    both(g, o, "integrate")
    n_leaves = same_map(g, o)
    assert n_upd > 3000 and n_leaves == n_upd
    st = g.stats()
    assert st["rays"] == 32768 and st["voxel_updates"] > 1_000_000


@pytest.mark.slow
def test_outdoor_stream_two_scans_bit_exact():
    """Two 262,144-point scans of the outdoor stream at 0.05 m, max range 50 m."""
    w = W.outdoor_stream_workload(2)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    for i in range(2):
        s = w.scan(i)
        both(g, o, "accumulate", "test", s.points, s.origin)
        if i == 0:
            same_update(g, o, "test")
        both(g, o, "integrate")
    assert same_map(g, o) > 1_000_000
