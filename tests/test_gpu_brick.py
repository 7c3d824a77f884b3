"""The brick path of insertPointCloud (rays cut into per-32^3-brick segments,
one workgroup per brick bitmap) against the oracle and against the update-set
hash traversal of the same library, bit for bit."""
import numpy as np
import pytest

import vdb_mapping_amd as V
from parity import both, pair, same_map
from test_gpu_parity import TIE_ENDS
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def exports_equal(a, b):
    ka, va, ma = a.export_leaves()
    kb, vb, mb = b.export_leaves()
    assert np.array_equal(ka, kb)
    assert np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    assert np.array_equal(ma, mb)
    return ka.shape[0]


def two_paths(res, max_range, **probs):
    maps = []
    for traversal in (0, 1):
        g = V.GpuMap(res)
        g.set_config(V.make_config(max_range=max_range, **probs))
        g.add_input_source("s", max_range)
        assert g.set_tuning(traversal=traversal) == V.OK
        maps.append(g)
    return maps


def test_room_scan_brick_equals_hash_and_oracle():
    """32,768 rays from one origin: the origin's bricks hold far more than one
    work item of segments, so they are split across workgroups."""
    w = W.room_scan_workload()
    s = w.scan(0)
    gb, gh = two_paths(w.resolution, w.max_range, **W.PROBS)
    for g in (gb, gh):
        assert g.insert("s", s.points, s.origin) == V.OK
    assert exports_equal(gb, gh) > 3000
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    both(g, o, "insert", "test", s.points, s.origin)
    same_map(g, o)
    st = g.stats()
    assert st["rays"] == 32768 and st["n_walk"] == 0  # profiling off


@pytest.mark.slow
def test_outdoor_stream_insert_brick_equals_hash():
    """Three 262,144-point outdoor scans through insert on both paths."""
    w = W.outdoor_stream_workload(3)
    gb, gh = two_paths(w.resolution, w.max_range, **W.PROBS)
    for i in range(3):
        s = w.scan(i)
        for g in (gb, gh):
            assert g.insert("s", s.points, s.origin) == V.OK
    assert exports_equal(gb, gh) > 1_000_000
    assert gb.stats()["visits"] == gh.stats()["visits"]


def test_scan_sizes_grow_and_shrink_the_cut_tables():
    """Scans of 600,000, 100,000 and 1,200,000 rays (2,344, 391 and 4,688
    workgroups of the brick cut, whose per-workgroup brick tables are stored by
    the count and reread by the fill): the tables grow past their first size
    and are reused by a smaller scan; brick path = hash path, bit for bit."""
    gb, gh = two_paths(0.1, 10.0, **W.PROBS)
    rng = np.random.default_rng(17)
    for n in (600_000, 100_000, 1_200_000):
        print(f"scan of {n} rays", flush=True)
        orig = rng.uniform(-1.0, 1.0, 3)
        pts = (orig + rng.normal(0.0, 3.0, (n, 3))).astype(np.float32)
        for g in (gb, gh):
            assert g.insert("s", pts, orig) == V.OK
    assert exports_equal(gb, gh) > 10_000
    assert gb.stats()["visits"] == gh.stats()["visits"]
    assert gb.stats()["rays"] == 1_900_000


@pytest.mark.parametrize("origin", [(0.25, 0.25, 0.25), (-31.75, 63.25, -0.75), (31.5, -32.5, 95.5)])
def test_tie_rays_across_brick_faces(origin):
    """Rays with exact crossing ties (diagonals, rational multiples) from
    origins on and next to brick faces, positive and negative coordinates."""
    g, o = pair(1.0, max_range=5000.0)
    ends = []
    for e in TIE_ENDS:
        for sx in (1, -1):
            for sy in (1, -1):
                ends.append((origin[0] + sx * e[0], origin[1] + sy * e[1], origin[2] + e[2]))
    both(g, o, "insert", "test", np.array(ends, np.float32), list(origin))
    same_map(g, o)


def test_long_rays_grow_every_brick_table():
    """Few very long rays into an empty map: the brick table, segment array
    and record list all start too small and are grown and re-run."""
    g, o = pair(0.02, max_range=60.0)
    rng = np.random.default_rng(14)
    d = rng.normal(size=(300, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    pts = (d * rng.uniform(20, 59, (300, 1))).astype(np.float32)
    both(g, o, "insert", "test", pts, [0.01, -0.03, 0.02])
    same_map(g, o)


def test_zero_length_truncated_and_nan_rays():
    """End points in the origin voxel (hit without a walk), points beyond max
    range (walk without a hit), NaN points (skipped)."""
    g, o = pair(0.1, max_range=3.0)
    pts = np.array([[0.01, 0.02, 0.03], [0.04, 0.0, 0.0], [5.0, 0.0, 0.0], [0.0, -7.0, 1.0],
                    [np.nan, 1.0, 1.0], [1.0, 1.0, 1.0], [-0.3, 0.2, 2.9]], np.float32)
    both(g, o, "insert", "test", pts, [0.0, 0.0, 0.0])
    both(g, o, "insert", "test", pts[::-1].copy(), [0.0, 0.0, 0.0])
    same_map(g, o)


@pytest.mark.parametrize("seed", range(4))
def test_random_inserts_brick_path(seed):
    rng = np.random.default_rng(300 + seed)
    res = [0.1, 0.05, 0.2, 0.5][seed]
    R = [5.0, 3.0, 12.0, 30.0][seed]
    g, o = pair(res, max_range=R, **W.PROBS)
    for _ in range(4):
        origin = rng.uniform(-3, 3, 3)
        n = int(rng.integers(1, 20000))
        pts = (origin + rng.normal(0, R * 0.7, (n, 3))).astype(np.float32)
        pts[rng.random(n) < 0.02] = np.nan
        both(g, o, "insert", "test", pts, origin)
    same_map(g, o)


def test_pending_updates_use_the_hash_path_in_order():
    """With another source's update pending, insert integrates both sources in
    id order (the brick path only runs when nothing is pending)."""
    g, o = pair(0.1, max_range=4.0, sources=("a", "b"))
    rng = np.random.default_rng(21)
    p1 = rng.normal(0, 2, (3000, 3)).astype(np.float32)
    p2 = rng.normal(0, 2, (3000, 3)).astype(np.float32)
    both(g, o, "accumulate", "b", p1, [0.1, 0.0, 0.0])
    both(g, o, "insert", "a", p2, [0.0, 0.2, 0.0])
    both(g, o, "insert", "a", p1, [0.0, 0.0, 0.3])
    same_map(g, o)


def test_brick_path_device_resident_and_profiled():
    w = W.room_scan_workload()
    s = w.scan(0)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    g.set_profiling(True)
    dev = V.DeviceBuffer.from_array(s.points)
    assert g.insert_device("test", dev.address, s.points.shape[0], 16, s.origin) == V.OK
    o.insert("test", s.points, s.origin)
    same_map(g, o)
    st = g.stats()
    assert st["n_walk"] == 1 and st["ms_walk"] > 0 and st["n_traverse"] == 1


def test_update_leaf_count_matches_hash_update_set():
    """stats.update_leaves of a brick-path insert equals the number of leaves
    of the same scan's update set on the hash path."""
    w = W.room_scan_workload()
    s = w.scan(0)
    gb, gh = two_paths(w.resolution, w.max_range, **W.PROBS)
    assert gh.accumulate("s", s.points, s.origin) == V.OK
    n_update = gh.export_update("s")[0].shape[0]
    assert gb.insert("s", s.points, s.origin) == V.OK
    assert gb.stats()["update_leaves"] == n_update > 3000
    assert gh.integrate() == V.OK
    assert gh.stats()["update_leaves"] == n_update


def test_streamed_inserts_interleaved_with_integrates():
    """Back-to-back inserts (the next scan's cast overlaps the previous map
    update) mixed with accumulate/integrate of a second source and reads of
    the map in between."""
    g, o = pair(0.1, max_range=8.0, sources=("lidar", "depth"), **W.PROBS)
    rng = np.random.default_rng(33)
    for k in range(10):
        origin = rng.uniform(-2, 2, 3)
        pts = (origin + rng.normal(0, 4, (int(rng.integers(2000, 30000)), 3))).astype(np.float32)
        both(g, o, "insert", "lidar", pts, origin)
        if k % 3 == 1:
            both(g, o, "accumulate", "depth", pts[::5] * 0.5, origin * 0.5)
        if k % 3 == 2:
            both(g, o, "integrate")
        if k == 6:
            same_map(g, o)
    same_map(g, o)
