"""GPU parity of raytrace (VDBMapping.hpp:675-721) against the oracle, bit for bit:
success flags and end points (float64, compared exactly)."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from golden_cases import digest, digests, popcount
from parity import same_update, both, pair, same_map
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def fast_mode_pair(res, max_range, **probs):
    g, o = pair(res, max_range=max_range, **probs)
    return g, o


def switch_to_fast_mode(g, o, max_range, **probs):
    assert g.set_config(V.make_config(max_range=max_range, fast_mode=True, **probs)) == V.OK
    assert o.set_config(O.make_config(max_range=max_range, fast_mode=True, **probs)) == O.OK
    both(g, o, "integrate")


def same_raytrace(g, o, origins, dirs, lens):
    sg, okg, eg = g.raytrace(origins, dirs, lens)
    so, oko, eo = o.raytrace(origins, dirs, lens)
    assert sg == so == 0, (sg, so, g.last_error())
    assert np.array_equal(okg, oko)
    assert np.array_equal(eg.view(np.uint64), eo.view(np.uint64))
    return okg.mean()


def test_raytrace_without_intersector():
    g, o = pair(1.0)
    both(g, o, "insert", "test", np.array([[0, 0, 3]], np.float32), [0, 0, 0])
    assert g.raytrace([[0, 0, 0]], [[0, 0, 1]], [5.0])[0] == V.NO_INTERSECTOR


def test_raytrace_wall():
    g, o = pair(1.0, max_range=50.0)
    ys, zs = np.meshgrid(np.arange(-4, 5), np.arange(-4, 5))
    wall = np.stack([np.full(ys.size, 10.0), ys.ravel(), zs.ravel()], 1).astype(np.float32)
    both(g, o, "insert", "test", wall, [0, 0, 0])
    switch_to_fast_mode(g, o, 50.0)
    same_map(g, o)
    orig = [[0.5, 0.5, 0.5], [0.5, 2.5, -1.5], [0.5, 0.5, 0.5], [0.5, 30.5, 0.5], [-20, 0.3, 0.2]]
    dirs = [[1, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [1, 0.01, -0.02]]
    lens = [20.0, 20.0, 4.0, 4.0, 60.0]
    same_raytrace(g, o, orig, dirs, lens)
    _, ok, end = g.raytrace(orig[:1], dirs[:1], lens[:1])
    assert ok[0] and end[0].tolist() == [10.0, 0.0, 0.0]


@pytest.mark.parametrize("seed", range(4))
def test_raytrace_random_maps(seed):
    rng = np.random.default_rng(500 + seed)
    res = [0.1, 0.05, 0.2, 0.02][seed]
    R = [6.0, 4.0, 15.0, 2.0][seed]
    g, o = pair(res, max_range=R, **W.PROBS)
    for _ in range(3):
        origin = rng.uniform(-1, 1, 3)
        pts = (origin + rng.normal(0, R * 0.6, (4000, 3))).astype(np.float32)
        both(g, o, "insert", "test", pts, origin)
    switch_to_fast_mode(g, o, R, **W.PROBS)
    n = 20000
    orig = rng.uniform(-R, R, (n, 3))
    dirs = rng.normal(size=(n, 3))
    dirs[::97] = 0.0  # zero direction: left unnormalised
    dirs[1::101, 0] = 0.0  # axis-parallel components
    lens = rng.uniform(0.1, 2 * R, n)
    same_raytrace(g, o, orig, dirs, lens)


def test_raytrace_rgbd_frames():
    """Two RGB-D frames (every 4th pixel) at 0.02 m, then 20k query rays of 5 m."""
    w = W.rgbd_workload(2)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    for i in range(2):
        s = w.scan(i)
        both(g, o, "insert", "test", s.points[::4], s.origin)
    switch_to_fast_mode(g, o, w.max_range, **W.PROBS)
    same_map(g, o)
    org, dirs, lens = W.raytrace_queries(20000)
    frac = same_raytrace(g, o, org, dirs, lens)
    assert frac > 0.3


def test_rgbd_workload_full_matches_oracle_digest():
    """The RGB-D workload end to end at full size: 200 frames of 640 x 480
    at 0.02 m inserted in batches of 16, fast mode switched on and integrated,
    then 1,000,000 raytrace queries; the map before and after the switch and
    the query results equal the oracle's (committed digests)."""
    dg = digests()["rgbd_200"]
    w = W.rgbd_workload(dg["frames"])
    g = V.GpuMap(w.resolution)
    assert g.set_config(V.make_config(max_range=w.max_range, **W.PROBS)) == V.OK
    g.add_input_source("lidar", w.max_range)
    for j in range(0, dg["frames"], 16):
        scans = [w.scan(i) for i in range(j, min(j + 16, dg["frames"]))]
        g.insert_batch("lidar", [s.points for s in scans], np.array([s.origin for s in scans]), check=True)
    o, v, m = g.export_leaves()
    assert (o.shape[0], popcount(m)) == (dg["leaves"], dg["active_voxels"])
    assert digest(o, v.view(np.uint32), m) == dg["sha256"]
    assert g.set_config(V.make_config(max_range=w.max_range, fast_mode=True, **W.PROBS)) == V.OK
    g.integrate(check=True)
    o, v, m = g.export_leaves()
    assert (o.shape[0], popcount(m)) == (dg["fast_leaves"], dg["fast_active_voxels"])
    assert digest(o, v.view(np.uint32), m) == dg["fast_sha256"]
    org, dirs, lens = W.raytrace_queries(dg["queries"])
    st, ok, ends = g.raytrace(org, dirs, lens)
    assert st == V.OK
    assert int(ok.sum()) == dg["raytrace_hits"]
    assert digest(ok.astype(np.uint8), np.ascontiguousarray(ends, np.float64).view(np.uint64)) == \
        dg["raytrace_sha256"]
    g.close()


def test_raytrace_device_arrays():
    w = W.rgbd_workload(1)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    s = w.scan(0)
    both(g, o, "insert", "test", s.points[::8], s.origin)
    switch_to_fast_mode(g, o, w.max_range, **W.PROBS)
    org, dirs, lens = W.raytrace_queries(5000)
    bo, bd, bl = (V.DeviceBuffer.from_array(np.ascontiguousarray(a, np.float64)) for a in (org, dirs, lens))
    bs, be = V.DeviceBuffer(5000), V.DeviceBuffer(5000 * 24)
    assert g.raytrace_device(5000, bo.address, bd.address, bl.address, bs.address, be.address) == V.OK
    ok = bs.download(np.zeros(5000, np.uint8)).astype(bool)
    end = be.download(np.zeros((5000, 3), np.float64))
    _, oko, eo = o.raytrace(org, dirs, lens)
    assert np.array_equal(ok, oko) and np.array_equal(end.view(np.uint64), eo.view(np.uint64))


def test_fast_mode_accumulate_after_switch():
    """After switching to fast mode and integrating, accumulates cast through
    the intersector (tests/test_gpu_fast.py covers the walk itself)."""
    g, o = pair(0.1, max_range=5.0)
    both(g, o, "insert", "test", np.array([[1, 0, 0]], np.float32), [0, 0, 0])
    switch_to_fast_mode(g, o, 5.0)
    both(g, o, "accumulate", "test", np.array([[0, 1, 0], [1.05, 0, 0]], np.float32), [0, 0, 0])
    same_update(g, o, "test")
