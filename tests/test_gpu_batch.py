"""Batched integration (vdbgpu_insert_batch) must equal inserting the scans one
after another: bit-exact against the oracle's sequential inserts and against
the GPU plain path. Every test runs on both batch paths: traversal 0 streams
the scans through the brick path, traversal 1 runs the (leaf, scan)-keyed
traversal and the scan-ordered fold."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
import parity
from golden_cases import digest, digests, popcount
from parity import both, same_map
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1], ids=["streamed", "fold"], autouse=True)
def traversal(request):
    return request.param


@pytest.fixture
def pair(traversal):
    def make(*args, **kw):
        g, o = parity.pair(*args, **kw)
        assert g.set_tuning(traversal=traversal) == V.OK
        return g, o
    return make


def oracle_sequential(o, source, clouds, origins, ranges):
    for c, org, r in zip(clouds, origins, ranges):
        o.add_input_source(source, r)
        o.insert(source, c, org)


def random_scans(rng, n_scans, R, sizes=(0, 3000)):
    clouds, origins, ranges = [], [], []
    for _ in range(n_scans):
        org = rng.uniform(-2, 2, 3)
        n = int(rng.integers(*sizes))
        pts = (org + rng.normal(0, R * 0.7, (n, 3))).astype(np.float32)
        if n:
            pts[rng.random(n) < 0.02] = np.nan
        clouds.append(np.concatenate([pts, np.ones((n, 1), np.float32)], 1))
        origins.append(org)
        ranges.append(float(rng.uniform(0.3 * R, R)))
    return clouds, np.array(origins), np.array(ranges)


@pytest.mark.parametrize("seed", range(4))
def test_batch_equals_sequential(seed, pair):
    rng = np.random.default_rng(900 + seed)
    res = [0.1, 0.05, 0.2, 0.5][seed]
    R = [5.0, 3.0, 12.0, 25.0][seed]
    g, o = pair(res, max_range=R, **W.PROBS)
    clouds, origins, ranges = random_scans(rng, [7, 20, 3, 40][seed], R)
    assert g.insert_batch("test", clouds, origins, ranges, check=True) == V.OK
    oracle_sequential(o, "test", clouds, origins, ranges)
    same_map(g, o)
    st = g.stats()
    assert st["scans"] == sum(1 for c in clouds if c.shape[0] > 0)


def test_batch_equals_gpu_plain_path(traversal):
    rng = np.random.default_rng(77)
    clouds, origins, ranges = random_scans(rng, 12, 6.0, sizes=(100, 5000))
    gb = V.GpuMap(0.05)
    gs = V.GpuMap(0.05)
    for g in (gb, gs):
        g.set_config(V.make_config(max_range=6.0, **W.PROBS))
        g.add_input_source("s", 6.0)
    assert gb.set_tuning(traversal=traversal) == V.OK
    gb.insert_batch("s", clouds, origins, ranges, check=True)
    for c, org, r in zip(clouds, origins, ranges):
        gs.add_input_source("s", r)
        gs.insert("s", c, org, check=True)
    a, b = gb.export_leaves(), gs.export_leaves()
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
    assert np.array_equal(a[2], b[2])


def test_batch_clamps_across_scans(pair):
    """Seven hits on one voxel inside one batch: the clamp and activation
    happen at the same scans as with sequential inserts."""
    g, o = pair(1.0, max_range=10.0, **W.PROBS)
    clouds = [np.array([[0, 0, 3, 1]], np.float32)] * 7 + [np.array([[0, 0, 6, 1]], np.float32)] * 2
    origins = np.zeros((9, 3))
    ranges = np.full(9, 10.0)
    g.insert_batch("test", clouds, origins, ranges, check=True)
    oracle_sequential(o, "test", clouds, origins, ranges)
    same_map(g, o)


def test_batch_more_than_128_scans(pair):
    rng = np.random.default_rng(5)
    clouds, origins, ranges = random_scans(rng, 300, 2.0, sizes=(0, 60))
    g, o = pair(0.1, max_range=2.0)
    g.insert_batch("test", clouds, origins, ranges, check=True)
    oracle_sequential(o, "test", clouds, origins, ranges)
    same_map(g, o)


def test_batch_after_pending_updates_of_other_sources(pair):
    """Pending update sets are integrated with the first scan, in source-id order."""
    g, o = pair(0.1, max_range=4.0, sources=("a", "m", "z"))
    rng = np.random.default_rng(8)
    pre = (rng.normal(0, 2, (800, 3))).astype(np.float32)
    for x in (g, o):
        x.accumulate("a", pre, [0, 0, 0])
        x.accumulate("z", pre[::-1] * 0.7, [0.1, 0, 0])
    clouds, origins, ranges = random_scans(rng, 5, 4.0)
    g.insert_batch("m", clouds, origins, ranges, check=True)
    oracle_sequential(o, "m", clouds, origins, ranges)
    same_map(g, o)


def test_batch_beyond_batched_key_range_falls_back(pair):
    """Voxel indices past 2^21 (but inside 2^23) cannot use (leaf, scan) keys;
    the batch is re-done scan by scan with the same result."""
    g, o = pair(1.0, max_range=0.0)
    for x in (g, o):
        x.add_input_source("far", 50.0)
    base = np.array([2_100_000.0, 0.0, 0.0])
    clouds = [np.array([[2_100_010, 3, 1, 1], [2_099_990, -2, 0, 1]], np.float32) for _ in range(3)]
    origins = np.stack([base + [0, i, 0] for i in range(3)])
    g.insert_batch("far", clouds, origins, None, check=True)
    for c, org in zip(clouds, origins):
        o.insert("far", c, org)
    same_map(g, o)


def test_batch_growth_paths(pair):
    """Long rays into an empty map: the update set overflows its first
    capacity and the map is empty at the fold, so both re-run paths run."""
    rng = np.random.default_rng(6)
    clouds = []
    for _ in range(6):
        d = rng.normal(size=(60, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        clouds.append(np.concatenate([(d * 55).astype(np.float32), np.ones((60, 1), np.float32)], 1))
    origins = rng.uniform(-1, 1, (6, 3))
    g, o = pair(0.02, max_range=60.0)
    g.insert_batch("test", clouds, origins, None, check=True)
    oracle_sequential(o, "test", clouds, origins, [60.0] * 6)
    same_map(g, o)


def test_mixed_batch_workload_subset(pair):
    """Eight scans of the mixed-size batch workload (every 8th point)."""
    w = W.mixed_batch_workload(8)
    scans = [w.scan(i) for i in range(8)]
    clouds = [s.points[::8] for s in scans]
    origins = np.array([s.origin for s in scans])
    ranges = np.array([s.max_range for s in scans])
    g, o = pair(w.resolution, max_range=50.0, **W.PROBS)
    g.insert_batch("test", clouds, origins, ranges, check=True)
    oracle_sequential(o, "test", clouds, origins, ranges)
    assert same_map(g, o) > 1000


def test_batch_tables_grow_between_scans(pair):
    """Small scans, then scans of long rays in the same batch: on the streamed
    path a later scan's cast is already enqueued when an earlier scan's brick
    tables or record set overflow, so the earlier scan is cast again from its
    points; the map equals sequential inserts."""
    rng = np.random.default_rng(11)
    clouds, origins = [], []
    for n, r in ((40, 0.8), (20000, 40.0), (30, 0.5), (30000, 45.0), (25000, 45.0)):
        org = rng.uniform(-1, 1, 3)
        d = rng.normal(size=(n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts = (org + d * rng.uniform(0.2, 1.0, (n, 1)) * r).astype(np.float32)
        clouds.append(np.concatenate([pts, np.ones((n, 1), np.float32)], 1))
        origins.append(org)
    g, o = pair(0.05, max_range=50.0)
    g.insert_batch("test", clouds, np.array(origins), None, check=True)
    oracle_sequential(o, "test", clouds, origins, [50.0] * len(clouds))
    assert same_map(g, o) > 10000


def test_batch_graph_replays_interleaved(pair, monkeypatch):
    """Batched inserts on one map, with plain inserts and exports between them:
    with VDBGPU_GRAPH=1 the streamed path replays its captured cast graphs
    (different scan sizes, every record set of the ring); each state equals the
    oracle's."""
    monkeypatch.setenv("VDBGPU_GRAPH", "1")
    rng = np.random.default_rng(21)
    g, o = pair(0.05, max_range=8.0)
    for _ in range(4):
        clouds, origins, ranges = random_scans(rng, 7, 8.0, sizes=(500, 6000))
        assert g.insert_batch("test", clouds, origins, ranges, check=True) == V.OK
        oracle_sequential(o, "test", clouds, origins, ranges)
        same_map(g, o)
        for m in (g, o):
            assert m.add_input_source("test", 8.0) == V.OK
        both(g, o, "insert", "test", clouds[1], origins[1])
        same_map(g, o)


def test_batch_graph_equals_launches(monkeypatch):
    """The streamed batch path with its casts replayed from graphs
    (VDBGPU_GRAPH=1) and with the same casts launched one by one gives the
    same map."""
    rng = np.random.default_rng(22)
    clouds, origins, ranges = random_scans(rng, 20, 10.0, sizes=(100, 8000))
    maps = []
    for on in ("1", "0"):
        monkeypatch.setenv("VDBGPU_GRAPH", on)
        m = V.GpuMap(0.05)
        m.set_config(V.make_config(max_range=10.0, **W.PROBS))
        m.add_input_source("s", 10.0)
        assert m.insert_batch("s", clouds, origins, ranges, check=True) == V.OK
        maps.append(m.export_leaves())
        m.close()
    (ao, av, am), (bo, bv, bm) = maps
    assert ao.shape[0] > 1000
    assert np.array_equal(ao, bo) and np.array_equal(av.view(np.uint32), bv.view(np.uint32))
    assert np.array_equal(am, bm)


def test_mixed_batch_workload_full_matches_oracle_digest(traversal):
    """All 64 scans of the mixed-size batch workload (10k-300k points, ranges
    U[20, 50] m, 5.8 M rays) in one batched insert: the map equals the
    oracle's after inserting them in order (committed digest)."""
    dg = digests()["mixed_batch_64"]
    w = W.mixed_batch_workload(dg["scans"])
    scans = [w.scan(i) for i in range(dg["scans"])]
    assert sum(s.points.shape[0] for s in scans) == dg["rays"]
    g = V.GpuMap(w.resolution)
    assert g.set_config(V.make_config(max_range=w.max_range, **W.PROBS)) == V.OK
    assert g.set_tuning(traversal=traversal) == V.OK
    g.add_input_source("lidar", w.max_range)
    g.insert_batch("lidar", [s.points for s in scans], np.array([s.origin for s in scans]),
                   np.array([s.max_range for s in scans]), check=True)
    o, v, m = g.export_leaves()
    assert (o.shape[0], popcount(m)) == (dg["leaves"], dg["active_voxels"])
    assert digest(o, v.view(np.uint32), m) == dg["sha256"]
    g.close()


def test_batch_device_resident(pair):
    rng = np.random.default_rng(3)
    clouds, origins, ranges = random_scans(rng, 9, 5.0, sizes=(1, 2000))
    cat = np.concatenate(clouds)
    dev = V.DeviceBuffer.from_array(cat)
    g, o = pair(0.1, max_range=5.0)
    g.insert_batch_device("test", dev.address, [c.shape[0] for c in clouds], 16, origins, ranges,
                          check=True)
    oracle_sequential(o, "test", clouds, origins, ranges)
    same_map(g, o)
