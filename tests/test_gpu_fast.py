"""Fast-mode raycasting (castRayIntoGridFast, VDBMapping.hpp:577-602) on the
GPU against the oracle, bit for bit: update sets and maps over several scans,
the empty-map and missing-intersector cases, batches, and a mode switch."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from parity import both, pair, same_map, same_update
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def fast_pair(res, R, **probs):
    g, o = pair(res, max_range=R, **probs)
    cg = V.make_config(max_range=R, fast_mode=True, **probs)
    co = O.make_config(max_range=R, fast_mode=True, **probs)
    assert g.set_config(cg) == o.set_config(co) == V.OK
    return g, o


@pytest.mark.parametrize("seed", range(4))
def test_fast_mode_random_scans(seed):
    rng = np.random.default_rng(500 + seed)
    res = [0.1, 0.05, 0.2, 0.5][seed]
    R = [5.0, 3.0, 12.0, 30.0][seed]
    g, o = fast_pair(res, R, **W.PROBS)
    for k in range(5):
        origin = rng.uniform(-1, 1, 3)
        n = int(rng.integers(500, 6000))
        pts = (origin + rng.normal(0, R * 0.6, (n, 3))).astype(np.float32)
        pts[rng.random(n) < 0.02] = np.nan
        both(g, o, "accumulate", "test", pts, origin)
        same_update(g, o, "test")  # scan 0: the map is empty, endpoints only
        both(g, o, "integrate")
        same_map(g, o)


def test_fast_mode_room_scans():
    w = W.room_scan_workload()
    s = w.scan(0)
    g, o = fast_pair(w.resolution, w.max_range, **W.PROBS)
    rng = np.random.default_rng(2)
    for k in range(3):
        origin = s.origin + rng.normal(0, 0.3, 3)
        both(g, o, "insert", "test", s.points[k::3], origin)
    same_map(g, o)
    both(g, o, "accumulate", "test", s.points[::5], s.origin)
    assert same_update(g, o, "test") > 100


def test_fast_mode_needs_an_intersector():
    g, o = pair(0.1, max_range=6.0, **W.PROBS)
    rng = np.random.default_rng(4)
    pts = rng.normal(0, 2, (3000, 3)).astype(np.float32)
    both(g, o, "insert", "test", pts, [0, 0, 0])  # normal mode: the map is not empty now
    cg = V.make_config(max_range=6.0, fast_mode=True, **W.PROBS)
    co = O.make_config(max_range=6.0, fast_mode=True, **W.PROBS)
    assert g.set_config(cg) == o.set_config(co) == V.OK
    assert both(g, o, "accumulate", "test", pts, [0.1, 0, 0]) == V.NO_INTERSECTOR
    both(g, o, "integrate")  # builds the intersectors
    both(g, o, "accumulate", "test", pts, [0.1, 0, 0])
    same_update(g, o, "test")
    both(g, o, "add_input_source", "late", 5.0)
    assert both(g, o, "accumulate", "late", pts, [0, 0, 0]) == V.NO_INTERSECTOR
    both(g, o, "integrate")
    same_map(g, o)


def test_fast_mode_batch_equals_sequential():
    rng = np.random.default_rng(9)
    g, o = fast_pair(0.1, 5.0, **W.PROBS)
    clouds, origins = [], []
    for _ in range(6):
        org = rng.uniform(-1, 1, 3)
        clouds.append((org + rng.normal(0, 2.5, (2000, 3))).astype(np.float32))
        origins.append(org)
    assert g.insert_batch("test", clouds, np.array(origins), None, check=True) == V.OK
    for c, org in zip(clouds, origins):
        o.insert("test", c, org)
    same_map(g, o)


def test_fast_mode_raytrace_after_fast_inserts():
    w = W.rgbd_workload(4)
    g, o = fast_pair(w.resolution, w.max_range, **W.PROBS)
    for i in range(4):
        s = w.scan(i)
        both(g, o, "insert", "test", s.points[::4], s.origin)
    same_map(g, o)
    org, dirs, lens = W.raytrace_queries(2000)
    sg, okg, eg = g.raytrace(org, dirs, lens)
    so, oko, eo = o.raytrace(org, dirs, lens)
    assert sg == so == V.OK
    assert np.array_equal(okg, oko) and np.array_equal(eg, eo)
