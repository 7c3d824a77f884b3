"""The allocation-free raytrace entry point (vdbgpu_raytrace_noalloc) against
vdbgpu_raytrace, through the C ABI: bit-identical success flags and end points
on two calls in a row, the second smaller than the first, so the second call
runs on scratch the first one left behind."""
import numpy as np
import pytest

import vdb_mapping_amd as V
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def fast_map():
    w = W.rgbd_workload(2)
    g = V.GpuMap(w.resolution)
    assert g.set_config(V.make_config(max_range=w.max_range, **W.PROBS)) == V.OK
    g.add_input_source("cam", w.max_range)
    for i in range(2):
        s = w.scan(i)
        assert g.insert("cam", s.points[::4], s.origin, check=True) == V.OK
    assert g.set_config(V.make_config(max_range=w.max_range, fast_mode=True, **W.PROBS)) == V.OK
    assert g.integrate(check=True) == V.OK
    return g


def test_noalloc_raytrace_matches_raytrace_twice():
    g = fast_map()
    org, dirs, lens = W.raytrace_queries(200000)
    for n in (200000, 777):  # the second call is smaller: stale scratch must not leak into it
        o, d, ln = org[:n] * (1.0 if n > 1000 else 0.5), dirs[:n], lens[:n]
        s0, ok0, e0 = g.raytrace(o, d, ln)
        s1, ok1, e1 = g.raytrace_noalloc(o, d, ln)
        assert s0 == s1 == V.OK, g.last_error()
        assert np.array_equal(ok0, ok1)
        assert np.array_equal(e0.view(np.uint64), e1.view(np.uint64))
        assert 0.05 < ok1.mean() < 1.0


def test_noalloc_raytrace_limits():
    g = fast_map()
    n = V.NOALLOC_MAX_QUERIES + 1
    z = np.zeros((n, 3))
    st, _, _ = g.raytrace_noalloc(z, z + 1.0, np.ones(n))
    assert st == V.BAD_ARGUMENT and "393216" in g.last_error()
    # a scan still waiting for its map update is refused, not applied (that may allocate)
    w = W.rgbd_workload(1)
    g2 = V.GpuMap(w.resolution)
    assert g2.set_config(V.make_config(max_range=w.max_range, **W.PROBS)) == V.OK
    g2.add_input_source("cam", w.max_range)
    s = w.scan(0)
    assert g2.insert("cam", s.points[::8], s.origin, check=True) == V.OK
    assert g2.set_config(V.make_config(max_range=w.max_range, fast_mode=True, **W.PROBS)) == V.OK
    assert g2.integrate(check=True) == V.OK
    assert g2.set_config(V.make_config(max_range=w.max_range, **W.PROBS)) == V.OK
    assert g2.insert("cam", s.points[1::8], s.origin, check=True) == V.OK
    org, dirs, lens = W.raytrace_queries(1000)
    st, _, _ = g2.raytrace_noalloc(org, dirs, lens)
    assert st == V.BAD_ARGUMENT and "vdbgpu_synchronize" in g2.last_error()
    assert g2.synchronize() == V.OK
    st, ok, end = g2.raytrace_noalloc(org, dirs, lens)
    s0, ok0, e0 = g2.raytrace(org, dirs, lens)
    assert st == s0 == V.OK
    assert np.array_equal(ok, ok0) and np.array_equal(end.view(np.uint64), e0.view(np.uint64))
