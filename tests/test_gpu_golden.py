"""The GPU library against the committed golden fixtures (tests/golden/),
without running the oracle: the reference's known-answer tests as data,
tie-heavy rays, a fast-mode map with raytrace queries, and digests of the
room-scan map. Both the brick path (plain inserts) and the update-set path
(accumulate / integrate) are covered."""
import numpy as np
import pytest

import vdb_mapping_amd as V
from golden_cases import cases, check_map, digest, digests, known_answers, popcount, \
    run_known_answer
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def make_gpu(traversal):
    def make(res, max_range, probs):
        g = V.GpuMap(res)
        assert g.set_config(V.make_config(max_range=max_range, **probs)) == V.OK
        assert g.add_input_source("test", max_range) == V.OK
        assert g.set_tuning(traversal=traversal) == V.OK
        return g
    return make


@pytest.mark.parametrize("traversal", [0, 1])
@pytest.mark.parametrize("name", [c["name"] for c in known_answers()["cases"]])
def test_known_answer_fixture(name, traversal):
    ka = known_answers()
    case = next(c for c in ka["cases"] if c["name"] == name)
    run_known_answer(make_gpu(traversal), case, ka["probs"])


def test_tie_rays_fixture():
    d = cases()
    g = make_gpu(1)(1.0, 5000.0, {})
    assert g.accumulate("test", d["tie_points"], d["tie_origin"]) == V.OK
    uo, uv, uh = g.export_update("test")
    assert np.array_equal(uo, d["tie_update_origins"])
    assert np.array_equal(uv, d["tie_update_visited"])
    assert np.array_equal(uh, d["tie_update_hit"])
    assert g.integrate() == V.OK
    check_map(g, d, "tie_map")
    gb = make_gpu(0)(1.0, 5000.0, {})
    assert gb.insert("test", d["tie_points"], d["tie_origin"]) == V.OK
    check_map(gb, d, "tie_map")


def test_fast_mode_raytrace_fixture():
    d = cases()
    g = V.GpuMap(0.1)
    assert g.set_config(V.make_config(max_range=6.0, fast_mode=True, **W.PROBS)) == V.OK
    g.add_input_source("test", 6.0)
    for o in d["fast_origins"]:
        assert g.insert("test", d["fast_points"], o) == V.OK
    check_map(g, d, "fast_map")
    st, succ, ends = g.raytrace(d["rt_origins"], d["rt_dirs"], d["rt_lengths"])
    assert st == V.OK
    assert np.array_equal(succ.astype(np.uint8), d["rt_success"])
    assert np.array_equal(ends.view(np.uint64), d["rt_ends"].view(np.uint64))


@pytest.mark.parametrize("traversal", [0, 1])
def test_room_scan_digests(traversal):
    dg = digests()
    w = W.room_scan_workload()
    g = make_gpu(traversal)(w.resolution, w.max_range, W.PROBS)
    s = w.scan(0)
    assert g.accumulate("test", s.points[:4096], s.origin) == V.OK
    uo, uv, uh = g.export_update("test")
    u = dg["room_scan_first_4096_update"]
    assert (uo.shape[0], popcount(uv)) == (u["leaves"], u["visited_voxels"])
    assert digest(uo, uv, uh) == u["sha256"]
    assert g.reset() == V.OK
    r = dg["room_scan"]
    for i in range(r["scans"]):
        s = w.scan(i)
        assert g.insert("test", s.points, s.origin) == V.OK
    o, v, a = g.export_leaves()
    assert (o.shape[0], popcount(a)) == (r["leaves"], r["active_voxels"])
    assert digest(o, v.view(np.uint32), a) == r["sha256"]


@pytest.mark.parametrize("traversal", [0, 1])
def test_edit_fixture(traversal):
    from golden_cases import check_edit_case, run_edit_case

    def make(res):
        g = V.GpuMap(res)
        assert g.set_tuning(traversal=traversal) == V.OK
        return g
    check_edit_case(run_edit_case(make, V.make_config), cases())
