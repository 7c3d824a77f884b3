"""The oracle against the committed golden fixtures (tests/golden/): the
reference's known-answer tests as data, and the oracle's own pinned outputs
(a change of the restated semantics shows up here first)."""
import numpy as np
import pytest

import oracle as O
from golden_cases import cases, check_map, digest, digests, known_answers, popcount, \
    run_known_answer
from vdb_mapping_amd import workloads as W


def make_oracle(res, max_range, probs):
    m = O.OracleMap(res)
    assert m.set_config(O.make_config(max_range=max_range, **probs)) == O.OK
    assert m.add_input_source("test", max_range) == O.OK
    return m


@pytest.mark.parametrize("name", [c["name"] for c in known_answers()["cases"]])
def test_known_answer_fixture(name):
    ka = known_answers()
    case = next(c for c in ka["cases"] if c["name"] == name)
    run_known_answer(make_oracle, case, ka["probs"])


def test_tie_rays_fixture():
    d = cases()
    m = make_oracle(1.0, 5000.0, {})
    assert m.accumulate("test", d["tie_points"], d["tie_origin"]) == O.OK
    uo, uv, uh = m.export_update("test")
    assert np.array_equal(uo, d["tie_update_origins"])
    assert np.array_equal(uv, d["tie_update_visited"])
    assert np.array_equal(uh, d["tie_update_hit"])
    assert m.integrate() == O.OK
    assert check_map(m, d, "tie_map") > 100


def test_fast_mode_raytrace_fixture():
    d = cases()
    m = O.OracleMap(0.1)
    assert m.set_config(O.make_config(max_range=6.0, fast_mode=True, **W.PROBS)) == O.OK
    m.add_input_source("test", 6.0)
    for o in d["fast_origins"]:
        assert m.insert("test", d["fast_points"], o) == O.OK
    check_map(m, d, "fast_map")
    st, succ, ends = m.raytrace(d["rt_origins"], d["rt_dirs"], d["rt_lengths"])
    assert st == O.OK
    assert np.array_equal(succ.astype(np.uint8), d["rt_success"])
    assert np.array_equal(ends.view(np.uint64), d["rt_ends"].view(np.uint64))
    assert 0 < succ.sum() < succ.size


def test_room_scan_digests():
    dg = digests()
    w = W.room_scan_workload()
    m = make_oracle(w.resolution, w.max_range, W.PROBS)
    s = w.scan(0)
    assert m.accumulate("test", s.points[:4096], s.origin) == O.OK
    uo, uv, uh = m.export_update("test")
    u = dg["room_scan_first_4096_update"]
    assert (uo.shape[0], popcount(uv)) == (u["leaves"], u["visited_voxels"])
    assert digest(uo, uv, uh) == u["sha256"]
    m.reset()
    r = dg["room_scan"]
    for i in range(r["scans"]):
        s = w.scan(i)
        assert m.insert("test", s.points, s.origin) == O.OK
    o, v, a = m.export_leaves()
    assert (o.shape[0], popcount(a)) == (r["leaves"], r["active_voxels"])
    assert digest(o, v.view(np.uint32), a) == r["sha256"]


def test_edit_fixture():
    from golden_cases import check_edit_case, run_edit_case
    m = run_edit_case(O.OracleMap, O.make_config)
    check_edit_case(m, cases())


def test_multi_robot_digest_first_map():
    """Drift check of the multi-robot serving digests: the oracle rebuilds map 0."""
    entry = digests()["multi_robot_serve"]["maps"][0]
    w = W.multi_robot_pool()
    m = make_oracle(w.resolution, w.max_range, W.PROBS)
    for k in entry["scans"]:
        s = w.scan(k)
        assert m.insert("test", s.points, s.origin) == O.OK
    o, v, a = m.export_leaves()
    assert (o.shape[0], popcount(a)) == (entry["leaves"], entry["active_voxels"])
    assert digest(o, v.view(np.uint32), a) == entry["sha256"]
