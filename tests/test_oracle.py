"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

Sources of truth, none of which is a retyping of the oracle's own formulas:
  * the reference's five known-answer tests (tests/mapping.cpp:6-158),
    re-expressed against the oracle;
  * float32 bit patterns of the log-odds constants (closed form, computed
    independently of the reference with Python's math.log);
  * world->index: for coordinates away from half-voxel boundaries the rounded
    index equals round-half-up(c / res) in exact rational arithmetic;
  * DDA: visit count 1+|dx|+|dy|+|dz|, 6-connectivity, start/end voxels, and an
    exact-rational walk on every |delta| <= 5 triple (each step must take an
    axis whose rational crossing time is minimal; equal crossing sequences must
    tie-break to the larger axis index, VDBMapping.hpp DDA via math::MinIndex);
  * update set: a brute-force Python set built from the visited voxel lists;
  * integration: per-voxel closed forms (threshold activation, clamping after
    six hits at p_hit = 0.7, hits never deactivate, misses never activate).
"""
import math
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

WORDS = np.uint64


def f2u(x):
    return struct.unpack("<I", struct.pack("<f", float(np.float32(x))))[0]


def logit32(p):
    return np.float32(math.log(p) - math.log(1 - p))


def mk(res, max_range=10.0, **probs):
    m = O.OracleMap(res)
    assert m.set_config(O.make_config(max_range=max_range, **probs)) == O.OK
    assert m.add_input_source("test", max_range, 0) == O.OK
    return m


# ---------------------------------------------------------- known answers --

def test_set_config_known_answer():
    """tests/mapping.cpp:6-30: insert before setConfig is a no-op; after it,
    (0,0,0) = log_miss and (0,0,1) = log_hit at resolution 1."""
    m = O.OracleMap(1.0)
    pts = np.array([[0, 0, 1]], np.float32)
    assert m.insert("test", pts, [0, 0, 0]) == O.UNKNOWN_SOURCE
    assert m.get_value([0, 0, 1]) == (0.0, False)
    assert m.set_config(O.make_config()) == O.OK
    m.add_input_source("test", 10.0, 0)
    assert m.insert("test", pts, [0, 0, 0]) == O.OK
    v0, _ = m.get_value([0, 0, 0])
    v1, _ = m.get_value([0, 0, 1])
    assert f2u(v0) == 0xC00C9F54 and f2u(v1) == 0x400C9F54
    assert v0 == logit32(0.1) and v1 == logit32(0.9)


@pytest.mark.parametrize("sign", [1, -1])
def test_insert_point_known_answer(sign):
    """tests/mapping.cpp:32-98: point (0,0,+-0.5) at 0.1 m -> z = 0..+-4 miss and
    inactive, z = +-5 hit and active."""
    m = mk(0.1)
    m.insert("test", np.array([[0, 0, sign * 5 * 0.1]], np.float32), [0, 0, 0])
    for i in range(5):
        v, a = m.get_value([0, 0, sign * i])
        assert v == logit32(0.1) and not a
    v, a = m.get_value([0, 0, sign * 5])
    assert v == logit32(0.9) and a


def test_insert_max_range_known_answer():
    """tests/mapping.cpp:100-131: max_range 0.5, point (0,0,0.7f) -> z = 0..5
    miss and inactive, z = 6 untouched (truncated ray: no hit)."""
    m = mk(0.1, max_range=0.5)
    m.insert("test", np.array([[0, 0, 7 * 0.1]], np.float32), [0, 0, 0])
    for i in range(0, int(0.5 / 0.1) + 1):
        v, a = m.get_value([0, 0, i])
        assert v == logit32(0.1) and not a
    assert m.get_value([0, 0, int(0.5 / 0.1) + 1]) == (0.0, False)


def test_reset_map_known_answer():
    """tests/mapping.cpp:133-158."""
    m = mk(1.0)
    m.insert("test", np.array([[0, 0, 1]], np.float32), [0, 0, 0])
    assert m.get_value([0, 0, 1])[0] == logit32(0.9)
    m.reset()
    assert m.get_value([0, 0, 1]) == (0.0, False)
    assert m.empty()


# ----------------------------------------------------------- closed forms --

def test_logodds_bit_patterns():
    """OccupancyVDBMapping.hpp:79-87 constants for the benchmark probabilities."""
    m = O.OracleMap(0.05)
    m.set_config(O.make_config(prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12, prob_thres_max=0.97))
    got = [f2u(x) for x in m.logodds()]
    assert got == [0x3F58E883, 0xBECF991F, 0xBFFF07F4, 0x405E7867, 0x40930B39, 0xC0930B39]


def test_config_partial_application():
    """VDBMapping.hpp:1458-1463 + OccupancyVDBMapping.hpp:62-77: max_range < 0 is
    rejected by the base but the log-odds are still set; a bad prob_miss returns
    after the base already marked the map configured."""
    m = O.OracleMap(0.1)
    assert m.set_config(O.make_config(max_range=-1.0)) == O.BAD_MAX_RANGE
    assert m.logodds()[0] == logit32(0.9)
    m2 = O.OracleMap(0.1)
    assert m2.set_config(O.make_config(prob_miss=0.6)) == O.BAD_PROB_MISS
    assert (m2.logodds() == 0).all()
    m2.add_input_source("s", 5.0)
    assert m2.accumulate("s", np.array([[1, 0, 0]], np.float32), [0, 0, 0]) == O.OK  # configured
    m3 = O.OracleMap(0.1)
    m3.add_input_source("s", 5.0)
    assert m3.accumulate("s", np.array([[1, 0, 0]], np.float32), [0, 0, 0]) == O.NOT_CONFIGURED


def _round_half_up_exact(c, res):
    q = Fraction(c) / Fraction(res)
    return math.floor(q + Fraction(1, 2)), q


def test_world_to_index_examples():
    """Survey A.3 examples: res 0.1 -> 0.5 -> 5, -0.5 -> -5, 0.7f -> 7, 0 -> 0, 1.0 -> 10."""
    out = O.world_to_index(0.1, [0.5, -0.5, float(np.float32(0.7))])
    assert list(out) == [5, -5, 7]
    assert list(O.world_to_index(0.1, [0.0, 1.0, -1.0])) == [0, 10, -10]


def test_world_to_index_rational_rounding():
    """Away from half-voxel boundaries the rounded index is the nearest integer of c/res."""
    rng = np.random.default_rng(7)
    for res in (0.05, 0.1, 0.02, 0.3, 1.0):
        c = rng.uniform(-500, 500, 3000)
        checked = 0
        for x in c:
            k, q = _round_half_up_exact(float(x), res)
            frac = q - math.floor(q)
            if abs(frac - Fraction(1, 2)) < Fraction(1, 10**6):
                continue
            got = O.world_to_index(res, [x, 0.0, 0.0])[0]
            assert got == k, (res, x)
            checked += 1
        assert checked > 2900


# ------------------------------------------------------------------- DDA --

def _check_walk(s, e, path):
    s, e = np.array(s), np.array(e)
    d = e - s
    assert len(path) == (0 if not d.any() else 1 + np.abs(d).sum())
    if not d.any():
        return
    assert (path[0] == s).all() and (path[-1] == e).all()
    steps = np.diff(path, axis=0)
    assert (np.abs(steps).sum(1) == 1).all()          # 6-connected
    assert (steps * np.sign(d) >= 0).all()             # monotone toward e


def test_dda_invariants_random():
    rng = np.random.default_rng(3)
    for _ in range(1500):
        s = rng.integers(-50, 50, 3)
        e = s + rng.integers(-300, 300, 3)
        _check_walk(s, e, O.cast_ray(s, e))


def test_dda_exact_rational_walk():
    """Every |delta| <= 5 triple against an exact-rational crossing walk."""
    rng = np.random.default_rng(11)
    for dx in range(-5, 6):
        for dy in range(-5, 6):
            for dz in range(-5, 6):
                d = np.array([dx, dy, dz])
                if not d.any():
                    continue
                s = rng.integers(-20, 20, 3)
                path = O.cast_ray(s, s + d)
                _check_walk(s, s + d, path)
                n = np.abs(d)
                c = [0, 0, 0]
                for st in np.diff(path, axis=0):
                    a = int(np.nonzero(st)[0][0])
                    heads = {b: Fraction(2 * c[b] + 1, 2 * int(n[b])) for b in range(3) if c[b] < n[b]}
                    tmin = min(heads.values())
                    assert heads[a] == tmin, (d, c, a)
                    tied = [b for b, t in heads.items() if t == tmin and n[b] == n[a]]
                    assert a == max(tied), (d, c, a)   # identical sequences: larger axis first
                    c[a] += 1


def test_dda_diagonal_steps_y_first():
    """(0,0,0) -> (7,7,0) steps y first, then x (ties to the larger index)."""
    p = O.cast_ray([0, 0, 0], [7, 7, 0])
    assert p[:4].tolist() == [[0, 0, 0], [0, 1, 0], [1, 1, 0], [1, 2, 0]]


# ------------------------------------------------------------ update set --

def _unpack_update(o, vis, hit):
    """Expand (origins, masks) into {voxel: hit} using the LeafNode offset layout."""
    out = {}
    for org, vw, hw in zip(o, vis, hit):
        for w in range(8):
            v, h = int(vw[w]), int(hw[w])
            while v:
                b = (v & -v).bit_length() - 1
                v &= v - 1
                n = w * 64 + b
                vox = (int(org[0]) + (n >> 6), int(org[1]) + ((n >> 3) & 7), int(org[2]) + (n & 7))
                out[vox] = bool((h >> b) & 1)
    return out


def _brute_update(points, origin, res, R):
    s = O.world_to_index(res, origin)
    vis, hits = set(), set()
    for p in points:
        p = np.array(p[:3], np.float64)
        if np.isnan(p).any():
            continue
        d = p - origin
        ln = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        trunc = R > 0 and ln > R
        if trunc:
            p = np.array([origin[a] + (d[a] / ln) * R for a in range(3)])
        e = O.world_to_index(res, p)
        for v in O.cast_ray(s, e):
            vis.add(tuple(int(x) for x in v))
        if not trunc:
            vis.add(tuple(int(x) for x in e))
            hits.add(tuple(int(x) for x in e))
    return {v: v in hits for v in vis}


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_update_set_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    res, R = [0.1, 0.05, 0.25][seed], [3.0, 2.0, 6.0][seed]
    origin = rng.uniform(-1, 1, 3)
    pts = (origin + rng.normal(0, R * 0.8, (300, 3))).astype(np.float32)
    pts[::37] = np.nan
    m = mk(res, max_range=R)
    m.accumulate("test", pts, origin)
    got = _unpack_update(*m.export_update("test"))
    assert got == _brute_update(pts, origin, res, R)


def test_hit_overrides_miss_regardless_of_order():
    m = mk(1.0)
    pts = np.array([[0, 0, 5], [0, 0, 2], [0, 0, 4]], np.float32)
    m.accumulate("test", pts, [0, 0, 0])
    u = _unpack_update(*m.export_update("test"))
    assert u[(0, 0, 2)] and u[(0, 0, 4)] and u[(0, 0, 5)] and not u[(0, 0, 3)]


def test_endpoint_equal_origin_is_a_hit():
    """raycastPointCloud marks the endpoint even when no traversal happens (VDBMapping.hpp:533-536)."""
    m = mk(1.0)
    m.accumulate("test", np.array([[0.2, 0.1, 0.3]], np.float32), [0, 0, 0])
    assert _unpack_update(*m.export_update("test")) == {(0, 0, 0): True}


def test_truncation_geometry():
    """A truncated endpoint lies at distance R from the origin along the ray."""
    rng = np.random.default_rng(5)
    res, R = 0.05, 7.5
    origin = np.array([0.3, -0.2, 1.1])
    for _ in range(200):
        p = origin + rng.normal(0, 20, 3)
        d = p - origin
        ln = math.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        if ln <= R:
            continue
        q = np.array([origin[a] + (d[a] / ln) * R for a in range(3)])
        assert abs(np.linalg.norm(q - origin) - R) < 1e-9
        assert np.allclose((q - origin) / R, d / ln, atol=1e-12)


def test_nan_and_unknown_source():
    m = mk(0.1)
    pts = np.array([[np.nan, 0, 0], [0, np.nan, 1]], np.float32)
    assert m.insert("test", pts, [0, 0, 0]) == O.OK
    assert m.export_leaves()[0].shape[0] == 0
    assert m.insert("nope", np.array([[1, 0, 0]], np.float32), [0, 0, 0]) == O.UNKNOWN_SOURCE
    assert m.insert("test", np.array([[1, 0, 0]], np.float32), [np.nan, 0, 0]) == O.OK
    assert m.empty()


# ----------------------------------------------------------- integration --

def test_clamp_after_six_hits():
    """Six hits at p_hit = 0.7: 6 * 0.847 = 5.08 > logit(0.99) = 4.59512 -> clamped;
    five hits stay below the clamp."""
    m = mk(1.0, prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12, prob_thres_max=0.97)
    lh = logit32(0.7)
    x = np.float32(0)
    for i in range(6):
        m.insert("test", np.array([[0, 0, 3]], np.float32), [0, 0, 0])
        v, a = m.get_value([0, 0, 3])
        x = np.float32(x + lh)
        if i < 5:
            assert v == x and v < np.float32(4.59512)
        else:
            assert f2u(v) == 0x40930B39
        assert a == (v > logit32(0.97))


def test_hits_never_deactivate_misses_never_activate():
    m = mk(1.0, prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12, prob_thres_max=0.97)
    for _ in range(5):
        m.insert("test", np.array([[0, 0, 3]], np.float32), [0, 0, 0])
    assert m.get_value([0, 0, 3])[1]
    # a miss that keeps the value above thres_min must not deactivate
    m.insert("test", np.array([[0, 0, 6]], np.float32), [0, 0, 0])
    v, a = m.get_value([0, 0, 3])
    assert a and v > logit32(0.12)
    # misses at an unseen voxel never activate it
    v, a = m.get_value([0, 0, 5])
    assert not a and v == logit32(0.4)


def test_multi_source_integrate_order():
    """integrateUpdate applies each source's set separately (VDBMapping.hpp:380-385):
    a voxel seen by two sources gets two updates."""
    m = mk(1.0)
    m.add_input_source("b", 10.0)
    m.accumulate("test", np.array([[0, 0, 3]], np.float32), [0, 0, 0])
    m.accumulate("b", np.array([[0, 0, 3]], np.float32), [0, 0, 0])
    m.integrate()
    assert m.get_value([0, 0, 3])[0] == np.float32(logit32(0.9) + logit32(0.9))
    assert m.get_value([0, 0, 1])[0] == np.float32(logit32(0.1) + logit32(0.1))


def test_accumulate_twice_is_or_merged():
    """Two scans in one update grid update a voxel once (UpdateGridT is a set)."""
    m = mk(1.0)
    p = np.array([[0, 0, 3]], np.float32)
    m.accumulate("test", p, [0, 0, 0])
    m.accumulate("test", p, [0, 0, 0])
    m.integrate()
    assert m.get_value([0, 0, 3])[0] == logit32(0.9)


def test_leaf_layout_and_sorting():
    m = mk(1.0, max_range=50.0)
    m.insert("test", np.array([[9, -3, 17], [-12, 4, 2]], np.float32), [0, 0, 0])
    o, v, msk = m.export_leaves()
    keys = [tuple(int(c) for c in x) for x in o]
    assert keys == sorted(keys)
    assert all((np.array(k) % 8 == 0).all() for k in keys)
    i = keys.index((8, -8, 16))
    n = (1 << 6) | (5 << 3) | 1
    assert v[i, n] == logit32(0.9)
    assert (int(msk[i, n >> 6]) >> (n & 63)) & 1


# -------------------------------------------------------------- raytrace --

def _fast_mode_map():
    """Wall of occupied voxels at x = 10 (res 1), built in normal mode, then the
    intersector is built by an integrate in fast mode (survey A.9 recipe)."""
    m = O.OracleMap(1.0)
    m.set_config(O.make_config(max_range=50.0))
    m.add_input_source("s", 50.0)
    ys, zs = np.meshgrid(np.arange(-4, 5), np.arange(-4, 5))
    wall = np.stack([np.full(ys.size, 10.0), ys.ravel(), zs.ravel()], 1).astype(np.float32)
    m.insert("s", wall, [0, 0, 0])
    m.set_config(O.make_config(max_range=50.0, fast_mode=True))
    m.integrate()
    return m


def test_raytrace_requires_intersector():
    m = mk(1.0)
    m.insert("test", np.array([[0, 0, 3]], np.float32), [0, 0, 0])
    st, _, _ = m.raytrace([[0, 0, 0]], [[0, 0, 1]], [5.0])
    assert st == O.NO_INTERSECTOR


def test_raytrace_hits_wall():
    m = _fast_mode_map()
    st, ok, end = m.raytrace([[0.5, 0.5, 0.5], [0.5, 2.5, -1.5]], [[1, 0, 0], [1, 0, 0]], [20.0, 20.0])
    assert st == O.OK and ok.all()
    assert end.tolist() == [[10.0, 0.0, 0.0], [10.0, 2.0, -2.0]]


def test_raytrace_leaf_run_without_active_voxel():
    """Ray from the origin voxel toward -x (length 4, res 1): the origin's leaf
    exists (miss voxels) and the next leaf does not, so the HDDA run is
    t in [0, 0.125]; the voxel DDA tests voxel -1 (inactive), steps past the
    run and reports voxel -2 with success (VDBMapping.hpp:702-713)."""
    m = _fast_mode_map()
    st, ok, end = m.raytrace([[0.5, 0.5, 0.5]], [[-1, 0, 0]], [4.0])
    assert st == O.OK and ok[0]
    assert end[0].tolist() == [-2.0, 0.0, 0.0]


def test_raytrace_miss_reports_segment_end():
    """No leaf along the ray: success false, end = origin + direction * length (VDBMapping.hpp:717-718)."""
    m = _fast_mode_map()
    st, ok, end = m.raytrace([[0.5, 30.5, 0.5]], [[0, 1, 0]], [4.0])
    assert st == O.OK and not ok[0]
    assert end[0].tolist() == [0.5, 34.5, 0.5]


# ------------------------------------------------------------ fast mode --
# castRayIntoGridFast (VDBMapping.hpp:577-602): pinned by closed forms and
# invariants; its agreement with OpenVDB's VolumeHDDA::hits itself (U11) is
# unpinned.

def _voxels(origins, words):
    """Set of voxel coords of an exported update (origins + 8 mask words)."""
    out = set()
    for o, w in zip(origins, words):
        for j in range(8):
            bits = int(w[j])
            while bits:
                b = (bits & -bits).bit_length() - 1
                bits &= bits - 1
                n = j * 64 + b
                out.add((int(o[0]) + (n >> 6), int(o[1]) + ((n >> 3) & 7), int(o[2]) + (n & 7)))
    return out


def _active_block_map(fast):
    """1 m voxels; the leaf-aligned block [0,16)^3 made active by one scan whose
    endpoints are the block's voxel centres (one hit activates at thres 0.8)."""
    m = O.OracleMap(1.0)
    m.set_config(O.make_config(max_range=1000.0, prob_hit=0.9, prob_miss=0.45, prob_thres_min=0.2,
                               prob_thres_max=0.8))
    m.add_input_source("s", 1000.0)
    g = np.stack(np.meshgrid(*(np.arange(16),) * 3, indexing="ij"), -1).reshape(-1, 3) + 0.25
    assert m.insert("s", g.astype(np.float32), [-30.25, -30.25, -30.25]) == O.OK
    _, _, act = m.export_leaves()
    block = {(x, y, z) for x in range(16) for y in range(16) for z in range(16)}
    on = set()
    o, _, a = m.export_leaves()
    on = _voxels(o, a)
    assert block <= on
    if fast:
        assert m.set_config(O.make_config(max_range=1000.0, fast_mode=True, prob_hit=0.9, prob_miss=0.45,
                                          prob_thres_min=0.2, prob_thres_max=0.8)) == O.OK
        assert m.integrate() == O.OK  # builds the intersector (updateVolumeRayIntersectors)
    return m, on


def test_fast_mode_axis_ray_through_active_block_closed_form():
    m, on = _active_block_map(True)
    # along x at y = 3, z = 5 from x = -40 to x = 40 (truncated at range 1000: no)
    pts = np.array([[40.25, 3.25, 5.25]], np.float32)
    assert m.accumulate("s", pts, [-39.75, 3.25, 5.25]) == O.OK
    o, vis, hit = m.export_update("s")
    got = _voxels(o, vis)
    want = {(x, 3, 5) for x in range(16)} | {(40, 3, 5)}  # the block voxels on the line + the endpoint
    assert got == want
    assert _voxels(o, hit) == {(40, 3, 5)}


def test_fast_mode_marks_only_previously_active_voxels():
    m, on = _active_block_map(True)
    rng = np.random.default_rng(5)
    org = np.array([-7.7, 21.3, 8.1])
    pts = (rng.uniform(-10, 26, (400, 3))).astype(np.float32)
    assert m.accumulate("s", pts, org) == O.OK
    o, vis, hit = m.export_update("s")
    ends = {tuple(int(v) for v in np.floor(p.astype(np.float64) + 0.5)) for p in pts}
    visited = _voxels(o, vis)
    assert visited - _voxels(o, hit) <= on
    assert _voxels(o, hit) <= ends
    assert len(visited & on) > 100


def test_fast_mode_empty_map_marks_endpoints_only():
    m = O.OracleMap(0.1)
    m.set_config(O.make_config(max_range=5.0, fast_mode=True))
    m.add_input_source("s", 5.0)
    pts = np.array([[1.0, 0.2, 0.3], [0.5, -0.5, 0.1], [9.0, 0.0, 0.0]], np.float32)
    assert m.accumulate("s", pts, [0, 0, 0]) == O.OK
    o, vis, hit = m.export_update("s")
    assert _voxels(o, vis) == _voxels(o, hit) == {(10, 2, 3), (5, -5, 1)}  # the truncated ray marks nothing


def test_fast_mode_zero_length_ray_probes_origin_voxel():
    m, on = _active_block_map(True)
    assert m.accumulate("s", np.array([[4.3, 4.4, 4.1]], np.float32), [4.2, 4.3, 4.4]) == O.OK
    o, vis, hit = m.export_update("s")
    assert _voxels(o, vis) == _voxels(o, hit) == {(4, 4, 4)}


def test_fast_mode_without_intersector_is_rejected():
    """Normal-mode map, then fast mode without an integrate: the reference
    dereferences a null intersector; the oracle reports NO_INTERSECTOR. A source
    added after the intersector was built has none either."""
    m, on = _active_block_map(False)
    assert m.set_config(O.make_config(max_range=1000.0, fast_mode=True)) == O.OK
    assert m.accumulate("s", np.array([[1, 1, 1]], np.float32), [0, 0, 0]) == O.NO_INTERSECTOR
    assert m.integrate() == O.OK
    assert m.accumulate("s", np.array([[1, 1, 1]], np.float32), [0, 0, 0]) == O.OK
    m.add_input_source("late", 100.0)
    assert m.accumulate("late", np.array([[1, 1, 1]], np.float32), [0, 0, 0]) == O.NO_INTERSECTOR
