"""Pins the oracle's map edits and map sections (VDBMapping.hpp:413-447,
:811-1085, :1152-1236; OccupancyVDBMapping.hpp:118-134) before they are used
as the checker.

Sources of truth, none a retyping of the oracle's code:
  * closed forms per voxel: addPointsToGrid sets (logit(0.99), active),
    removePointsFromGrid (logit(0.01), inactive), restoreMapIntegrity
    active = value > t_max; voxels not named keep their bits;
  * twin maps: a map with artificial areas against the same map without
    them (values equal everywhere, active states equal except the artificial
    voxels, which are forced on after every non-empty update);
  * the artificial grid against the union of castRayIntoGrid walks (pinned by
    the DDA tests) over the polygon edges and height layers;
  * map sections against a numpy crop of the full leaf export, round trips
    (extract -> apply into an empty map -> extract) and hand-built maps for
    the internal-tile visit of applyMapSectionGrid;
  * createIndexBoundingBox on exactly representable boxes and rotations.
"""
import math

import numpy as np
import pytest

import oracle as O

PROBS = dict(prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12, prob_thres_max=0.97)
C_MAX = np.float32(math.log(0.99) - math.log(0.01))
C_MIN = np.float32(math.log(0.01) - math.log(0.99))


def mk(res=0.1, R=5.0, **probs):
    m = O.OracleMap(res)
    assert m.set_config(O.make_config(max_range=R, **(probs or PROBS))) == O.OK
    assert m.add_input_source("s", R) == O.OK
    return m


def as_dict(m):
    """Voxel dict {(x,y,z): (value bits, active)} of every voxel of every leaf
    and of every leaf-level tile (createMapFromPointCloud prunes uniform leaves)."""
    o, v, a = m.export_leaves()
    out = {}
    for k in range(o.shape[0]):
        bits = v[k].view(np.uint32)
        for n in range(512):
            c = (int(o[k, 0]) + (n >> 6), int(o[k, 1]) + ((n >> 3) & 7), int(o[k, 2]) + (n & 7))
            out[c] = (int(bits[n]), bool((int(a[k, n >> 6]) >> (n & 63)) & 1))
    to, tv, ta = m.export_tiles()
    for k in range(to.shape[0]):
        st = (f32bits(tv[k]), bool(ta[k]))
        for n in range(512):
            out[(int(to[k, 0]) + (n >> 6), int(to[k, 1]) + ((n >> 3) & 7), int(to[k, 2]) + (n & 7))] = st
    return out


def leaf_set(m):
    return {tuple(int(c) for c in r) for r in m.export_leaves()[0]}


def vox(p, res):
    return tuple(int(math.floor(float(np.float32(c)) * (1.0 / res))) for c in p)


def f32bits(x):
    return int(np.array([x], np.float32).view(np.uint32)[0])


def scanned_map(seed=5, n=3000):
    m = mk()
    rng = np.random.default_rng(seed)
    for k in range(3):
        o = rng.uniform(-0.5, 0.5, 3)
        assert m.insert("s", (o + rng.normal(0, 1.2, (n, 3))).astype(np.float32), o) == O.OK
    return m


# -------------------------------------------------------------- point edits --

def test_add_and_remove_points_closed_form():
    m = scanned_map()
    before = as_dict(m)
    leaves0 = leaf_set(m)
    rng = np.random.default_rng(1)
    add = rng.uniform(-3, 3, (500, 3)).astype(np.float32)
    assert m.add_points_to_grid(add) == O.OK
    after = as_dict(m)
    targets = {vox(p, 0.1) for p in add}
    for c, st in after.items():
        if c in targets:
            assert st == (f32bits(C_MAX), True)
        else:
            assert st == before.get(c, (0, False))
    assert leaf_set(m) == leaves0 | {tuple((x // 8) * 8 for x in c) for c in targets}
    rem = np.concatenate([add[::3], rng.uniform(-3, 3, (100, 3)).astype(np.float32)])
    assert m.remove_points_from_grid(rem) == O.OK
    after2 = as_dict(m)
    rtargets = {vox(p, 0.1) for p in rem}
    for c, st in after2.items():
        if c in rtargets:
            assert st == (f32bits(C_MIN), False)
        else:
            assert st == after.get(c, (0, False))


def test_point_edits_are_idempotent_and_order_free():
    rng = np.random.default_rng(2)
    pts = rng.uniform(-2, 2, (400, 3)).astype(np.float32)
    a, b = scanned_map(), scanned_map()
    assert a.add_points_to_grid(pts) == O.OK
    assert b.add_points_to_grid(pts[::-1].copy()) == O.OK
    assert b.add_points_to_grid(pts) == O.OK
    assert as_dict(a) == as_dict(b)


def test_point_edit_rejects_nan_before_any_change():
    m = scanned_map()
    before = as_dict(m)
    pts = np.array([[0.5, 0.5, 0.5], [np.nan, 0, 0]], np.float32)
    assert m.add_points_to_grid(pts) == O.OUT_OF_RANGE
    assert as_dict(m) == before


def test_remove_on_empty_map_creates_leaves():
    """setNodeToFree changes a background voxel (value becomes logit(0.01)),
    so the leaf is created."""
    m = mk()
    assert m.remove_points_from_grid(np.array([[0.05, 0.05, 0.05]], np.float32)) == O.OK
    assert leaf_set(m) == {(0, 0, 0)}
    assert m.get_value([0, 0, 0]) == (C_MIN, False)
    assert not m.empty()


# ------------------------------------------------------- artificial areas --

def art_voxels(m):
    o, msk = m.export_artificial()
    out = set()
    for k in range(o.shape[0]):
        for n in range(512):
            if (int(msk[k, n >> 6]) >> (n & 63)) & 1:
                out.add((int(o[k, 0]) + (n >> 6), int(o[k, 1]) + ((n >> 3) & 7), int(o[k, 2]) + (n & 7)))
    return out


POLYS = [np.array([[-1.0, -1.0, 0.0, 1.0], [1.2, -1.0, 0.0, 1.0], [1.2, 0.8, 0.0, 1.0]]),
         np.array([[2.05, 2.0, -0.3], [-2.0, 2.55, -0.3]]),
         np.array([[0.33, 0.33, 0.1]])]  # a one-vertex polygon: wall from p to p marks nothing


def expected_walls(polys, neg, pos, res):
    out = set()
    k0, k1 = int(neg / res), int(pos / res)
    for p in polys:
        for i in range(len(p)):
            s = O.world_to_index(res, p[i][:3])
            e = O.world_to_index(res, p[(i + 1) % len(p)][:3])
            for k in range(k0, k1):
                dz = np.array([0, 0, k], np.int32)
                for c in O.cast_ray(s + dz, e + dz):
                    out.add(tuple(int(x) for x in c))
    return out


@pytest.mark.parametrize("neg,pos", [(-0.3, 0.45), (0.0, 0.25), (0.2, 0.1)])
def test_artificial_grid_is_union_of_wall_walks(neg, pos):
    m = mk()
    assert m.add_artificial_areas(POLYS, neg, pos) == O.OK
    want = expected_walls(POLYS, neg, pos, 0.1)
    assert art_voxels(m) == want
    if pos > neg + 0.1:
        assert len(want) > 50


def test_artificial_voxels_activate_after_nonempty_updates_only():
    a, b = scanned_map(), scanned_map()
    assert b.add_artificial_areas(POLYS, -0.3, 0.45) == O.OK
    art = art_voxels(b)
    assert as_dict(a) == as_dict(b)  # nothing changes before the next update
    assert b.integrate() == O.OK     # no pending update: updateMap returns early
    assert as_dict(a) == as_dict(b)
    rng = np.random.default_rng(9)
    for k in range(2):
        o = rng.uniform(-0.5, 0.5, 3)
        pts = (o + rng.normal(0, 1.0, (1500, 3))).astype(np.float32)
        assert a.insert("s", pts, o) == O.OK
        assert b.insert("s", pts, o) == O.OK
    da, db = as_dict(a), as_dict(b)
    for c, (bits, act) in db.items():
        if c in art:
            assert act
            assert bits == da.get(c, (0, False))[0]
        else:
            assert (bits, act) == da.get(c, (0, False))
    assert art <= set(db)


def test_restore_map_integrity_sets_threshold_state():
    m = scanned_map()
    assert m.add_artificial_areas(POLYS, -0.3, 0.45) == O.OK
    art = art_voxels(m)
    assert m.insert("s", np.array([[0.3, 0.2, 0.1]], np.float32), [0, 0, 0]) == O.OK
    t_max = m.logodds()[3]
    before = as_dict(m)
    assert m.restore_map_integrity() == O.OK
    after = as_dict(m)
    assert art_voxels(m) == set()
    for c, (bits, act) in after.items():
        if c in art:
            v = np.array([bits], np.uint32).view(np.float32)[0]
            assert (bits, act) == (before[c][0], bool(v > t_max))
        else:
            assert (bits, act) == before[c]


def test_add_artificial_areas_restores_previous_areas_first():
    m = scanned_map()
    assert m.add_artificial_areas(POLYS[:1], -0.2, 0.2) == O.OK
    first = art_voxels(m)
    assert m.insert("s", np.array([[0.3, 0.2, 0.1]], np.float32), [0, 0, 0]) == O.OK
    assert m.add_artificial_areas(POLYS[1:2], -0.1, 0.1) == O.OK
    assert art_voxels(m) == expected_walls(POLYS[1:2], -0.1, 0.1, 0.1)
    t_max = m.logodds()[3]
    d = as_dict(m)
    for c in first - art_voxels(m):
        v = np.array([d[c][0]], np.uint32).view(np.float32)[0]
        assert d[c][1] == bool(v > t_max)


# ------------------------------------------------------------ map sections --

def crop(m, bbox, full):
    """Brute-force section of the map export."""
    o, v, a = m.export_leaves()
    lo, hi = np.array(bbox[:3]), np.array(bbox[3:])
    n = np.arange(512)
    off = np.stack([n >> 6, (n >> 3) & 7, n & 7], axis=1)
    ro, rv, ra = [], [], []
    for k in range(o.shape[0]):
        c = o[k] + off
        inside = ((c >= lo) & (c <= hi)).all(axis=1)
        act = np.array([(int(a[k, j >> 6]) >> (j & 63)) & 1 for j in range(512)], bool)
        if full:
            keep = inside & (act | (v[k] != 0))
            vals = np.where(inside, v[k], 0).astype(np.float32)
        else:
            keep = inside & act
            vals = np.where(inside & act, 1.0, 0).astype(np.float32)
        if not keep.any():
            continue
        bits = inside & act
        words = np.zeros(8, np.uint64)
        for j in np.nonzero(bits)[0]:
            words[j >> 6] |= np.uint64(1 << int(j & 63))
        ro.append(o[k])
        rv.append(vals)
        ra.append(words)
    if not ro:
        return np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32), np.zeros((0, 8), np.uint64)
    return np.array(ro, np.int32), np.array(rv, np.float32), np.array(ra, np.uint64)


BOXES = [(-7, -3, -9, 12, 20, 5), (0, 0, 0, 0, 0, 0), (3, -40, -40, 2, 40, 40), (-100, -100, -100, 100, 100, 100)]


@pytest.mark.parametrize("bbox", BOXES)
@pytest.mark.parametrize("full", [True, False])
def test_export_section_matches_crop(bbox, full):
    m = scanned_map()
    m.add_points_to_grid(np.random.default_rng(3).uniform(-2, 2, (300, 3)).astype(np.float32))
    got = m.export_section(bbox, full)
    want = crop(m, bbox, full)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)


def test_section_round_trip_into_empty_map():
    m = scanned_map()
    bbox = (-15, -12, -9, 14, 17, 10)
    so, sv, sa = m.export_section(bbox, True)
    e = mk()
    assert e.apply_section(so, sv, sa) == O.OK
    o, v, a = e.export_leaves()
    assert np.array_equal(o, so) and np.array_equal(v, sv) and np.array_equal(a, sa)
    for g, w in zip(e.export_section(bbox, True), (so, sv, sa)):
        assert np.array_equal(g, w)


def test_apply_section_overwrites_whole_leaves():
    """Every voxel of a section leaf is written, also those outside the box
    the section was cut with (they carry background 0, inactive)."""
    m = scanned_map()
    bbox = (-3, -3, -3, 3, 3, 3)
    so, sv, sa = m.export_section(bbox, True)
    t = scanned_map(seed=11)
    before = as_dict(t)
    assert t.apply_section(so, sv, sa) == O.OK
    after = as_dict(t)
    for k in range(so.shape[0]):
        bits = sv[k].view(np.uint32)
        for n in range(512):
            c = (int(so[k, 0]) + (n >> 6), int(so[k, 1]) + ((n >> 3) & 7), int(so[k, 2]) + (n & 7))
            assert after[c] == (int(bits[n]), bool((int(sa[k, n >> 6]) >> (n & 63)) & 1))
    touched = {tuple(int(x) for x in r) for r in so}
    for c, st in after.items():
        if tuple((x // 8) * 8 for x in c) not in touched and not all(x % 8 == 0 for x in c):
            assert st == before.get(c, (0, False))


def test_apply_section_visits_internal_tiles():
    """A section with a single leaf at (8, 0, 0): its 128^3 node has an
    inactive tile at every other 8^3 slot, its 4096^3 node one at every other
    128^3 slot. Each visit writes (0, off) at the tile origin of the map."""
    t = mk()
    probe = np.array([[0.0, 0.0, 0.0], [0.15, 0.0, 0.0], [12.8, 0.0, 0.0], [13.65, 0.0, 0.0],
                      [0.8, 0.0, 0.0]], np.float32) + 0.001
    assert t.add_points_to_grid(probe) == O.OK
    sv = np.zeros((1, 512), np.float32)
    sa = np.zeros((1, 8), np.uint64)
    sv[0, 0] = 2.5
    sa[0, 0] = 1
    assert t.apply_section(np.array([[8, 0, 0]], np.int32), sv, sa) == O.OK
    assert t.get_value([0, 0, 0]) == (0.0, False)      # 8^3 tile origin of the section's 128^3 node
    assert t.get_value([0, 0, 0])[0] == 0.0
    assert t.get_value([1, 0, 0]) == (C_MAX, True)     # not a tile origin
    assert t.get_value([128, 0, 0]) == (0.0, False)    # 128^3 tile origin of its 4096^3 node
    assert t.get_value([136, 0, 0]) == (C_MAX, True)   # inside that tile, not its origin
    assert t.get_value([8, 0, 0]) == (np.float32(2.5), True)  # the section leaf itself


def test_apply_section_update():
    m = scanned_map()
    src = scanned_map(seed=21)
    bbox = (-8, -8, -8, 9, 9, 9)
    so, sv, sa = src.export_section(bbox, False)
    before = as_dict(m)
    assert m.apply_section_update(bbox, so, sa) == O.OK
    after = as_dict(m)
    sect = set()
    for k in range(so.shape[0]):
        for n in range(512):
            if (int(sa[k, n >> 6]) >> (n & 63)) & 1:
                sect.add((int(so[k, 0]) + (n >> 6), int(so[k, 1]) + ((n >> 3) & 7), int(so[k, 2]) + (n & 7)))
    for c, (bits, act) in after.items():
        inside = all(bbox[i] <= c[i] <= bbox[3 + i] for i in range(3))
        b0, a0 = before.get(c, (0, False))
        assert bits == b0
        assert act == (c in sect or (a0 and not inside))


# ------------------------------------------------------------------ bboxes --

def test_index_bbox_identity_and_rotation():
    assert O.index_bbox(0.5, [0, 0, 0], [1, 2, 3]).tolist() == [0, 0, 0, 2, 4, 6]
    rz = np.array([[0, -1, 0, 0], [1, 0, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)
    assert O.index_bbox(0.5, [0, 0, 0], [1, 2, 3], rz).tolist() == [-4, 0, 0, 0, 2, 6]
    tr = np.eye(4)
    tr[:3, 3] = [10.25, -3.5, 0.75]
    assert O.index_bbox(0.25, [-1, -1, -1], [1, 1, 1], tr).tolist() == [37, -18, -1, 45, -10, 7]
    # corners are rounded to float first: 0.1f * 10 > 1 -> index 1
    assert O.index_bbox(0.1, [0.1, 0.1, 0.1], [0.1, 0.1, 0.1]).tolist() == [1, 1, 1, 1, 1, 1]


def test_walls_and_polygons_append_without_restore():
    m = scanned_map()
    assert m.add_artificial_areas(POLYS[:1], -0.2, 0.2) == O.OK
    first = art_voxels(m)
    a, b = np.array([2.05, 2.0, -0.3]), np.array([-2.0, 2.55, -0.3])
    assert m.add_artificial_wall(a, b, -0.1, 0.3) == O.OK
    wall = set()
    for k in range(int(-0.1 / 0.1), int(0.3 / 0.1)):
        dz = np.array([0, 0, k], np.int32)
        wall |= {tuple(int(x) for x in c)
                 for c in O.cast_ray(O.world_to_index(0.1, a) + dz, O.world_to_index(0.1, b) + dz)}
    assert art_voxels(m) == first | wall
    assert m.add_artificial_polygon(POLYS[0], 0.0, 0.1) == O.OK
    assert art_voxels(m) == first | wall | expected_walls(POLYS[:1], 0.0, 0.1, 0.1)


def test_wall_direction_is_kept():
    """castRayIntoGrid(s, e) and (e, s) can take different voxels on exact
    ties; a wall is cast from its start vertex."""
    m1, m2 = mk(1.0), mk(1.0)
    a, b = np.array([0.5, 0.5, 0.5]), np.array([3.5, 1.5, 0.5])
    assert m1.add_artificial_wall(a, b, 0.0, 1.0) == O.OK
    assert m2.add_artificial_wall(b, a, 0.0, 1.0) == O.OK
    s, e = O.world_to_index(1.0, a), O.world_to_index(1.0, b)
    assert art_voxels(m1) == {tuple(int(x) for x in c) for c in O.cast_ray(s, e)}
    assert art_voxels(m2) == {tuple(int(x) for x in c) for c in O.cast_ray(e, s)}


# ------------------------------------------------------------ persistence --

def tree_key(c):
    return (c[0] >> 12, c[1] >> 12, c[2] >> 12, c[0] >> 7, c[1] >> 7, c[2] >> 7, c[0] >> 3, c[1] >> 3, c[2] >> 3,
            ((c[0] & 7) << 6) | ((c[1] & 7) << 3) | (c[2] & 7))


def test_active_points_closed_form_and_tree_order():
    m = scanned_map()
    m.add_points_to_grid(np.random.default_rng(8).uniform(-3, 3, (400, 3)).astype(np.float32))
    m.add_points_to_grid(np.array([[-410.0, 0.3, 7.2], [409.7, -0.2, -6.1]], np.float32))  # other root children
    act = sorted((c for c, (b, a) in as_dict(m).items() if a), key=tree_key)
    want = np.array([[np.float32(ci * 0.1 + 0.05) for ci in c] for c in act], np.float32)
    got = m.export_active_points()
    assert got.shape == want.shape and got.shape[0] > 400
    assert np.array_equal(got, want)


def test_create_map_from_points():
    rng = np.random.default_rng(12)
    pts = rng.uniform(-0.9, 1.3, (300, 3)).astype(np.float32)
    cells = {vox(p, 0.1) for p in pts}
    m = mk()
    assert m.create_map_from_points(pts, False, False) == O.OK
    d = as_dict(m)
    assert {c for c, (b, a) in d.items() if a} == cells
    assert all(d[c] == (f32bits(C_MAX), True) for c in cells)
    assert leaf_set(m) == {tuple((x // 8) * 8 for x in c) for c in cells}
    m2 = mk()
    assert m2.create_map_from_points(pts, True, False) == O.OK
    lo = [min(c[a] for c in cells) for a in range(3)]
    hi = [max(c[a] for c in cells) for a in range(3)]
    d2 = as_dict(m2)
    for c, st in d2.items():
        inside = all(lo[a] <= c[a] <= hi[a] for a in range(3))
        if c in cells:
            assert st == (f32bits(C_MAX), True)
        elif inside:
            assert st == (f32bits(C_MIN), False)
        else:
            assert st == (0, False)
    leaves = {(x, y, z) for x in range((lo[0] // 8) * 8, hi[0] + 1, 8) for y in range((lo[1] // 8) * 8, hi[1] + 1, 8)
              for z in range((lo[2] // 8) * 8, hi[2] + 1, 8)}
    assert leaf_set(m2) == leaves


def test_create_map_from_points_into_a_scanned_map():
    m = scanned_map()
    before = as_dict(m)
    act0 = {c for c, (b, a) in before.items() if a}
    pts = np.array([[0.31, -0.22, 0.05], [0.75, 0.12, 0.33]], np.float32)
    assert m.create_map_from_points(pts, True, False) == O.OK
    new = {vox(p, 0.1) for p in pts}
    allact = act0 | new
    lo = [min(c[a] for c in allact) for a in range(3)]
    hi = [max(c[a] for c in allact) for a in range(3)]
    for c, st in as_dict(m).items():
        inside = all(lo[a] <= c[a] <= hi[a] for a in range(3))
        if c in new:
            assert st == (f32bits(C_MAX), True)
        elif c in act0:
            assert st == before[c]
        elif inside:
            assert st == (f32bits(C_MIN), False)
        else:
            assert st == before.get(c, (0, False))
    m3 = scanned_map()
    assert m3.create_map_from_points(pts, False, True) == O.OK  # clear first
    assert {c for c, (b, a) in as_dict(m3).items() if a} == new
    bad = np.array([[np.nan, 0, 0]], np.float32)
    assert m3.create_map_from_points(bad, True, True) == O.OUT_OF_RANGE
    assert {c for c, (b, a) in as_dict(m3).items() if a} == new  # unchanged
