"""Direct map edits, artificial areas and map sections on the GPU against the
oracle, bit for bit: addPointsToGrid / removePointsFromGrid
(VDBMapping.hpp:413-447), addArtificialAreas / restoreMapIntegrity
(:1152-1236) with the activation at the end of every non-empty map update
(:785-789) on the brick, update-set, batch and fast-mode paths, and
getMapSection / applyMapSectionGrid / applyMapSectionUpdateGrid (:884-1085)."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from parity import both, pair, same_map
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu

POLYS = [np.array([[-1.0, -1.0, 0.0], [1.2, -1.0, 0.0], [1.2, 0.8, 0.0]]),
         np.array([[2.05, 2.0, -0.3], [-2.0, 2.55, -0.3]]),
         np.array([[0.33, 0.33, 0.1]])]


def same_art(g, o):
    go, gm = g.export_artificial()
    oo, om = o.export_artificial()
    assert np.array_equal(go, oo) and np.array_equal(gm, om)
    return go.shape[0]


def same_section(g, o, bbox, full):
    gs = g.export_section(bbox, full)
    for a, b in zip(gs, o.export_section(bbox, full)):
        assert a.shape == b.shape
        assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                              b.view(np.uint32) if b.dtype == np.float32 else b)
    return gs[0].shape[0]


def scanned(traversal=0, seed=5, res=0.1, R=5.0, fast=False):
    g, o = pair(res, max_range=R, fast_mode=fast, **W.PROBS)
    assert g.set_tuning(traversal=traversal) == V.OK
    rng = np.random.default_rng(seed)
    for _ in range(3):
        orig = rng.uniform(-0.5, 0.5, 3)
        pts = (orig + rng.normal(0, 1.2, (3000, 3))).astype(np.float32)
        both(g, o, "insert", "test", pts, orig)
    return g, o, rng


@pytest.mark.parametrize("traversal", [0, 1])
def test_add_and_remove_points(traversal):
    g, o, rng = scanned(traversal)
    add = rng.uniform(-3, 3, (5000, 3)).astype(np.float32)
    both(g, o, "add_points_to_grid", add)
    same_map(g, o)
    rem = np.concatenate([add[::3], rng.uniform(-3, 3, (700, 3)).astype(np.float32)])
    both(g, o, "remove_points_from_grid", rem)
    same_map(g, o)
    # edits followed by scans and more edits, streamed
    for k in range(4):
        orig = rng.uniform(-0.5, 0.5, 3)
        both(g, o, "insert", "test", (orig + rng.normal(0, 1.5, (4000, 3))).astype(np.float32), orig)
        both(g, o, ["add_points_to_grid", "remove_points_from_grid"][k % 2],
             rng.uniform(-4, 4, (300, 3)).astype(np.float32))
    same_map(g, o)


def test_point_edit_errors_and_empty_map():
    g, o = pair(0.1, max_range=5.0, **W.PROBS)
    both(g, o, "remove_points_from_grid", np.array([[0.05, 0.05, 0.05]], np.float32))
    same_map(g, o)
    assert g.empty() == o.empty() is False
    bad = np.array([[0.5, 0.5, 0.5], [np.nan, 0, 0]], np.float32)
    assert g.add_points_to_grid(bad) == o.add_points_to_grid(bad) == V.OUT_OF_RANGE
    far = np.array([[1e6, 0, 0]], np.float32)
    assert g.add_points_to_grid(far) == o.add_points_to_grid(far) == V.OUT_OF_RANGE
    same_map(g, o)
    both(g, o, "add_points_to_grid", np.zeros((0, 3), np.float32))
    # large edit: the edit set grows
    pts = np.random.default_rng(4).uniform(-40, 40, (200000, 3)).astype(np.float32)
    both(g, o, "add_points_to_grid", pts)
    same_map(g, o)


@pytest.mark.parametrize("neg,pos", [(-0.3, 0.45), (0.0, 0.25), (0.2, 0.1)])
def test_artificial_grid(neg, pos):
    g, o = pair(0.1, max_range=5.0, **W.PROBS)
    both(g, o, "add_artificial_areas", POLYS, neg, pos)
    same_art(g, o)


@pytest.mark.parametrize("traversal", [0, 1])
def test_artificial_areas_through_inserts_and_restore(traversal):
    g, o, rng = scanned(traversal)
    both(g, o, "add_artificial_areas", POLYS, -0.3, 0.45)
    assert same_art(g, o) > 10
    same_map(g, o)           # nothing happens before the next update
    both(g, o, "integrate")  # no pending update: no activation
    same_map(g, o)
    for k in range(3):
        orig = rng.uniform(-0.5, 0.5, 3)
        both(g, o, "insert", "test", (orig + rng.normal(0, 1.0, (2500, 3))).astype(np.float32), orig)
    same_map(g, o)
    both(g, o, "add_artificial_areas", POLYS[1:2], -0.1, 0.2)  # restores the first set
    same_art(g, o)
    same_map(g, o)
    orig = rng.uniform(-0.5, 0.5, 3)
    both(g, o, "insert", "test", (orig + rng.normal(0, 1.0, (2500, 3))).astype(np.float32), orig)
    same_map(g, o)
    both(g, o, "restore_map_integrity")
    same_map(g, o)
    assert g.export_artificial()[0].shape[0] == 0


def test_artificial_areas_with_batch_accumulate_and_reset():
    g, o, rng = scanned(1)
    both(g, o, "add_artificial_areas", POLYS, -0.2, 0.3)
    clouds, origins = [], []
    for k in range(4):
        orig = rng.uniform(-0.5, 0.5, 3)
        clouds.append((orig + rng.normal(0, 1.0, (2000, 3))).astype(np.float32))
        origins.append(orig)
    assert g.insert_batch("test", clouds, origins) == V.OK
    for c, orig in zip(clouds, origins):
        assert o.insert("test", c, orig) == O.OK
    same_map(g, o)
    both(g, o, "accumulate", "test", clouds[0], origins[0])
    both(g, o, "integrate")
    same_map(g, o)
    # resetMap keeps the artificial grid; the next update re-creates its leaves
    assert g.reset() == V.OK
    o.reset()
    both(g, o, "insert", "test", clouds[1], origins[1])
    same_map(g, o)


def test_artificial_areas_fast_mode():
    g, o, rng = scanned(0, fast=True)
    both(g, o, "add_artificial_areas", POLYS, -0.3, 0.45)
    for k in range(3):
        orig = rng.uniform(-0.5, 0.5, 3)
        both(g, o, "insert", "test", (orig + rng.normal(0, 1.0, (2500, 3))).astype(np.float32), orig)
        same_map(g, o)


def test_artificial_area_errors():
    g, o = pair(0.1, max_range=5.0, **W.PROBS)
    bad = [np.array([[0.0, 0.0, 0.0], [np.nan, 1.0, 0.0]])]
    assert g.add_artificial_areas(bad, 0.0, 0.3) == o.add_artificial_areas(bad, 0.0, 0.3) == V.OUT_OF_RANGE
    same_art(g, o)
    assert g.add_artificial_areas(POLYS, -1e300, 0.3) == V.BAD_ARGUMENT
    assert o.add_artificial_areas(POLYS, -1e300, 0.3) == O.BAD_ARGUMENT
    both(g, o, "add_artificial_areas", [], 0.0, 1.0)
    same_art(g, o)


BOXES = [(-7, -3, -9, 12, 20, 5), (0, 0, 0, 0, 0, 0), (3, -40, -40, 2, 40, 40),
         (-100, -100, -100, 100, 100, 100), (-13, 2, -5, 30, 9, 17)]


@pytest.mark.parametrize("bbox", BOXES)
def test_export_sections(bbox):
    g, o, rng = scanned(0)
    both(g, o, "add_points_to_grid", rng.uniform(-2, 2, (300, 3)).astype(np.float32))
    for full in (True, False):
        same_section(g, o, bbox, full)


def test_apply_section_round_trip_and_tiles():
    src_g, src_o, rng = scanned(0, seed=8)
    dst_g, dst_o, _ = scanned(0, seed=13)
    for bbox in BOXES[:3]:
        so, sv, sa = src_o.export_section(bbox, True)
        go, gv, ga = src_g.export_section(bbox, True)
        assert np.array_equal(so, go) and np.array_equal(sv.view(np.uint32), gv.view(np.uint32))
        both(dst_g, dst_o, "apply_section", so, sv, sa)
        same_map(dst_g, dst_o)
    # a sparse section (values 1.0) applied as a grid, and the internal-tile visits
    so, sv, sa = src_o.export_section(BOXES[4], False)
    both(dst_g, dst_o, "apply_section", so, sv, sa)
    same_map(dst_g, dst_o)
    t_g, t_o = pair(0.1, max_range=5.0, **W.PROBS)
    probe = np.array([[0.0, 0.0, 0.0], [0.15, 0.0, 0.0], [12.8, 0.0, 0.0], [13.65, 0.0, 0.0],
                      [0.8, 0.0, 0.0]], np.float32) + 0.001
    both(t_g, t_o, "add_points_to_grid", probe)
    v = np.zeros((1, 512), np.float32)
    a = np.zeros((1, 8), np.uint64)
    v[0, 0], a[0, 0] = 2.5, 1
    both(t_g, t_o, "apply_section", np.array([[8, 0, 0]], np.int32), v, a)
    same_map(t_g, t_o)
    assert t_g.export_leaves()[2][:, 0].tolist() == t_o.export_leaves()[2][:, 0].tolist()


def test_apply_section_errors():
    g, o = pair(0.1, max_range=5.0, **W.PROBS)
    v = np.zeros((2, 512), np.float32)
    a = np.zeros((2, 8), np.uint64)
    assert g.apply_section(np.array([[8, 0, 0], [8, 0, 0]], np.int32), v, a) == V.BAD_ARGUMENT
    assert g.apply_section(np.array([[8, 0, 0], [4, 0, 0]], np.int32), v, a) == V.BAD_ARGUMENT
    assert g.apply_section(np.array([[8, 0, 0], [1 << 23, 0, 0]], np.int32), v, a) == V.OUT_OF_RANGE
    same_map(g, o)
    both(g, o, "apply_section", np.zeros((0, 3), np.int32), np.zeros((0, 512), np.float32),
         np.zeros((0, 8), np.uint64))
    # an all-zero inactive section leaf creates nothing
    both(g, o, "apply_section", np.array([[16, 0, 0]], np.int32), v[:1], a[:1])
    same_map(g, o)
    assert g.empty() and o.empty()


def test_apply_section_update():
    g, o, rng = scanned(0)
    src = scanned(0, seed=21)[1]
    for bbox in BOXES[:3]:
        so, sv, sa = src.export_section(bbox, False)
        both(g, o, "apply_section_update", bbox, so, sa)
        same_map(g, o)
    both(g, o, "apply_section_update", BOXES[3], np.zeros((0, 3), np.int32), np.zeros((0, 8), np.uint64))
    same_map(g, o)


def test_index_bbox_matches_oracle():
    rng = np.random.default_rng(17)
    for _ in range(200):
        lo = rng.uniform(-5, 5, 3)
        hi = lo + rng.uniform(0, 4, 3)
        q = rng.normal(size=4)
        q /= np.linalg.norm(q)
        w, x, y, z = q
        R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                      [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                      [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])
        tf = np.eye(4)
        tf[:3, :3] = R
        tf[:3, 3] = rng.uniform(-10, 10, 3)
        res = float(rng.choice([0.05, 0.1, 0.25]))
        assert np.array_equal(V.index_bbox(res, lo, hi, tf), O.index_bbox(res, lo, hi, tf))


def test_sections_between_gpu_maps_outdoor():
    """A section of a streamed outdoor map (262,144-point scans at 0.05 m)
    moved into a second map, GPU against oracle."""
    w = W.outdoor_stream_workload(2)
    g, o = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    for i in range(2):
        s = w.scan(i)
        both(g, o, "insert", "test", s.points, s.origin)
    bbox = V.index_bbox(w.resolution, [-10, -10, -2], [10, 10, 3])
    assert same_section(g, o, bbox, True) > 1000
    same_section(g, o, bbox, False)
    so, sv, sa = g.export_section(bbox, True)
    g2, o2 = pair(w.resolution, max_range=w.max_range, **W.PROBS)
    both(g2, o2, "apply_section", so, sv, sa)
    same_map(g2, o2)


@pytest.mark.parametrize("traversal", [0, 1])
def test_walls_and_polygons_append(traversal):
    g, o, rng = scanned(traversal)
    both(g, o, "add_artificial_areas", POLYS[:1], -0.2, 0.2)
    both(g, o, "add_artificial_wall", [2.05, 2.0, -0.3], [-2.0, 2.55, -0.3], -0.1, 0.3)
    same_art(g, o)
    orig = rng.uniform(-0.5, 0.5, 3)
    both(g, o, "insert", "test", (orig + rng.normal(0, 1.0, (2500, 3))).astype(np.float32), orig)
    same_map(g, o)
    both(g, o, "add_artificial_polygon", POLYS[0] * 1.7, 0.0, 0.1)
    walls = [(rng.uniform(-3, 3, 3), rng.uniform(-3, 3, 3)) for _ in range(40)]
    both(g, o, "add_artificial_walls", walls, -0.25, 0.15)
    same_art(g, o)
    both(g, o, "insert", "test", (orig + rng.normal(0, 1.0, (2500, 3))).astype(np.float32), orig)
    same_map(g, o)
    bad = [(np.array([0.0, 0.0, 0.0]), np.array([np.nan, 0.0, 0.0]))]
    assert g.add_artificial_walls(bad, 0.0, 0.3) == o.add_artificial_walls(bad, 0.0, 0.3) == V.OUT_OF_RANGE
    same_art(g, o)
    both(g, o, "restore_map_integrity")
    same_map(g, o)


def test_active_points_export():
    g, o, rng = scanned(0)
    both(g, o, "add_points_to_grid", rng.uniform(-3, 3, (2000, 3)).astype(np.float32))
    both(g, o, "add_points_to_grid", np.array([[-410.0, 0.3, 7.2], [409.7, -0.2, -6.1]], np.float32))
    gp, op = g.export_active_points(), o.export_active_points()
    assert gp.shape == op.shape and gp.shape[0] > 1900
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))


@pytest.mark.parametrize("set_background,clear_map", [(False, False), (True, False), (True, True), (False, True)])
def test_create_map_from_points(set_background, clear_map):
    g, o, rng = scanned(0)
    pts = rng.uniform(-1.5, 2.0, (5000, 3)).astype(np.float32)
    both(g, o, "create_map_from_points", pts, set_background, clear_map)
    same_map(g, o)
    assert g.empty() == o.empty()
    # the map keeps integrating afterwards
    orig = rng.uniform(-0.5, 0.5, 3)
    both(g, o, "insert", "test", (orig + rng.normal(0, 1.0, (2500, 3))).astype(np.float32), orig)
    same_map(g, o)
    bad = np.array([[0.1, 0.2, 0.3], [np.nan, 0, 0]], np.float32)
    assert g.create_map_from_points(bad, True, True) == o.create_map_from_points(bad, True, True) == V.OUT_OF_RANGE
    same_map(g, o)


def test_save_load_round_trip_through_points():
    """saveMapToPCD's points loaded back with createMapFromPointCloud give
    the same active voxels (no background, cleared map)."""
    g, o, rng = scanned(0)
    both(g, o, "add_points_to_grid", rng.uniform(-3, 3, (3000, 3)).astype(np.float32))
    pts = g.export_active_points()
    g2, o2 = pair(0.1, max_range=5.0, **W.PROBS)
    both(g2, o2, "create_map_from_points", pts, False, True)
    same_map(g2, o2)
    act = lambda m: {tuple(r) for r in (m.export_active_points() * 1).tolist()}
    assert act(g2) == act(g)
