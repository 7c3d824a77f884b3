"""pruneGrid at the end of createMapFromPointCloud (OccupancyVDBMapping.hpp:173)
on the GPU against the oracle, bit for bit: the leaf-level tiles it forms,
leaves and tiles after scans on every insert path (brick, update-set, batch,
fast mode), point edits, artificial areas, map sections, the saveMapToPCD point
list and raytrace through a pruned map."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from parity import both, pair, same_map
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu

RES = 0.1


def same_tiles(g, o):
    go, gv, ga = g.export_tiles()
    oo, ov, oa = o.export_tiles()
    assert go.shape == oo.shape, (go.shape, oo.shape)
    assert np.array_equal(go, oo)
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(ga, oa)
    return go.shape[0]


def same_all(g, o):
    n = same_map(g, o)
    t = same_tiles(g, o)
    assert g.leaf_count() == o.L.orc_leaf_count(o.h) == n
    gp, op = g.export_active_points(), o.export_active_points()
    assert np.array_equal(gp.view(np.uint32), op.view(np.uint32))
    return n, t


def block_points():
    """Leaf (0,0,0) full of points, one more voxel at (23, 0, 0): with the
    background fill, an active tile, an inactive tile and one leaf."""
    gr = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    vox = np.concatenate([gr, [[23, 0, 0]]])
    return ((vox + 0.5) * RES).astype(np.float32)


def room_points(rng):
    """A 9.6 x 9.6 x 4.8 m room at 0.1 m: walls one voxel thick with holes, a
    solid leaf-aligned 1.6 m block (full leaves: active tiles) and scattered
    points; the background fill makes the free interior inactive tiles."""
    n = np.array([96, 96, 48])
    vox = []
    for a in range(3):
        dims = [d for k, d in enumerate(n) if k != a]
        gr = np.stack(np.meshgrid(np.arange(dims[0]), np.arange(dims[1]), indexing="ij"), -1).reshape(-1, 2)
        gr = gr[rng.random(len(gr)) < 0.93]
        for side in (0, n[a] - 1):
            vox.append(np.insert(gr, a, side, axis=1))
    blk = np.stack(np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij"), -1).reshape(-1, 3)
    vox.append(blk + [64, 8, 8])
    vox.append(rng.integers(0, n, (60, 3)))
    vox = np.unique(np.concatenate(vox), axis=0)
    return ((vox + 0.5) * RES).astype(np.float32), n * RES


def room_scans(rng, size, k, n_pts=3000):
    """Scans from near the room centre: most points on the walls, a few inside."""
    scans = []
    for _ in range(k):
        org = size / 2 + rng.uniform(-0.6, 0.6, 3)
        face = rng.integers(0, 6, n_pts)
        p = rng.uniform(0.05, 1.0, (n_pts, 3)) * (size - 0.1)
        ax = face % 3
        p[np.arange(n_pts), ax] = np.where(face < 3, 0.05, size[ax] - 0.05)
        inner = rng.random(n_pts) < 0.05
        p[inner] = org + rng.normal(0, 0.8, (inner.sum(), 3))
        scans.append((p.astype(np.float32), org))
    return scans


def test_block_tiles_and_their_transitions():
    g, o = pair(RES, max_range=5.0, **W.PROBS)
    both(g, o, "create_map_from_points", block_points(), True, False)
    n, t = same_all(g, o)
    assert (n, t) == (1, 2)
    # hits inside the inactive tile change voxels (a leaf again); a hit inside
    # the active tile is clamped back (still a tile)
    both(g, o, "insert", "test", np.array([[1.425, 0.325, 0.325]], np.float32), [1.225, 0.325, 0.325])
    assert same_all(g, o) == (2, 1)
    both(g, o, "insert", "test", np.array([[0.325, 0.325, 0.325]], np.float32), [0.325, 0.325, 0.325])
    assert same_all(g, o) == (2, 1)
    both(g, o, "insert", "test", np.array([[0.325, 0.325, 0.725]], np.float32), [0.325, 0.325, 0.025])
    assert same_all(g, o) == (3, 0)


def test_block_tiles_point_edits_and_sections():
    g, o = pair(RES, max_range=5.0, **W.PROBS)
    both(g, o, "create_map_from_points", block_points(), True, False)
    both(g, o, "remove_points_from_grid", np.array([[1.05, 0.05, 0.05]], np.float32))
    assert same_all(g, o) == (1, 2)
    both(g, o, "add_points_to_grid", np.array([[1.05, 0.05, 0.05]], np.float32))
    assert same_all(g, o) == (2, 1)
    for full in (True, False):
        a, b = g.export_section([0, 0, 0, 30, 7, 7], full), o.export_section([0, 0, 0, 30, 7, 7], full)
        assert all(np.array_equal(x, y) for x, y in zip(a, b))
        assert a[0].tolist() == [[8, 0, 0], [16, 0, 0]]  # the active tile is not a section leaf
    # the active tile's origin outside the box: untouched; inside: only that voxel goes off
    none = (np.zeros((0, 3), np.int32), np.zeros((0, 8), np.uint64))
    both(g, o, "apply_section_update", [1, 0, 0, 30, 7, 7], *none)
    assert same_all(g, o)[1] == 1
    both(g, o, "apply_section_update", [0, 0, 0, 30, 7, 7], *none)
    assert same_all(g, o)[1] == 0


def test_reset_and_clear_drop_tiles():
    g, o = pair(RES, max_range=5.0, **W.PROBS)
    both(g, o, "create_map_from_points", block_points(), True, False)
    assert same_all(g, o)[1] == 2
    both(g, o, "create_map_from_points", block_points()[:5], False, True)
    assert same_all(g, o)[1] == 0
    both(g, o, "create_map_from_points", block_points(), True, False)
    assert g.reset() == V.OK
    o.reset()
    assert same_all(g, o) == (0, 0)


@pytest.mark.parametrize("traversal", [0, 1])
def test_pruned_room_then_scans_edits_and_areas(traversal):
    rng = np.random.default_rng(41 + traversal)
    g, o = pair(RES, max_range=8.0, **W.PROBS)
    assert g.set_tuning(traversal=traversal) == V.OK
    pts, size = room_points(rng)
    both(g, o, "create_map_from_points", pts, True, True)
    n0, t0 = same_all(g, o)
    assert t0 > 300, t0
    # the dirty export lists leaves only
    do = g.export_dirty_leaves()[0]
    tiles = {tuple(r) for r in g.export_tiles()[0].tolist()}
    assert not tiles & {tuple(r) for r in do.tolist()}
    for pts_s, org in room_scans(rng, size, 4):
        both(g, o, "insert", "test", pts_s, org)
        same_all(g, o)
    # several scans in one batch: scan by scan on a pruned map
    scans = room_scans(rng, size, 3)
    clouds = [np.concatenate([p, np.ones((len(p), 1), np.float32)], 1) for p, _ in scans]
    origins = np.array([org for _, org in scans])
    ranges = np.array([8.0, 4.5, 6.5])
    assert g.insert_batch("test", clouds, origins, ranges, check=True) == V.OK
    for c, org, r in zip(clouds, origins, ranges):
        o.add_input_source("test", r)
        o.insert("test", c, org)
    assert g.add_input_source("test", 8.0) == o.add_input_source("test", 8.0)
    same_all(g, o)
    # artificial areas: activation after every scan
    polys = [np.array([[1.6, 1.6, 0.0], [7.9, 1.7, 0.0], [6.5, 7.8, 0.0]])]
    both(g, o, "add_artificial_areas", polys, -0.3, 1.2)
    for pts_s, org in room_scans(rng, size, 2):
        both(g, o, "insert", "test", pts_s, org)
        same_all(g, o)
    both(g, o, "restore_map_integrity")
    same_all(g, o)
    both(g, o, "add_points_to_grid", rng.uniform(0, 9.6, (300, 3)).astype(np.float32))
    both(g, o, "remove_points_from_grid", rng.uniform(0, 9.6, (300, 3)).astype(np.float32))
    n1, t1 = same_all(g, o)
    assert 0 < t1 < t0
    # sections of the pruned map, and applied back
    box = [3, 3, 3, 70, 60, 30]
    for full in (True, False):
        a, b = g.export_section(box, full), o.export_section(box, full)
        assert all(np.array_equal(x.view(np.uint32) if x.dtype == np.float32 else x,
                                  y.view(np.uint32) if y.dtype == np.float32 else y) for x, y in zip(a, b))
    so, sv, sm = o.export_section(box, True)
    g2, o2 = pair(RES, max_range=8.0, **W.PROBS)
    both(g2, o2, "create_map_from_points", pts, True, True)
    both(g2, o2, "apply_section", so, sv, sm)
    same_all(g2, o2)
    so, _, sm = o.export_section(box, False)
    both(g2, o2, "apply_section_update", box, so, sm)
    same_all(g2, o2)


def test_fast_mode_and_raytrace_through_tiles():
    rng = np.random.default_rng(77)
    g, o = pair(RES, max_range=8.0, fast_mode=True, **W.PROBS)
    pts, size = room_points(rng)
    both(g, o, "create_map_from_points", pts, True, True)
    both(g, o, "integrate")  # the intersector of the pruned map
    same_all(g, o)

    def queries(n):
        org = rng.uniform(-1.0, 10.6, (n, 3))
        tgt = rng.uniform(0.3, 9.3, (n, 3))
        d = tgt - org
        ln = np.linalg.norm(d, axis=1) * rng.uniform(0.5, 1.5, n)
        return org, d / np.linalg.norm(d, axis=1, keepdims=True), ln

    def same_raytrace():
        q = queries(4000)
        sg, okg, eg = g.raytrace(*q)
        so, oko, eo = o.raytrace(*q)
        assert sg == so == 0, (sg, so, g.last_error())
        assert np.array_equal(okg, oko)
        assert np.array_equal(eg[okg].view(np.uint64), eo[oko].view(np.uint64))
        return okg.mean()

    assert 0.2 < same_raytrace() < 1.0
    for pts_s, org in room_scans(rng, size, 3):
        both(g, o, "insert", "test", pts_s, org)
        same_all(g, o)
    same_raytrace()
