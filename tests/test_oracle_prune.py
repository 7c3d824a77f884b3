"""pruneGrid at the end of createMapFromPointCloud (OccupancyVDBMapping.hpp:173)
in the oracle: closed forms on hand-built maps, and raytrace / fast-mode casts
through the pruned map against the exact-rational brute force.

The hand-built map (resolution 0.1, the reference's default probabilities):
points fill the leaf at origin (0, 0, 0) completely and add one voxel at
(23, 0, 0); with set_background the active box is [0, 23] x [0, 7] x [0, 7], so
leaf (0,0,0) holds 512 voxels at (max log-odds, on) -> an active tile, leaf
(8,0,0) holds 512 voxels at (min log-odds, off) -> an inactive tile, and leaf
(16,0,0) keeps one active voxel -> stays a leaf.
"""
import math

import numpy as np
import pytest

import oracle as O
import rt_exact as R

C_MAX = np.float32(math.log(0.99) - math.log(0.01))
C_MIN = np.float32(math.log(0.01) - math.log(0.99))
RES = 0.1


def mk(res=RES, R_=5.0, fast=False):
    m = O.OracleMap(res)
    assert m.set_config(O.make_config(max_range=R_, fast_mode=fast)) == O.OK
    assert m.add_input_source("s", R_) == O.OK
    return m


def block_points():
    g = np.stack(np.meshgrid(np.arange(8), np.arange(8), np.arange(8), indexing="ij"), -1).reshape(-1, 3)
    vox = np.concatenate([g, [[23, 0, 0]]])
    return ((vox + 0.5) * RES).astype(np.float32), vox


def test_uniform_leaves_become_tiles():
    pts, _ = block_points()
    m = mk()
    assert m.create_map_from_points(pts, True, False) == O.OK
    o, v, a = m.export_leaves()
    assert o.tolist() == [[16, 0, 0]]
    assert m.L.orc_leaf_count(m.h) == 1
    to, tv, ta = m.export_tiles()
    assert to.tolist() == [[0, 0, 0], [8, 0, 0]]
    assert tv.view(np.uint32).tolist() == np.array([C_MAX, C_MIN], np.float32).view(np.uint32).tolist()
    assert ta.tolist() == [True, False]
    # voxel reads are unchanged by pruning
    assert m.get_value([3, 4, 5]) == (C_MAX, True)
    assert m.get_value([12, 1, 1]) == (C_MIN, False)
    # without the background fill no leaf is uniform except the full one
    m2 = mk()
    assert m2.create_map_from_points(pts, False, False) == O.OK
    assert m2.export_tiles()[0].tolist() == [[0, 0, 0]]
    assert m2.export_leaves()[0].tolist() == [[16, 0, 0]]


def test_pcd_points_of_an_active_tile():
    """saveMapToPCD iterates cbeginValueOn: an active tile is one value, at its origin."""
    pts, vox = block_points()
    m = mk()
    assert m.create_map_from_points(pts, True, False) == O.OK
    got = m.export_active_points()
    want = (np.array([[0, 0, 0], [23, 0, 0]]) * RES + 0.5 * RES).astype(np.float32)
    assert np.array_equal(got, want)


def test_tiles_turn_into_leaves_only_when_a_voxel_changes():
    """A miss through the inactive tile at min log-odds is clamped back (no change:
    it stays a tile); a hit changes a voxel (a leaf again, with the tile's values
    elsewhere). The same for the active tile with hit / miss exchanged."""
    pts, _ = block_points()
    m = mk()
    assert m.create_map_from_points(pts, True, False) == O.OK
    # ray from voxel (12, 3, 3) to (14, 3, 3): misses at 12, 13, hit at 14, all in the inactive tile
    # -> the hit changes voxel 14: the tile becomes a leaf (quarter-voxel coordinates keep the
    # rounded worldToIndex away from voxel faces)
    assert m.insert("s", np.array([[1.425, 0.325, 0.325]], np.float32), [1.225, 0.325, 0.325]) == O.OK
    assert [t for t in m.export_tiles()[0].tolist()] == [[0, 0, 0]]
    o, v, a = m.export_leaves()
    assert o.tolist() == [[8, 0, 0], [16, 0, 0]]
    vals = v[0].view(np.uint32)
    off = lambda x, y, z: ((x & 7) << 6) | ((y & 7) << 3) | (z & 7)  # noqa: E731
    assert vals[off(14, 3, 3)] != C_MIN.view(np.uint32)
    assert vals[off(12, 3, 3)] == C_MIN.view(np.uint32) and vals[off(9, 0, 0)] == C_MIN.view(np.uint32)
    # a hit inside the active tile is clamped back (it stays a tile); misses through it
    # change voxels (a leaf again)
    m2 = mk()
    assert m2.create_map_from_points(pts, True, False) == O.OK
    assert m2.insert("s", np.array([[0.325, 0.325, 0.325]], np.float32), [0.325, 0.325, 0.325]) == O.OK  # hit
    assert m2.export_tiles()[0].tolist() == [[0, 0, 0], [8, 0, 0]]
    m3 = mk()
    assert m3.create_map_from_points(pts, True, False) == O.OK
    assert m3.insert("s", np.array([[0.325, 0.325, 0.725]], np.float32), [0.325, 0.325, 0.025]) == O.OK
    assert m3.export_tiles()[0].tolist() == [[8, 0, 0]]


def test_point_edits_and_sections_respect_tiles():
    pts, _ = block_points()
    m = mk()
    assert m.create_map_from_points(pts, True, False) == O.OK
    # removePointsFromGrid on the inactive tile writes (min log-odds, off): no change
    assert m.remove_points_from_grid(np.array([[1.05, 0.05, 0.05]], np.float32)) == O.OK
    assert m.export_tiles()[0].tolist() == [[0, 0, 0], [8, 0, 0]]
    # addPointsToGrid there changes the voxel: a leaf again
    assert m.add_points_to_grid(np.array([[1.05, 0.05, 0.05]], np.float32)) == O.OK
    assert m.export_tiles()[0].tolist() == [[0, 0, 0]]
    # getMapSection iterates leaves only: the active tile is not part of a section
    o, v, a = m.export_section([0, 0, 0, 7, 7, 7], True)
    assert o.shape[0] == 0
    # applyMapSectionUpdateGrid visits the active tile once, at its origin
    assert m.apply_section_update([0, 0, 0, 30, 7, 7], np.zeros((0, 3), np.int32), np.zeros((0, 8), np.uint64)) == O.OK
    assert m.export_tiles()[0].shape[0] == 0
    assert m.get_value([0, 0, 0]) == (C_MAX, False) and m.get_value([1, 0, 0]) == (C_MAX, True)


def pruned_scene(seed):
    """A room-corner-like point set at resolution 1: three walls of points with
    random holes inside a 48^3 box, created with the background fill, so the box
    interior is pruned into inactive tiles and some full wall leaves into active
    tiles. Returns the map (fast mode, intersector built) and (leaves, active)
    as the intersector sees them: leaves + active tiles, active voxels."""
    rng = np.random.default_rng(seed)
    vox = []
    for a in range(3):
        g = np.stack(np.meshgrid(np.arange(48), np.arange(48), indexing="ij"), -1).reshape(-1, 2)
        keep = rng.random(len(g)) < rng.choice([0.9, 1.0])
        g = g[keep]
        wall = np.insert(g, a, 47 if a == 0 else 0, axis=1)
        vox.append(wall)
    vox.append(rng.integers(0, 48, (40, 3)))
    vox = np.unique(np.concatenate(vox), axis=0)
    m = O.OracleMap(1.0)
    assert m.set_config(O.make_config(max_range=500.0, fast_mode=True)) == O.OK
    assert m.create_map_from_points((vox + 0.5).astype(np.float32), True, False) == O.OK
    assert m.integrate() == O.OK
    o, v, a = m.export_leaves()
    to, tv, ta = m.export_tiles()
    leaves = {tuple(int(x) for x in r) for r in o} | {tuple(int(x) for x in r) for r, on in zip(to, ta) if on}
    active = {tuple(int(x) for x in c) for c in vox}
    return m, leaves, active, len(to), int(ta.sum())


@pytest.mark.parametrize("seed", [31, 32])
def test_raytrace_and_fast_casts_through_tiles_match_brute_force(seed):
    m, leaves, active, n_tiles, n_active_tiles = pruned_scene(seed)
    assert n_tiles > 50
    rng = np.random.default_rng(seed)
    n = 400
    org = rng.uniform(-10, 58, (n, 3))
    tgt = rng.uniform(2, 46, (n, 3))
    dirs = tgt - org
    lens = np.linalg.norm(dirs, axis=1) * rng.uniform(0.6, 1.6, n)
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    st, ok, ends = m.raytrace(org, dirs, lens)
    assert st == O.OK
    clean = bad = 0
    for i in range(n):
        E, D = R.exact_ray(org[i], dirs[i], lens[i])
        try:
            want = R.raytrace(leaves, active, E, D)
        except R.Unclean:
            continue
        clean += 1
        if want is None:
            bad += int(ok[i])
        else:
            v = tuple(int(x) for x in np.round(ends[i]))
            bad += int(not ok[i] or v not in want)
    assert clean > 0.9 * n and bad == 0
    S = rng.integers(-6, 54, (300, 3)).astype(np.int32)
    Ev = rng.integers(-6, 54, (300, 3)).astype(np.int32)
    seen = 0
    for s, e in zip(S, Ev):
        try:
            want = R.fast_marks(active, s, e)
        except R.Unclean:
            continue
        ln = R.fast_clip_length(active, s, e)
        if ln is not None and 0 <= ln < 2 * R.RAY_EPS:
            continue
        seen += 1
        assert {tuple(int(c) for c in v) for v in m.cast_ray_fast(s, e)} == want
    assert seen > 50
