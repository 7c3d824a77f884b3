"""Pins the oracle's morphology operators (VDBMapping.hpp:1095-1148:
morphologicalDilateMap / ErodeMap over tools::dilateActiveValues /
erodeActiveValues with NN_FACE_EDGE_VERTEX, and their Close / Open
compositions) before they are used as the checker.

Source of truth for the voxel states: scipy.ndimage.binary_dilation and
binary_erosion with a 3x3x3 structuring element on a dense array (voxels in
no leaf are inactive, the array border is far away). The leaf set the
oracle returns (input leaves kept, a leaf created where dilation activates a
voxel in no leaf) is OpenVDB's tree behaviour and stays unpinned; the tests
check it against that rule, stated on the dense result.
"""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
from morph_cases import OPS, random_leafset, to_dense, unpack

BOX = np.ones((3, 3, 3), bool)


def dense_ref(op, it, d):
    """Textbook result of op on the dense array d, and the dilated array whose
    leaves dilation creates (the last dilation stage)."""
    dil = lambda a: ndimage.binary_dilation(a, BOX, iterations=it) if it > 0 else a
    ero = lambda a: ndimage.binary_erosion(a, BOX, iterations=it, border_value=0) if it > 0 else a
    if op == "dilate":
        r = dil(d)
        return r, r
    if op == "erode":
        return ero(d), d
    if op == "close":
        g = dil(d)
        return ero(g), g
    g = dil(ero(d))
    return g, g


def leaves_of(d, lo):
    """Origins of the 8^3 leaves of dense array d (offset lo) holding a voxel."""
    n = d.shape[0] // 8
    blk = d.reshape(n, 8, n, 8, n, 8).any(axis=(1, 3, 5))
    return {tuple(int(c) for c in lo + 8 * np.array(ijk)) for ijk in np.argwhere(blk)}


def check(op, it, origins, masks, lo, size):
    oo, om = O.morphology(op, it, origins, masks)
    n = len(origins)
    # input leaves first, in input order, each still there
    assert np.array_equal(oo[:n], origins)
    # voxel states: the textbook operator on the dense array
    want, grown = dense_ref(op, it, to_dense(origins, masks, lo, size))
    assert np.array_equal(to_dense(oo, om, lo, size), want), (op, it)
    # leaf set: the new leaves are the ones the dilation stage put a voxel in
    new = {tuple(int(c) for c in o) for o in oo[n:]}
    assert new == leaves_of(grown, lo) - {tuple(int(c) for c in o) for o in origins}, (op, it)
    assert [tuple(o) for o in oo[n:]] == sorted(tuple(o) for o in oo[n:])
    return oo, om


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("it", [1, 2, 3])
@pytest.mark.parametrize("seed", [0, 1])
def test_matches_scipy(op, it, seed):
    o, m = random_leafset(seed, span=3, lo=-8)
    check(op, it, o, m, lo=-8 - 16, size=24 + 32)


@pytest.mark.parametrize("op", ["dilate", "close", "open"])
def test_more_than_a_leaf_of_steps(op):
    # 9 steps reach past the neighbouring leaf: two rings of new leaves
    o, m = random_leafset(3, span=2, keep=0.7, lo=0)
    oo, _ = check(op, 9, o, m, lo=-24, size=16 + 48)
    if op != "open":  # some new leaf lies two leaves away from every input leaf
        far = lambda q: min(int(np.abs(q - p).max()) for p in o) >= 16
        assert any(far(q) for q in oo[len(o):])


def test_random_voxels_and_sparse_leaves():
    o, m = random_leafset(5, span=3, keep=0.3, density=0.05, blobs=False, lo=-8)
    for op in OPS:
        check(op, 2, o, m, lo=-24, size=56)


def test_single_voxel_closed_forms():
    o = np.array([[0, 0, 0]], np.int32)
    m = np.zeros((1, 8), np.uint64)
    m[0, 0] = 1  # voxel (0, 0, 0)
    oo, om = O.morphology("dilate", 1, o, m)
    assert len(oo) == 8  # the 3^3 cube around a leaf corner touches 8 leaves
    assert sum(int(unpack(w).sum()) for w in om) == 27
    oo, om = O.morphology("close", 1, o, m)
    assert len(oo) == 8 and int(unpack(om[0]).sum()) == 1 and not om[1:].any()
    for op in ("erode", "open"):
        oo, om = O.morphology(op, 1, o, m)
        assert len(oo) == 1 and not om.any()
    # 2 steps: a 5^3 cube
    oo, om = O.morphology("dilate", 2, o, m)
    assert sum(int(unpack(w).sum()) for w in om) == 125


def test_full_leaf_erodes_to_its_interior():
    o = np.array([[8, -16, 0]], np.int32)
    m = np.full((1, 8), np.uint64(0xFFFFFFFFFFFFFFFF))
    oo, om = O.morphology("erode", 1, o, m)
    v = unpack(om[0])
    assert v.sum() == 6 ** 3 and v[1:7, 1:7, 1:7].all()
    oo, om = O.morphology("erode", 4, o, m)
    assert not om.any() and len(oo) == 1  # emptied, kept


def test_zero_iterations_and_empty():
    o, m = random_leafset(7, span=2, lo=0)
    for op in OPS:
        oo, om = O.morphology(op, 0, o, m)
        assert np.array_equal(oo, o) and np.array_equal(om, m)
    oo, om = O.morphology("close", 2, np.zeros((0, 3), np.int32), np.zeros((0, 8), np.uint64))
    assert len(oo) == 0


def test_bad_input():
    m = np.zeros((2, 8), np.uint64)
    with pytest.raises(ValueError):
        O.morphology("dilate", 1, np.array([[0, 0, 0], [0, 0, 0]], np.int32), m)  # repeated
    with pytest.raises(ValueError):
        O.morphology("dilate", 1, np.array([[0, 0, 4], [8, 0, 0]], np.int32), m)  # not a multiple of 8
    with pytest.raises(ValueError):
        O.morphology("dilate", 1, np.array([[(1 << 23) - 8, 0, 0]], np.int32), m[:1])  # leaves the key range
