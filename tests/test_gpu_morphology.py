"""The morphology operators on the GPU (vdbgpu_morphology: VDBMapping.hpp:1095-1148,
dilate / erode / close / open with 26-neighbour connectivity) against the
oracle, bit for bit, and the smoothed section applies built on them
(applyMapSectionGrid / UpdateGrid with smooth_map, VDBMapping.hpp:1022-1085)."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from morph_cases import OPS, random_leafset
from parity import both, pair, same_map
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def same_morph(g, op, it, o, m):
    print(f"morphology {op} x{it} on {len(o)} leaves", flush=True)
    go, gm = g.morphology(op, it, o, m)
    oo, om = O.morphology(op, it, o, m)
    assert np.array_equal(go, oo), (op, it)
    assert np.array_equal(gm, om), (op, it)
    return go, gm


@pytest.fixture(scope="module")
def gmap():
    return V.GpuMap(0.1)


@pytest.mark.parametrize("op", OPS)
@pytest.mark.parametrize("it", [1, 2, 3, 8, 9, 17])
def test_matches_oracle(gmap, op, it):
    for seed in (0, 1):
        o, m = random_leafset(seed, span=3, lo=-8)
        same_morph(gmap, op, it, o, m)


def test_random_voxels_sparse_and_far_apart(gmap):
    o, m = random_leafset(5, span=4, keep=0.3, density=0.05, blobs=False, lo=-1000)
    for op in OPS:
        same_morph(gmap, op, 2, o, m)
    # isolated leaves far apart (no shared neighbours), negative and positive
    o = np.array([[-4096, 64, 8], [4096, -64, -8], [0, 0, 0]], np.int32)
    m = np.zeros((3, 8), np.uint64)
    m[0, 0], m[1, 7], m[2, 3] = 1, 1 << 63, 0x00FF00FF00FF00FF
    for op in OPS:
        same_morph(gmap, op, 3, o, m)


def test_closed_forms(gmap):
    o = np.array([[0, 0, 0]], np.int32)
    m = np.zeros((1, 8), np.uint64)
    m[0, 0] = 1
    go, gm = same_morph(gmap, "dilate", 2, o, m)
    assert sum(bin(int(w)).count("1") for w in gm.ravel()) == 125
    full = np.full((1, 8), np.uint64(0xFFFFFFFFFFFFFFFF))
    go, gm = same_morph(gmap, "erode", 1, o, full)
    assert sum(bin(int(w)).count("1") for w in gm.ravel()) == 216
    go, gm = same_morph(gmap, "erode", 4, o, full)
    assert len(go) == 1 and not gm.any()


def test_zero_iterations_empty_and_errors(gmap):
    o, m = random_leafset(7, span=2, lo=0)
    for op in OPS:
        go, gm = gmap.morphology(op, 0, o, m)
        assert np.array_equal(go, o) and np.array_equal(gm, m)
    go, gm = gmap.morphology("close", 2, np.zeros((0, 3), np.int32), np.zeros((0, 8), np.uint64))
    assert len(go) == 0
    z = np.zeros((2, 8), np.uint64)
    for bad in ([[0, 0, 0], [0, 0, 0]], [[0, 0, 4], [8, 0, 0]]):
        with pytest.raises(V.VdbGpuError):
            gmap.morphology("dilate", 1, np.array(bad, np.int32), z)
    with pytest.raises(V.VdbGpuError):
        gmap.morphology("dilate", 1, np.array([[(1 << 23) - 8, 0, 0]], np.int32), z[:1])


def test_large_set(gmap):
    """~5,000 leaves, 3 steps each way: several thousand leaves created."""
    o, m = random_leafset(11, span=20, keep=0.6, lo=-80)
    for op in ("close", "open"):
        same_morph(gmap, op, 3, o, m)


def scanned(seed, res=0.1, R=5.0):
    g, o = pair(res, max_range=R, **W.PROBS)
    rng = np.random.default_rng(seed)
    for _ in range(3):
        orig = rng.uniform(-0.5, 0.5, 3)
        pts = (orig + rng.normal(0, 1.2, (3000, 3))).astype(np.float32)
        both(g, o, "insert", "test", pts, orig)
    return g, o


BOXES = [np.array(b, np.int32) for b in ([-20, -20, -20, 20, 20, 20], [-7, 3, -30, 9, 17, 1])]


@pytest.mark.parametrize("it", [1, 2, 3])
def test_smoothed_section_apply(it):
    """applyMapSectionGrid(section, smooth_map=True, it): the section closed by
    `it` steps (created leaves hold value 0), then applied."""
    src_g, src_o = scanned(8)
    dst_g, dst_o = scanned(13)
    for bbox in BOXES:
        for full in (True, False):
            so, sv, sa = src_o.export_section(bbox, full)
            co, ca = same_morph(dst_g, "close", it, so, sa)
            cv = np.zeros((len(co), 512), np.float32)
            cv[:len(so)] = sv
            both(dst_g, dst_o, "apply_section", co, cv, ca)
            same_map(dst_g, dst_o)


@pytest.mark.parametrize("it", [1, 2])
def test_smoothed_section_update_apply(it):
    dst_g, dst_o = scanned(5)
    src_o = scanned(21)[1]
    for bbox in BOXES:
        so, _, sa = src_o.export_section(bbox, False)
        co, ca = same_morph(dst_g, "close", it, so, sa)
        both(dst_g, dst_o, "apply_section_update", bbox, co, ca)
        same_map(dst_g, dst_o)
