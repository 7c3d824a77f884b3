"""Loads the committed golden fixtures (tests/golden/) and checks a map
implementation against them. Used with the oracle (drift check, CPU) and
with the GPU library (parity without running the oracle)."""
import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def known_answers():
    with open(os.path.join(GOLDEN, "known_answers.json")) as f:
        return json.load(f)


def digests():
    with open(os.path.join(GOLDEN, "digests.json")) as f:
        return json.load(f)


def cases():
    return np.load(os.path.join(GOLDEN, "oracle_cases.npz"), allow_pickle=False)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def popcount(words) -> int:
    """Set bits of an integer mask array (any width)."""
    return int(np.unpackbits(np.ascontiguousarray(words).view(np.uint8)).sum())


def voxel_of(origins, values, masks, ijk):
    """(value bits, active) of voxel ijk in a leaf export; (0, False) when its
    leaf is absent."""
    o = np.array([(c // 8) * 8 for c in ijk], np.int32)
    hit = np.nonzero((origins == o).all(axis=1))[0]
    if hit.size == 0:
        return 0, False
    k = int(hit[0])
    off = ((ijk[0] & 7) << 6) | ((ijk[1] & 7) << 3) | (ijk[2] & 7)
    bits = int(values[k].view(np.uint32)[off])
    act = bool((int(masks[k][off >> 6]) >> (off & 63)) & 1)
    return bits, act


def check_map(m, data, prefix):
    o, v, a = m.export_leaves()
    assert np.array_equal(o, data[prefix + "_origins"])
    assert np.array_equal(v.view(np.uint32), data[prefix + "_values"])
    assert np.array_equal(a, data[prefix + "_masks"])
    return o.shape[0]


EDIT_BOX = (-12, -9, -6, 15, 10, 7)
EDIT_PROBS = dict(prob_hit=0.7, prob_miss=0.4, prob_thres_min=0.12, prob_thres_max=0.97)


def edit_case():
    """Inputs of the edit fixture (res 0.1, R 5): a scan, points to add and
    remove, and two artificial polygons."""
    rng = np.random.default_rng(91)
    scan = (rng.normal(0, 1.5, (4000, 3))).astype(np.float32)
    add = rng.uniform(-3, 3, (600, 3)).astype(np.float32)
    rem = np.concatenate([scan[::7], rng.uniform(-3, 3, (200, 3)).astype(np.float32)])
    polygons = [np.array([[-1.0, -1.0, 0.0], [1.2, -1.0, 0.0], [1.2, 0.8, 0.0]]),
                np.array([[2.05, 2.0, -0.3], [-2.0, 2.55, -0.3]])]
    return scan, add, rem, polygons


def run_edit_case(make_map, make_config):
    """The edit sequence of the fixture on a map class with the reference's
    method names: scan, addPointsToGrid, removePointsFromGrid,
    addArtificialAreas, another scan."""
    scan, add, rem, polygons = edit_case()
    m = make_map(0.1)
    assert m.set_config(make_config(max_range=5.0, **EDIT_PROBS)) == 0
    assert m.add_input_source("test", 5.0) == 0
    assert m.insert("test", scan, [0.02, 0.01, -0.03]) == 0
    assert m.add_points_to_grid(add) == 0
    assert m.remove_points_from_grid(rem) == 0
    assert m.add_artificial_areas(polygons, -0.3, 0.45) == 0
    assert m.insert("test", scan[::2] * 0.8, [0.3, -0.2, 0.1]) == 0
    return m


def check_edit_case(m, data):
    check_map(m, data, "edit_map")
    o, a = m.export_artificial()
    assert np.array_equal(o, data["edit_art_origins"]) and np.array_equal(a, data["edit_art_masks"])
    for full in (0, 1):
        o, v, a = m.export_section(EDIT_BOX, bool(full))
        assert np.array_equal(o, data["edit_section%d_origins" % full])
        assert np.array_equal(v.view(np.uint32), data["edit_section%d_values" % full])
        assert np.array_equal(a, data["edit_section%d_masks" % full])


def run_known_answer(make_map, case, probs):
    """One reference known-answer case against map factory make_map(res,
    max_range, probs) -> map with source "test"."""
    m = make_map(case["res"], case["max_range"], probs)
    pts = np.array(case["points"], np.float32)
    assert m.insert("test", pts, case["origin"]) == 0
    if case.get("reset"):
        o, v, a = m.export_leaves()
        for ijk, bits, act in case["expect_before_reset"]:
            assert voxel_of(o, v, a, ijk) == (bits, act)
        m.reset()
        assert m.empty()
    o, v, a = m.export_leaves()
    for ijk, bits, act in case["expect"]:
        assert voxel_of(o, v, a, ijk) == (bits, act), (case["name"], ijk)
