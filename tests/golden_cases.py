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
    return int(sum(bin(int(x)).count("1") for x in np.asarray(words).ravel()))


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
