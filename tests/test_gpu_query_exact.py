"""GPU raytrace and fast-mode casts on the maps of the exact brute force.

The maps and rays of tests/test_oracle_query_exact.py (active leaves of mixed
density, leaves with inactive voxels only inside and outside the active box,
rays entering the box through the faces next to such leaves) run through the
C ABI and must equal the oracle bit for bit; the oracle itself is checked
against the exact-rational brute force on the CPU.
"""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from parity import both, same_map, same_update
import rt_exact as R

pytestmark = pytest.mark.gpu

SEEDS = [11, 12, 13, 14, 15]
N_SOURCES = 6


def scene_pair(seed):
    add_v, rem_v = R.random_scene(seed)
    g, o = V.GpuMap(1.0), O.OracleMap(1.0)
    assert g.set_config(V.make_config(max_range=500.0, fast_mode=True)) == V.OK
    assert o.set_config(O.make_config(max_range=500.0, fast_mode=True)) == O.OK
    for k in range(N_SOURCES):
        both(g, o, "add_input_source", f"s{k}", 500.0)
    both(g, o, "add_points_to_grid", (add_v + 0.5).astype(np.float32))
    both(g, o, "remove_points_from_grid", (rem_v + 0.5).astype(np.float32))
    both(g, o, "integrate")
    same_map(g, o)
    return g, o, R.scene_sets(add_v, rem_v)


@pytest.mark.parametrize("seed", SEEDS)
def test_raytrace_exact_scenes(seed):
    g, o, (leaves, active) = scene_pair(seed)
    org, dirs, lens = R.raytrace_rays(seed, leaves, active, n=4000)
    sg, okg, eg = g.raytrace(org, dirs, lens)
    so, oko, eo = o.raytrace(org, dirs, lens)
    assert sg == so == V.OK, g.last_error()
    assert np.array_equal(okg, oko)
    assert np.array_equal(eg.view(np.uint64), eo.view(np.uint64))
    assert 0.3 < okg.mean() < 1.0


@pytest.mark.parametrize("seed", SEEDS)
def test_fast_casts_exact_scenes(seed):
    """Fast-mode accumulates of the brute force's cast segments: each source casts
    from one start voxel to many end voxels (integer world coordinates at
    resolution 1 are voxel corners, so worldToIndex gives the voxel itself)."""
    g, o, (leaves, active) = scene_pair(seed)
    S, E = R.fast_rays(seed, active, n=600)
    for k in range(N_SOURCES):
        sel = np.arange(k, len(S), N_SOURCES)
        origin = S[sel[0]].astype(np.float64)
        pts = E[sel].astype(np.float32)
        both(g, o, "accumulate", f"s{k}", pts, origin)
        assert same_update(g, o, f"s{k}") > 0
