"""Helpers shared by the GPU parity tests: run the same calls on the GPU map
and on the oracle, compare exports bit for bit."""
import numpy as np

import oracle as O
import vdb_mapping_amd as V


def pair(res, max_range=10.0, sources=("test",), source_ranges=None, **probs):
    g = V.GpuMap(res)
    o = O.OracleMap(res)
    cg = V.make_config(max_range=max_range, **probs)
    co = O.make_config(max_range=max_range, **probs)
    assert g.set_config(cg) == o.set_config(co)
    for i, s in enumerate(sources):
        r = max_range if source_ranges is None else source_ranges[i]
        assert g.add_input_source(s, r) == o.add_input_source(s, r)
    return g, o


def same_update(g, o, src):
    go, gv, gh = g.export_update(src)
    oo, ov, oh = o.export_update(src)
    assert go.shape == oo.shape, (go.shape, oo.shape)
    assert np.array_equal(go, oo)
    assert np.array_equal(gv, ov)
    assert np.array_equal(gh, oh)
    return go.shape[0]


def same_map(g, o):
    go, gv, gm = g.export_leaves()
    oo, ov, om = o.export_leaves()
    assert go.shape == oo.shape, (go.shape, oo.shape)
    assert np.array_equal(go, oo)
    assert np.array_equal(gv.view(np.uint32), ov.view(np.uint32))
    assert np.array_equal(gm, om)
    return go.shape[0]


def both(g, o, method, *args):
    rg = getattr(g, method)(*args)
    ro = getattr(o, method)(*args)
    assert rg == ro, (method, rg, ro, g.last_error())
    return rg
