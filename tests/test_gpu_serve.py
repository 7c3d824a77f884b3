"""Serving: many streams, one device map each, continuous batching per map.
Each map must equal the map built by inserting its requests one by one in
submission order (alone vs in the mix), bit for bit."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def requests(seed, n_maps, n_req, res_range=6.0):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_req):
        m = int(rng.integers(0, n_maps))
        org = rng.uniform(-2, 2, 3)
        n = int(np.exp(rng.uniform(np.log(50), np.log(4000))))
        pts = (org + rng.normal(0, res_range * 0.6, (n, 3))).astype(np.float32)
        out.append((m, np.concatenate([pts, np.ones((n, 1), np.float32)], 1), org))
    return out


def export(g):
    o, v, m = g.export_leaves()
    return o, v.view(np.uint32), m


def test_server_matches_sequential_per_map():
    n_maps = 4
    cfg = V.make_config(max_range=6.0, **W.PROBS)
    srv = V.Server(n_maps, 0.1, cfg, 6.0, hbm_budget_bytes=0, max_batch=16)
    reqs = requests(11, n_maps, 60)
    for m, pts, org in reqs:
        srv.submit(m, pts, org)
    assert srv.drain() == V.OK
    recs = srv.records()
    assert len(recs) == len(reqs) and all(r["status"] == V.OK for r in recs)
    assert max(r["batch_size"] for r in recs) > 1  # requests were batched
    for i in range(n_maps):
        ref = V.GpuMap(0.1)
        ref.set_config(cfg)
        ref.add_input_source("s", 6.0)
        orc = O.OracleMap(0.1)
        orc.set_config(O.make_config(max_range=6.0, **W.PROBS))
        orc.add_input_source("s", 6.0)
        for m, pts, org in reqs:
            if m == i:
                ref.insert("s", pts, org, check=True)
                orc.insert("s", pts, org)
        a, b, c = export(srv.map(i)), export(ref), orc.export_leaves()
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert np.array_equal(a[0], c[0]) and np.array_equal(a[1], c[1].view(np.uint32))
        assert np.array_equal(a[2], c[2])
    srv.close()


def test_server_hbm_budget_rejects_growth():
    cfg = V.make_config(max_range=4.0, **W.PROBS)
    srv = V.Server(2, 0.1, cfg, 4.0, hbm_budget_bytes=1, max_batch=8)
    reqs = requests(3, 2, 2)
    srv.submit(0, reqs[0][1], reqs[0][2])
    srv.submit(1, reqs[1][1], reqs[1][2])
    assert srv.drain() == V.OK  # maps without memory may always start
    srv.submit(0, reqs[0][1], reqs[0][2])
    assert srv.drain() == V.OVER_BUDGET
    assert srv.records()[-1]["status"] == V.OVER_BUDGET
    srv.close()
