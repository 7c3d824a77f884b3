"""Serving: many streams, one device map each, continuous batching per map,
HBM admission. Each map must equal the map built by inserting its requests one
by one in submission order (alone vs in the mix), bit for bit: against a GPU
map and the oracle for small random clouds, against the oracle's committed
digests at the multi-robot configuration (16 maps at 0.05 m, 20k-262k-point
room scans)."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from golden_cases import digest, digests, popcount
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


def test_server_rejects_only_when_maps_fill_the_budget():
    """A one-byte budget: the first batch runs alone (nothing holds memory yet),
    then the maps fill the budget and every later request fails with
    VDBGPU_OVER_BUDGET, whichever map it is for, without changing any map."""
    cfg = V.make_config(max_range=4.0, **W.PROBS)
    srv = V.Server(2, 0.1, cfg, 4.0, hbm_budget_bytes=1, max_batch=8)
    reqs = requests(3, 2, 6)
    srv.submit(reqs[0][0], reqs[0][1], reqs[0][2])
    assert srv.drain() == V.OK
    leaves = [srv.map(i).leaf_count() for i in range(2)]
    assert sum(leaves) > 0
    for m, pts, org in reqs[1:]:
        srv.submit(m, pts, org)
    assert srv.drain() == V.OVER_BUDGET
    recs = srv.records()
    assert len(recs) == 6 and recs[0]["status"] == V.OK
    assert all(r["status"] == V.OVER_BUDGET for r in recs[1:])
    assert [srv.map(i).leaf_count() for i in range(2)] == leaves
    srv.close()


# ---------------------------------------------------------------- multi-robot configuration --
# 16 maps at 0.05 m, room-scene scans of 20k-262k points, one burst of requests per
# map; every map must equal the oracle's map after the same scans in order
# (committed digests, tests/golden/make_golden.py).

def submit_bursts(srv, pool, order):
    for mi, k in order:
        s = pool[k]
        srv.submit(mi, s.points, s.origin)


def check_against_digests(srv, dg):
    for entry in dg["maps"]:
        o, v, m = export(srv.map(entry["map"]))
        assert (o.shape[0], popcount(m)) == (entry["leaves"], entry["active_voxels"]), entry["map"]
        assert digest(o, v, m) == entry["sha256"], entry["map"]


@pytest.fixture(scope="module")
def robot_pool():
    w = W.multi_robot_pool()
    return [w.scan(i) for i in range(w.n_scans)]


def test_server_multi_robot_matches_oracle_digests(robot_pool):
    dg = digests()["multi_robot_serve"]
    cfg = V.make_config(max_range=50.0, **W.PROBS)
    srv = V.Server(16, 0.05, cfg, 50.0, hbm_budget_bytes=0, max_batch=32)
    submit_bursts(srv, robot_pool, W.multi_robot_bursts())
    assert srv.drain() == V.OK
    recs = srv.records()
    assert len(recs) == 16 * 7 and all(r["status"] == V.OK for r in recs)
    assert max(r["batch_size"] for r in recs) > 1  # requests that queued behind a busy map ran together
    assert np.mean([r["batch_size"] for r in recs]) > 1
    check_against_digests(srv, dg)
    srv.close()


def test_server_multi_robot_waits_for_memory(robot_pool):
    """The same requests under a budget that holds the maps but only about one
    batch's working memory at a time: batches wait for memory (released by the
    lanes that finish or sit idle) instead of failing, and every map still
    equals the oracle's."""
    dg = digests()["multi_robot_serve"]
    cfg = V.make_config(max_range=50.0, **W.PROBS)
    probe = V.Server(16, 0.05, cfg, 50.0, hbm_budget_bytes=0, max_batch=32)
    submit_bursts(probe, robot_pool, W.multi_robot_bursts())
    assert probe.drain() == V.OK
    stats = [probe.map(i).stats() for i in range(16)]
    maps_bytes = sum(s["device_bytes"] - s["scratch_bytes"] for s in stats)
    scratch_max = max(s["scratch_bytes"] for s in stats)
    probe.close()
    assert scratch_max > 0
    budget = maps_bytes + 2 * scratch_max
    print(f"maps {maps_bytes / 1e6:.1f} MB, largest lane scratch {scratch_max / 1e6:.1f} MB, "
          f"budget {budget / 1e6:.1f} MB", flush=True)
    srv = V.Server(16, 0.05, cfg, 50.0, hbm_budget_bytes=budget, max_batch=32)
    submit_bursts(srv, robot_pool, W.multi_robot_bursts())
    rc = srv.drain()
    recs = srv.records()
    print("mem_waits", sum(r["mem_waits"] for r in recs), "statuses", sorted({r["status"] for r in recs}),
          "peak device bytes now", srv.device_bytes(), flush=True)
    assert rc == V.OK, sorted({r["status"] for r in recs})
    assert len(recs) == 16 * 7 and all(r["status"] == V.OK for r in recs)
    assert sum(r["mem_waits"] for r in recs) > 0  # admission held batches back
    check_against_digests(srv, dg)
    srv.close()
