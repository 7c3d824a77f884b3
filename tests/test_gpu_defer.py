"""Deferred brick-path scans (tuning.defer): the map update of up to 8
consecutive plain inserts runs as one group pass that applies every scan to
each leaf in insert order. The result must equal one update per scan
(VDBMapping.hpp:399-406 and updateMap :764-783 per scan), bit for bit, against
the oracle and against the same library with defer = 1, whatever the map
sees in between: reads, edits, artificial areas, sections, config changes,
resets and hash-path inserts all apply the waiting scans first."""
import numpy as np
import pytest

import oracle as O
import vdb_mapping_amd as V
from golden_cases import digest, digests, popcount
from parity import both, pair, same_map
from vdb_mapping_amd import workloads as W

pytestmark = pytest.mark.gpu


def scans(seed, n, pts=3000, spread=1.2):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        orig = rng.uniform(-0.6, 0.6, 3)
        out.append(((orig + rng.normal(0, spread, (pts, 3))).astype(np.float32), orig))
    return out


def deferred_pair(defer, res=0.1, R=5.0, **probs):
    g, o = pair(res, max_range=R, **(probs or W.PROBS))
    assert g.set_tuning(defer=defer) == V.OK
    assert g.get_tuning()["defer"] == defer
    return g, o


@pytest.mark.parametrize("defer", [1, 2, 3, 8])
def test_stream_then_read(defer):
    g, o = deferred_pair(defer)
    for k, (p, orig) in enumerate(scans(1, 19)):
        print(f"defer {defer} scan {k}", flush=True)
        both(g, o, "insert", "test", p, orig)
    same_map(g, o)
    st = g.stats()
    assert st["scans"] == 19


@pytest.mark.parametrize("defer", [2, 8])
def test_reads_between_scans(defer):
    """Leaf exports, section extracts and active-point exports between
    partial groups."""
    g, o = deferred_pair(defer)
    ss = scans(2, 13)
    box = np.array([-6, -4, -3, 5, 6, 4], np.int32)
    for k, (p, orig) in enumerate(ss):
        both(g, o, "insert", "test", p, orig)
        if k % 5 == 4:
            same_map(g, o)
        if k % 3 == 1:
            for a, b in zip(g.export_section(box, True), o.export_section(box, True)):
                assert np.array_equal(a.view(np.uint32) if a.dtype == np.float32 else a,
                                      b.view(np.uint32) if b.dtype == np.float32 else b)
        if k == 7:
            assert np.array_equal(g.export_active_points(), o.export_active_points())
    same_map(g, o)


def test_edits_sections_and_artificial_areas_between_scans():
    g, o = deferred_pair(8)
    src_g, src_o = deferred_pair(8)
    ss = scans(3, 24)
    polys = [np.array([[-1.0, -1.0, 0.0], [1.2, -1.0, 0.0], [1.2, 0.8, 0.0]])]
    rng = np.random.default_rng(4)
    for k, (p, orig) in enumerate(ss):
        both(g, o, "insert", "test", p, orig)
        both(src_g, src_o, "insert", "test", p[::-1] * 0.9, orig)
        if k == 2:
            both(g, o, "add_points_to_grid", rng.uniform(-2, 2, (500, 3)).astype(np.float32))
        if k == 5:
            both(g, o, "add_artificial_areas", polys, -0.2, 0.4)
        if k == 9:
            both(g, o, "remove_points_from_grid", rng.uniform(-2, 2, (400, 3)).astype(np.float32))
        if k == 12:
            bbox = np.array([-12, -9, -6, 10, 8, 7], np.int32)
            so, sv, sa = src_o.export_section(bbox, True)
            both(g, o, "apply_section", so, sv, sa)
        if k == 15:
            both(g, o, "restore_map_integrity")
        if k == 18:
            bbox = np.array([-5, -5, -5, 6, 6, 6], np.int32)
            so, _, sa = src_o.export_section(bbox, False)
            both(g, o, "apply_section_update", bbox, so, sa)
    same_map(g, o)
    same_map(src_g, src_o)


def test_config_change_reset_and_hash_path_between_scans():
    g, o = deferred_pair(8)
    ss = scans(5, 20)
    for k, (p, orig) in enumerate(ss):
        both(g, o, "insert", "test", p, orig)
        if k == 4:  # new log-odds constants apply to later scans only
            cfg = dict(W.PROBS, prob_hit=0.8, prob_miss=0.3)
            assert g.set_config(V.make_config(max_range=5.0, **cfg)) == o.set_config(O.make_config(max_range=5.0, **cfg))
        if k == 8:  # an accumulate leaves a pending update set: the next inserts take the hash path
            both(g, o, "accumulate", "test", p[::2], orig)
        if k == 10:
            both(g, o, "integrate")
        if k == 13:
            same_map(g, o)
            g.reset()
            o.reset()
        if k == 16:
            assert g.set_tuning(defer=3) == V.OK  # applies the waiting scans, then groups of 3
    same_map(g, o)


def test_defer_settings_agree():
    """Room scans from moving origins: defer 1 .. 8 give the same map."""
    w = W.room_scan_workload()
    base = w.scan(0)
    maps = []
    for defer in (1, 4, 8):
        g = V.GpuMap(w.resolution)
        g.set_config(V.make_config(max_range=w.max_range, **W.PROBS))
        g.add_input_source("s", w.max_range)
        assert g.set_tuning(defer=defer) == V.OK
        for k in range(11):
            sh = np.array([0.37 * k, -0.21 * k, 0.05 * k])
            pts = base.points.copy()
            pts[:, :3] += sh.astype(np.float32)
            assert g.insert("s", pts, sh) == V.OK
        maps.append(g)
    ref = maps[0].export_leaves()
    for g in maps[1:]:
        got = g.export_leaves()
        assert np.array_equal(ref[0], got[0])
        assert np.array_equal(ref[1].view(np.uint32), got[1].view(np.uint32))
        assert np.array_equal(ref[2], got[2])


def test_bad_defer():
    g = V.GpuMap(0.1)
    assert g.set_tuning(defer=9) == V.BAD_ARGUMENT
    assert g.set_tuning(defer=-1) == V.BAD_ARGUMENT
    assert g.set_tuning(defer=0) == V.OK


@pytest.mark.slow
def test_outdoor_stream_deferred_equals_per_scan():
    """Ten 262,144-point outdoor scans: defer 8 against defer 1."""
    w = W.outdoor_stream_workload(10)
    maps = []
    for defer in (1, 8):
        g = V.GpuMap(w.resolution)
        g.set_config(V.make_config(max_range=w.max_range, **W.PROBS))
        g.add_input_source("s", w.max_range)
        assert g.set_tuning(defer=defer) == V.OK
        maps.append(g)
    for i in range(10):
        s = w.scan(i)
        for g in maps:
            assert g.insert("s", s.points, s.origin) == V.OK
    a, b = maps
    ka, va, ma = a.export_leaves()
    kb, vb, mb = b.export_leaves()
    assert ka.shape[0] > 1_000_000
    assert np.array_equal(ka, kb) and np.array_equal(va.view(np.uint32), vb.view(np.uint32))
    assert np.array_equal(ma, mb)
    assert a.stats()["voxel_updates"] == b.stats()["voxel_updates"]


def test_outdoor_stream_24_scans_matches_oracle_digest():
    """The timed configuration (outdoor stream, 262,144-point scans at 0.05 m,
    brick path, defer 8) over its first 24 scans, three full group passes: the
    map equals the oracle's after the same scans (committed digest)."""
    dg = digests()["outdoor_stream_24"]
    w = W.outdoor_stream_workload(dg["scans"])
    g = V.GpuMap(w.resolution)
    assert g.set_config(V.make_config(max_range=w.max_range, **W.PROBS)) == V.OK
    g.add_input_source("s", w.max_range)
    assert g.get_tuning()["defer"] == 8 and g.get_tuning()["traversal"] == 0
    for i in range(dg["scans"]):
        s = w.scan(i)
        assert g.insert("s", s.points, s.origin) == V.OK
    o, v, m = g.export_leaves()
    assert (o.shape[0], popcount(m)) == (dg["leaves"], dg["active_voxels"])
    assert digest(o, v.view(np.uint32), m) == dg["sha256"]


def test_plain_leaf_pool_equals_reserved_ranges(monkeypatch):
    """The leaf pool grown by copy (VDBGPU_PLAIN_POOL=1, the fallback when the
    driver refuses a mapping) and grown in place give the same map through
    several pool growths."""
    ss = scans(9, 16, pts=6000, spread=2.5)
    maps = []
    for plain in ("1", "0"):
        monkeypatch.setenv("VDBGPU_PLAIN_POOL", plain)
        g, o = deferred_pair(8, res=0.05)
        for p, orig in ss:
            both(g, o, "insert", "test", p, orig)
        same_map(g, o)
        maps.append(g)
    a, b = (m.export_leaves() for m in maps)
    assert a[0].shape[0] > 4096  # past the first pool size
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32))
