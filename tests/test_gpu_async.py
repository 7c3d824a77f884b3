"""The asynchronous front end (vdbgpu_async_*: addDataToAccumulate with the
accumulation and integration threads, VDBMapping.hpp:355-370, :1383-1430):
the latest-only slot drops a cloud its worker has not taken, clouds
accumulated within a period OR-merge, and a flush integrates; every outcome
is compared with the oracle's sequence of accumulates and integrates."""
import time

import numpy as np
import pytest

import vdb_mapping_amd as V
from parity import pair, same_map

pytestmark = pytest.mark.gpu


def clouds(seed, k, n=3000):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(k):
        org = rng.uniform(-1, 1, 3)
        out.append(((org + rng.normal(0, 2.0, (n, 3))).astype(np.float32), org))
    return out


def wait_for(g, key, value, timeout=60.0):
    t0 = time.time()
    while g.async_counters()[key] < value:
        assert time.time() - t0 < timeout, g.async_counters()
        time.sleep(0.001)


def test_one_cloud_equals_insert():
    g, o = pair(0.1, max_range=6.0)
    (pts, org), = clouds(1, 1)
    assert g.async_start(0.05) == V.OK
    assert g.add_data_to_accumulate("test", pts, org) == V.OK
    assert g.async_flush() == V.OK
    o.insert("test", pts, org)
    same_map(g, o)
    c = g.async_counters()
    assert c["submitted"] == 1 and c["accumulated"] == 1 and c["integrations"] >= 1
    assert g.async_stop() == V.OK


def test_latest_only_slot_and_or_merge():
    """Worker period 2 s (max_rate 0.5), integration period 60 s: cloud A is
    accumulated at once; B and C arrive while the worker waits, C replaces B;
    the worker then accumulates C into the same update set; the flush
    integrates A | C once."""
    g, o = pair(0.1, max_range=6.0)
    (a, oa), (b, ob), (c, oc) = clouds(2, 3)
    assert g.async_start(60.0) == V.OK
    wait_for(g, "integrations", 1)  # the thread's first (empty) integrate
    assert g.add_data_to_accumulate("test", a, oa, max_rate=0.5) == V.OK
    wait_for(g, "accumulated", 1)
    assert g.add_data_to_accumulate("test", b, ob, max_rate=0.5) == V.OK
    assert g.add_data_to_accumulate("test", c, oc, max_rate=0.5) == V.OK
    assert g.async_flush() == V.OK
    cnt = g.async_counters()
    assert cnt["submitted"] == 3 and cnt["replaced"] == 1 and cnt["accumulated"] == 2
    o.accumulate("test", a, oa)
    o.accumulate("test", c, oc)
    o.integrate()
    same_map(g, o)
    assert g.async_stop() == V.OK


def test_two_sources_and_unknown_source():
    g, o = pair(0.1, max_range=6.0, sources=("a", "b"))
    (p1, o1), (p2, o2) = clouds(3, 2)
    assert g.async_start(60.0) == V.OK
    assert g.add_data_to_accumulate("nope", p1, o1) == V.UNKNOWN_SOURCE
    assert g.add_data_to_accumulate("b", p1, o1) == V.OK
    assert g.add_data_to_accumulate("a", p2, o2) == V.OK
    assert g.async_flush() == V.OK
    o.accumulate("b", p1, o1)
    o.accumulate("a", p2, o2)
    o.integrate()  # sources are integrated in id order either way
    same_map(g, o)
    g.async_stop()


def test_destroy_stops_the_threads():
    g = V.GpuMap(0.1)
    g.set_config(V.make_config(max_range=5.0))
    g.add_input_source("s", 5.0)
    assert g.async_start(0.01) == V.OK
    for pts, org in clouds(4, 5, n=500):
        assert g.add_data_to_accumulate("s", pts, org) == V.OK
    time.sleep(0.05)
    g.close()  # joins the worker and integration threads


def test_has_source_is_a_pure_lookup():
    """The source check of addDataToAccumulate (VDBMapping.hpp:359-364) neither
    applies waiting scans nor touches the update sets."""
    g, o = pair(0.1, max_range=6.0)
    assert g.has_source("test") and not g.has_source("nope")
    (pts, org), = clouds(7, 1)
    assert g.accumulate("test", pts, org) == V.OK
    before = g.export_update("test")
    assert g.has_source("test")
    after = g.export_update("test")
    assert all(np.array_equal(a, b) for a, b in zip(before, after))


def test_stopped_engine_takes_no_more_clouds():
    g, o = pair(0.1, max_range=6.0)
    (pts, org), (pts2, org2) = clouds(8, 2)
    assert g.async_start(0.05) == V.OK
    assert g.add_data_to_accumulate("test", pts, org) == V.OK
    assert g.async_flush() == V.OK
    assert g.async_stop() == V.OK
    assert g.add_data_to_accumulate("test", pts2, org2) == V.NOT_CONFIGURED
    assert g.async_flush() == V.OK  # no engine: nothing to wait for
    o.insert("test", pts, org)
    same_map(g, o)
