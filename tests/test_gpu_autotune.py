"""The autotuner on a sample of the mixed-size batch workload (64 scans of
10k-300k points at 0.05 m): the setting it picks gives the same map as the
default tuning, bit for bit."""
import numpy as np
import pytest

import vdb_mapping_amd as V
from vdb_mapping_amd import workloads as W
from vdb_mapping_amd.autotune import autotune

pytestmark = pytest.mark.gpu


def test_tuned_setting_gives_the_same_map():
    w = W.mixed_batch_workload(64)
    sample = [0, 9, 21, 40]
    scans = [w.scan(i) for i in sample]
    clouds = [s.points[::4] for s in scans]
    origins = np.array([s.origin for s in scans])
    ranges = np.array([s.max_range for s in scans])
    cfg = dict(max_range=w.max_range, **w.probs)
    best, table = autotune(w.resolution, cfg, clouds, origins, ranges, batched=True,
                           grid={"block": (256,), "steps_per_lane": (32, 64), "waves_per_cu": (16,),
                                 "load_factor": (0.5, 0.7)})
    assert len(table) == 5 and best in [row[0] for row in table]
    maps = []
    for tuning in (None, best):
        g = V.GpuMap(w.resolution)
        g.set_config(V.make_config(**cfg))
        g.add_input_source("s", w.max_range)
        if tuning:
            assert g.set_tuning(**tuning) == V.OK
        assert g.insert_batch("s", clouds, origins, ranges, check=True) == V.OK
        maps.append(g.export_leaves())
    (ao, av, am), (bo, bv, bm) = maps
    assert ao.shape[0] > 1000
    assert np.array_equal(ao, bo) and np.array_equal(av.view(np.uint32), bv.view(np.uint32))
    assert np.array_equal(am, bm)
