"""Launch-parameter autotuner of the integration kernels.

Times the brick path (traversal 0) and a sweep of the update-set hash
traversal (traversal 1: workgroup size, visits per lane work item, resident
waves per CU, hash load factor) on a sample of the caller's own scans, on
fresh maps, and returns the fastest setting (vdbgpu_set_tuning). Results never
depend on the tuning (tests/test_gpu_parity.py::test_tuning_does_not_change_results,
tests/test_gpu_brick.py), only the time does.
"""
from __future__ import annotations

import itertools
import time

import numpy as np

from . import GpuMap, make_config

DEFAULT_GRID = {
    "block": (128, 256),
    "steps_per_lane": (16, 32, 64),
    "waves_per_cu": (8, 16),
    "load_factor": (0.5, 0.7),
}


def _time_once(resolution, cfg_kwargs, tuning, clouds, origins, ranges, batched, repeats):
    best = float("inf")
    for _ in range(repeats):
        g = GpuMap(resolution)
        g.set_config(make_config(**cfg_kwargs))
        g.add_input_source("tune", cfg_kwargs["max_range"])
        if g.set_tuning(**tuning) != 0:
            g.close()
            return None
        # one untimed pass sizes the tables, the timed pass starts from an empty map
        run(g, clouds, origins, ranges, batched)
        g.reset()
        g.synchronize()
        t0 = time.perf_counter()
        run(g, clouds, origins, ranges, batched)
        g.synchronize()
        best = min(best, time.perf_counter() - t0)
        g.close()
    return best


def run(g, clouds, origins, ranges, batched):
    if batched:
        g.insert_batch("tune", clouds, origins, ranges, check=True)
    else:
        for c, o, r in zip(clouds, origins, ranges):
            g.add_input_source("tune", float(r))
            g.insert("tune", c, o, check=True)


def autotune(resolution, cfg_kwargs, clouds, origins, ranges=None, batched=True, grid=None,
             repeats=1, verbose=False):
    """Returns (best_tuning_dict, table) where table rows are (tuning, seconds)."""
    grid = dict(DEFAULT_GRID, **(grid or {}))
    ranges = np.full(len(clouds), cfg_kwargs["max_range"]) if ranges is None else np.asarray(ranges)
    keys = list(grid)
    table = []
    candidates = [dict(traversal=0)] + [dict(zip(keys, values), traversal=1)
                                        for values in itertools.product(*(grid[k] for k in keys))]
    for tuning in candidates:
        t = _time_once(resolution, cfg_kwargs, tuning, clouds, origins, ranges, batched, repeats)
        if t is None:
            continue
        table.append((tuning, t))
        if verbose:
            print(f"# autotune {tuning} {t * 1e3:.2f} ms", flush=True)
    best = min(table, key=lambda row: row[1])[0]
    return best, table
