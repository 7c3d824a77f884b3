"""The launch autotuner's selection logic with stubbed timings (CPU): every
candidate of the grid is timed once, refused settings are skipped, and the
fastest one is returned with the full table."""
import vdb_mapping_amd.autotune as A


def test_picks_fastest_and_skips_refused(monkeypatch):
    seen = []

    def fake_time(resolution, cfg_kwargs, tuning, clouds, origins, ranges, batched, repeats):
        seen.append(dict(tuning))
        if tuning.get("block") == 128 and tuning.get("waves_per_cu") == 8:
            return None  # a setting the library refuses
        if tuning == {"traversal": 0}:
            return 0.010
        # hash traversal: fastest at block 256, 32 visits per lane, 16 waves, load 0.7
        t = 0.020
        t -= 0.004 * (tuning["block"] == 256) + 0.003 * (tuning["steps_per_lane"] == 32)
        t -= 0.002 * (tuning["waves_per_cu"] == 16) + 0.002 * (tuning["load_factor"] == 0.7)
        return t

    monkeypatch.setattr(A, "_time_once", fake_time)
    best, table = A.autotune(0.05, {"max_range": 50.0}, clouds=[], origins=[])
    grid = A.DEFAULT_GRID
    n_hash = 1
    for v in grid.values():
        n_hash *= len(v)
    assert len(seen) == 1 + n_hash
    refused = sum(1 for t in seen if t.get("block") == 128 and t.get("waves_per_cu") == 8)
    assert len(table) == len(seen) - refused
    assert best == {"block": 256, "steps_per_lane": 32, "waves_per_cu": 16, "load_factor": 0.7, "traversal": 1}
    assert min(t for _, t in table) == dict((repr(k), v) for k, v in table)[repr(best)]


def test_brick_path_wins_when_fastest(monkeypatch):
    monkeypatch.setattr(A, "_time_once", lambda *a, **k: 0.001 if a[2] == {"traversal": 0} else 0.002)
    best, table = A.autotune(0.05, {"max_range": 50.0}, clouds=[], origins=[], grid={"block": (256,)})
    assert best == {"traversal": 0}
    assert all(row[0].get("block") == 256 for row in table[1:])
