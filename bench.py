"""Benchmark of the gfx950 occupancy-integration path.

Workload: the outdoor stream -- 100 scans of 262,144 float32 points (128 rings
x 2048 columns, elevation [-22.5, 22.5] deg) along a 1 m-step trajectory,
resolution 0.05 m, max range 50 m. One step = one insertPointCloud of one scan
(prep, traversal, update-set compaction, map insert and log-odds apply), with
the scan's points already resident in HBM. Warmup steps integrate the first
scans of the stream, the timed steps integrate the following ones into the
same map (scans are reused cyclically when warmup + steps > 100).

Prints one JSON line (see --help); `value` is rays/s of the whole job.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
PMC_SUMMARY = os.path.join(ROOT, "profiles", "traverse_outdoor_stream_pmc.json")


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=95)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--scans", type=int, default=100, help="distinct scans in the stream")
    p.add_argument("--cpu-scans", type=int, default=2, help="scans timed on the CPU oracle (0 = skip)")
    p.add_argument("--no-profiling", action="store_true", help="do not time kernels with HIP events")
    p.add_argument("--tuning", type=str, default="", help="block,steps_per_lane,waves_per_cu,load_factor")
    return p.parse_args()


def cpu_baseline(w, n_scans):
    """The CPU restatement of the reference path (oracle/), one core, on the
    first n_scans scans of the same stream: full insert (raycast + integrate)."""
    import oracle as O
    m = O.OracleMap(w.resolution)
    m.set_config(O.make_config(max_range=w.max_range, **w.probs))
    m.add_input_source("lidar", w.max_range)
    rays = 0
    t = 0.0
    for i in range(n_scans):
        s = w.scan(i)
        t0 = time.perf_counter()
        m.insert("lidar", s.points, s.origin)
        t += time.perf_counter() - t0
        rays += int((~np.isnan(s.points[:, :3]).any(1)).sum())
    return m, rays, t


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {a.gpus}); this build has no "
              "multi-GPU path: run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import vdb_mapping_amd as V
    from vdb_mapping_amd import workloads as W

    if V.device_count() < 1:
        print("no gfx950 GPU visible", file=sys.stderr)
        sys.exit(1)
    w = W.outdoor_stream_workload(a.scans)
    t_gen = time.perf_counter()
    scans = [w.scan(i) for i in range(a.scans)]
    t_gen = time.perf_counter() - t_gen
    n_pts = scans[0].points.shape[0]
    stride = scans[0].points.strides[0]
    dev = V.DeviceBuffer(sum(s.points.nbytes for s in scans))
    offs = []
    off = 0
    for s in scans:
        dev.upload(s.points, off)
        offs.append(off)
        off += s.points.nbytes
    print(f"# generated {a.scans} scans in {t_gen:.1f}s, uploaded {off / 1e6:.0f} MB", file=sys.stderr, flush=True)

    g = V.GpuMap(w.resolution)
    g.set_config(V.make_config(max_range=w.max_range, **w.probs))
    g.add_input_source("lidar", w.max_range)
    if a.tuning:
        b, spl, wpc, lf = a.tuning.split(",")
        assert g.set_tuning(int(b), int(spl), int(wpc), float(lf)) == V.OK

    def step(i):
        j = i % a.scans
        g.insert_device("lidar", dev.address + offs[j], n_pts, stride, scans[j].origin, check=True)

    for i in range(a.warmup):
        step(i)
    g.synchronize()
    g.set_profiling(not a.no_profiling)
    st0 = g.stats()
    lat = []
    g.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        ts = time.perf_counter()
        step(a.warmup + k)
        g.synchronize()
        lat.append(time.perf_counter() - ts)
    g.synchronize()
    elapsed = time.perf_counter() - t0
    st1 = g.stats()
    d = {k: st1[k] - st0[k] for k in st1 if isinstance(st1[k], (int, float))}
    rays, visits, upd = d["rays"], d["visits"], d["voxel_updates"]
    lat_ms = np.array(lat) * 1e3

    roofline = None
    if d["n_traverse"] > 0 and d["ms_traverse"] > 0:
        # traversal: 12 B of point data per ray + one 8-byte key touch per voxel visit
        alg_bytes = 12.0 * rays + 8.0 * visits
        achieved = alg_bytes / (d["ms_traverse"] * 1e-3) / 1e9
        traffic = None
        if os.path.exists(PMC_SUMMARY):
            with open(PMC_SUMMARY) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": "k_traverse", "launches": d["n_traverse"],
                    "avg_launch_ms": round(d["ms_traverse"] / d["n_traverse"], 4),
                    "alg_bytes_per_launch": round(alg_bytes / d["n_traverse"])}

    cpu = None
    checked = None
    if a.cpu_scans > 0:
        om, crays, ct = cpu_baseline(w, a.cpu_scans)
        cpu = {"value": round(crays / ct, 1), "unit": "rays/s", "cores": 1, "kind": "port",
               "sample": f"first {a.cpu_scans} of the {a.scans} outdoor scans ({crays} rays), full "
                         f"insert (raycast + integrate) on the oracle, {ct:.1f} s"}
        # the same scans on a fresh GPU map must give the oracle's map bit for bit
        gc = V.GpuMap(w.resolution)
        gc.set_config(V.make_config(max_range=w.max_range, **w.probs))
        gc.add_input_source("lidar", w.max_range)
        for i in range(a.cpu_scans):
            gc.insert_device("lidar", dev.address + offs[i], n_pts, stride, scans[i].origin, check=True)
        go, gv, gm = gc.export_leaves()
        oo, ov, om_ = om.export_leaves()
        checked = bool(go.shape == oo.shape and np.array_equal(go, oo) and
                       np.array_equal(gv.view(np.uint32), ov.view(np.uint32)) and np.array_equal(gm, om_))
        gc.close()

    out = {
        "metric": "rays/s + voxel updates/s at 0.05 m (bit-exact vs CPU); p99 scan latency",
        "value": round(rays / elapsed, 1),
        "unit": "rays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(elapsed / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "outdoor_stream_262k_0.05m", "scans": a.scans,
                   "points_per_scan": n_pts, "resolution": w.resolution, "max_range": w.max_range,
                   "input": "device-resident", "path": "plain per-scan insertPointCloud"},
        "voxel_updates_per_s": round(upd / elapsed, 1),
        "visits_per_s": round(visits / elapsed, 1),
        "p50_scan_ms": round(float(np.percentile(lat_ms, 50)), 4),
        "p99_scan_ms": round(float(np.percentile(lat_ms, 99)), 4),
        "kernel_ms_per_step": {k: round(d["ms_" + k] / max(a.steps, 1), 4)
                               for k in ("prep", "traverse", "compact", "apply")},
        "map_leaves": st1["leaf_count"],
        "device_gb": round(st1["device_bytes"] / 1e9, 2),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "bit_exact_check": checked,
    }
    print(json.dumps(out), flush=True)
    g.close()
    dev.free()


if __name__ == "__main__":
    main()
