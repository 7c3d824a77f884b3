"""Benchmark of the gfx950 occupancy-integration path.

Workload: the outdoor stream -- 100 scans of 262,144 float32 points (128 rings
x 2048 columns, elevation [-22.5, 22.5] deg) along a 1 m-step trajectory,
resolution 0.05 m, max range 50 m. One step = one insertPointCloud of one scan
(prep, traversal, update-set compaction, map insert and log-odds apply), with
the scan's points already resident in HBM. Warmup steps integrate the first
scans of the stream, the timed steps integrate the following ones into the
same map (scans are reused cyclically when warmup + steps > 100).

Prints one JSON line (see --help); `value` is rays/s of the whole job.

Other modes (--mode, not the default line) measure batching, queries and serving:
  batch  the outdoor stream integrated in batches of --batch scans vs the plain path
  mixed  64 scans of 10k-300k points with random poses and ranges: one batch vs per-scan
  rgbd   200 RGB-D frames at 0.02 m, then 1M raytrace queries (device arrays)
  serve  16 robot streams, Poisson arrivals at 10 Hz each, one map per stream
         (--serve-scene room: scans from a shared room pool; outdoor: each robot
         drives its own outdoor world, so its map grows by GBs)
  fast   the outdoor stream in fast mode (castRayIntoGridFast): only voxels already
         active in the map are re-marked along each ray, so the update follows the map
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


def parse():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=95)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--scans", type=int, default=100, help="distinct scans in the stream")
    p.add_argument("--cpu-scans", type=int, default=2, help="scans timed on the CPU oracle (0 = skip)")
    p.add_argument("--host-scans", type=int, default=24,
                   help="scans streamed from host memory for the PCIe-inclusive rate (0 = skip)")
    p.add_argument("--latency-scans", type=int, default=50,
                   help="scans inserted one at a time after the timed region for p50/p99 (0 = skip)")
    p.add_argument("--no-profiling", action="store_true", help="do not time kernels with HIP events")
    p.add_argument("--tuning", type=str, default="",
                   help="block,steps_per_lane,waves_per_cu,load_factor[,traversal[,defer]]")
    p.add_argument("--mode", choices=("plain", "batch", "mixed", "rgbd", "serve", "copy", "fast"),
                   default="plain")
    p.add_argument("--copy-mib", type=int, default=1024, help="copy mode: bytes per copy (MiB)")
    p.add_argument("--batch", type=int, default=8, help="scans per batch (batch mode)")
    p.add_argument("--autotune", action="store_true", help="mixed mode: autotune first")
    p.add_argument("--seconds", type=float, default=10.0, help="serve mode: arrival window")
    p.add_argument("--rate", type=float, default=10.0, help="serve mode: Poisson arrivals per stream (Hz)")
    p.add_argument("--serve-scene", choices=("room", "outdoor"), default="room",
                   help="serve mode: room pool scans, or one outdoor world per robot (large maps)")
    p.add_argument("--serve-budget-gb", type=float, default=200.0, help="serve mode: HBM budget (GB)")
    p.add_argument("--serve-check", type=int, default=-1,
                   help="serve mode: maps rebuilt alone and compared (-1: all room maps, 4 outdoor maps)")
    p.add_argument("--frames", type=int, default=200, help="rgbd mode: depth frames")
    p.add_argument("--queries", type=int, default=1_000_000, help="rgbd mode: raytrace queries")
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


def new_map(V, w, tuning=None):
    g = V.GpuMap(w.resolution)
    g.set_config(V.make_config(max_range=w.max_range, **w.probs))
    g.add_input_source("lidar", w.max_range)
    if isinstance(tuning, dict):
        assert g.set_tuning(**tuning) == V.OK
    elif tuning:
        parts = tuning.split(",")
        b, spl, wpc, lf = parts[:4]
        trav = int(parts[4]) if len(parts) > 4 else 0
        assert g.set_tuning(int(b), int(spl), int(wpc), float(lf), trav) == V.OK
    return g


def upload_scans(V, scans):
    dev = V.DeviceBuffer(sum(s.points.nbytes for s in scans))
    offs, off = [], 0
    for s in scans:
        dev.upload(s.points, off)
        offs.append(off)
        off += s.points.nbytes
    return dev, offs


def mode_batch(a, V, W):
    """Outdoor stream: plain per-scan path vs batches of a.batch scans (same scans, fresh maps)."""
    w = W.outdoor_stream_workload(a.scans)
    scans = [w.scan(i) for i in range(a.scans)]
    dev, offs = upload_scans(V, scans)
    n_pts = scans[0].points.shape[0]
    rays = sum(int((~np.isnan(s.points[:, :3]).any(1)).sum()) for s in scans)
    out = {}
    for path in ("plain", "batch", "graph", "fold"):
        # batch: the streamed path (casts launched one by one); graph: the same
        # path with its casts replayed from captured HIP graphs
        os.environ["VDBGPU_GRAPH"] = "1" if path == "graph" else "0"
        g = new_map(V, w, a.tuning)
        if path == "fold":  # the (leaf, scan)-keyed traversal and scan-ordered fold
            assert g.set_tuning(traversal=1) == V.OK
        # warm: size every table on the same stream, then start from an empty map
        for j in range(0, a.scans, a.batch):
            k = min(a.batch, a.scans - j)
            g.insert_batch_device("lidar", dev.address + offs[j], [n_pts] * k, 16,
                                  np.array([s.origin for s in scans[j:j + k]]), check=True)
        g.reset()
        g.synchronize()
        t0 = time.perf_counter()
        if path == "plain":
            for j, s in enumerate(scans):
                g.insert_device("lidar", dev.address + offs[j], n_pts, 16, s.origin, check=True)
        else:
            for j in range(0, a.scans, a.batch):
                k = min(a.batch, a.scans - j)
                g.insert_batch_device("lidar", dev.address + offs[j], [n_pts] * k, 16,
                                      np.array([s.origin for s in scans[j:j + k]]), check=True)
        g.synchronize()
        out[path] = time.perf_counter() - t0
        out[path + "_leaves"] = g.leaf_count()
        g.close()
    print(json.dumps({"metric": "batched integration of the outdoor stream", "value": round(rays / out["batch"], 1),
                      "unit": "rays/s", "plain_rays_per_s": round(rays / out["plain"], 1),
                      "fold_rays_per_s": round(rays / out["fold"], 1),
                      "graph_rays_per_s": round(rays / out["graph"], 1),
                      "speedup_vs_plain": round(out["plain"] / out["batch"], 3),
                      "graph_speedup_vs_plain": round(out["plain"] / out["graph"], 3), "batch_scans": a.batch,
                      "scans": a.scans,
                      "same_leaf_count": out["plain_leaves"] == out["batch_leaves"] == out["graph_leaves"]
                      == out["fold_leaves"],
                      "config": {"workload": "outdoor_stream_262k_0.05m", "input": "device-resident"}}),
          flush=True)


def mode_fast(a, V, W):
    """Outdoor stream in fast mode vs the normal path, each from an empty map:
    every scan inserted back to back; the first scan of a fast-mode map marks
    endpoints only (the map is empty), later scans walk the topology snapshot
    of the previous integrate."""
    w = W.outdoor_stream_workload(a.scans)
    scans = [w.scan(i) for i in range(a.scans)]
    dev, offs = upload_scans(V, scans)
    n_pts = scans[0].points.shape[0]
    rays = sum(int((~np.isnan(s.points[:, :3]).any(1)).sum()) for s in scans)
    out = {}
    for path in ("normal", "fast"):
        g = V.GpuMap(w.resolution)
        g.set_config(V.make_config(max_range=w.max_range, fast_mode=(path == "fast"), **w.probs))
        g.add_input_source("lidar", w.max_range)
        for j in range(min(5, a.scans)):  # warm: size every table, then start from an empty map
            g.insert_device("lidar", dev.address + offs[j], n_pts, 16, scans[j].origin, check=True)
        g.reset()
        g.synchronize()
        v0 = g.stats()["voxel_updates"]
        t0 = time.perf_counter()
        for j, s in enumerate(scans):
            g.insert_device("lidar", dev.address + offs[j], n_pts, 16, s.origin, check=True)
        g.synchronize()
        out[path] = time.perf_counter() - t0
        out[path + "_voxels"] = int(g.stats()["voxel_updates"] - v0)
        out[path + "_leaves"] = g.leaf_count()
        g.close()
    print(json.dumps({"metric": "fast-mode integration of the outdoor stream", "value": round(rays / out["fast"], 1),
                      "unit": "rays/s", "normal_rays_per_s": round(rays / out["normal"], 1),
                      "fast_ms_per_scan": round(out["fast"] * 1e3 / a.scans, 3),
                      "normal_ms_per_scan": round(out["normal"] * 1e3 / a.scans, 3),
                      "fast_voxel_updates": out["fast_voxels"], "normal_voxel_updates": out["normal_voxels"],
                      "fast_map_leaves": out["fast_leaves"], "normal_map_leaves": out["normal_leaves"],
                      "scans": a.scans,
                      "config": {"workload": "outdoor_stream_262k_0.05m", "input": "device-resident"}}),
          flush=True)


def mode_mixed(a, V, W):
    """64 scans, 10k-300k points, random poses, ranges U[20, 50]: one batch vs per-scan inserts."""
    w = W.mixed_batch_workload(64)
    scans = [w.scan(i) for i in range(64)]
    counts = [s.points.shape[0] for s in scans]
    dev, offs = upload_scans(V, scans)
    origins = np.array([s.origin for s in scans])
    ranges = np.array([s.max_range for s in scans])
    tuning = a.tuning or None
    tune_table = None
    if a.autotune:
        from vdb_mapping_amd.autotune import autotune
        sample = list(range(0, 64, 4))
        best, table = autotune(w.resolution, dict(max_range=w.max_range, **w.probs),
                               [scans[i].points for i in sample], origins[sample], ranges[sample],
                               batched=True)
        tuning = best
        tune_table = [(t, round(s * 1e3, 3)) for t, s in sorted(table, key=lambda r: r[1])[:5]]
    times = {}
    for path in ("plain", "batch", "graph", "fold"):
        # batch: casts launched one by one; graph: replayed from captured graphs
        os.environ["VDBGPU_GRAPH"] = "1" if path == "graph" else "0"
        g = new_map(V, w, tuning)
        if path == "fold":  # the (leaf, scan)-keyed traversal and scan-ordered fold
            assert g.set_tuning(traversal=1) == V.OK
        best = float("inf")
        for rep in range(3):
            g.reset()
            g.synchronize()
            t0 = time.perf_counter()
            if path != "plain":
                g.insert_batch_device("lidar", dev.address, counts, 16, origins, ranges, check=True)
            else:
                for j in range(64):
                    g.add_input_source("lidar", float(ranges[j]))
                    g.insert_device("lidar", dev.address + offs[j], counts[j], 16, origins[j], check=True)
            g.synchronize()
            best = min(best, time.perf_counter() - t0)
        times[path] = best
        g.close()
    rays = int(sum(counts))
    print(json.dumps({"metric": "batched integration of 64 mixed-size scans (one call)",
                      "value": round(rays / times["batch"], 1), "unit": "rays/s",
                      "plain_rays_per_s": round(rays / times["plain"], 1),
                      "fold_rays_per_s": round(rays / times["fold"], 1),
                      "graph_rays_per_s": round(rays / times["graph"], 1),
                      "speedup_vs_plain": round(times["plain"] / times["batch"], 3),
                      "graph_speedup_vs_plain": round(times["plain"] / times["graph"], 3),
                      "batch_ms": round(times["batch"] * 1e3, 3), "plain_ms": round(times["plain"] * 1e3, 3),
                      "graph_ms": round(times["graph"] * 1e3, 3),
                      "rays": rays, "tuning": tuning, "autotune_top5": tune_table,
                      "config": {"workload": "mixed_batch_64_0.05m", "input": "device-resident"}}), flush=True)


def mode_rgbd(a, V, W):
    """RGB-D frames at 0.02 m (batches of 16), then raytrace queries on device arrays."""
    w = W.rgbd_workload(a.frames)
    scans = [w.scan(i) for i in range(a.frames)]
    dev, offs = upload_scans(V, scans)
    n_pts = scans[0].points.shape[0]
    g = new_map(V, w, a.tuning)
    rays = sum(int((~np.isnan(s.points[:, :3]).any(1)).sum()) for s in scans)
    g.synchronize()
    t0 = time.perf_counter()
    for j in range(0, a.frames, 16):
        k = min(16, a.frames - j)
        g.insert_batch_device("lidar", dev.address + offs[j], [n_pts] * k, 16,
                              np.array([s.origin for s in scans[j:j + k]]), check=True)
    g.synchronize()
    t_int = time.perf_counter() - t0
    g.set_config(V.make_config(max_range=w.max_range, fast_mode=True, **w.probs))
    g.integrate(check=True)
    org, dirs, lens = W.raytrace_queries(a.queries)
    bo, bd, bl = (V.DeviceBuffer.from_array(np.ascontiguousarray(x, np.float64)) for x in (org, dirs, lens))
    bs, be = V.DeviceBuffer(a.queries), V.DeviceBuffer(a.queries * 24)
    g.raytrace_device(a.queries, bo.address, bd.address, bl.address, bs.address, be.address)  # warm
    g.set_profiling(True)
    st0 = g.stats()
    t0 = time.perf_counter()
    assert g.raytrace_device(a.queries, bo.address, bd.address, bl.address, bs.address, be.address) == V.OK
    t_rt = time.perf_counter() - t0
    st1 = g.stats()
    ok = bs.download(np.zeros(a.queries, np.uint8))
    print(json.dumps({"metric": "RGB-D integration + raytrace queries", "value": round(a.queries / t_rt, 1),
                      "unit": "queries/s", "integrate_rays_per_s": round(rays / t_int, 1),
                      "raytrace_kernel_ms": round(st1["ms_raytrace"] - st0["ms_raytrace"], 3),
                      "success_fraction": round(float(ok.mean()), 4), "frames": a.frames,
                      "map_leaves": g.leaf_count(),
                      "config": {"workload": "rgbd_640x480_0.02m", "queries": a.queries}}), flush=True)


def mode_serve(a, V, W):
    """16 streams, Poisson arrivals (10 Hz each) of 20k-262k-point scans, one map each:
    room scans drawn from a shared pool, or (--serve-scene outdoor) each robot's own
    scans in path order through its own outdoor world, so that its map grows by GBs.
    Afterwards maps are rebuilt by inserting their requests one by one on a fresh map
    (alone) and compared bit for bit with the served maps (in the mix)."""
    streams, rate = 16, a.rate
    rng = np.random.default_rng(1005)
    times = []
    for sidx in range(streams):
        t, ts = rng.exponential(1 / rate), []
        while t < a.seconds:
            ts.append(t)
            t += rng.exponential(1 / rate)
        times.append(ts)
    if a.serve_scene == "room":
        w = W.multi_robot_pool(48)
        pool = [w.scan(i) for i in range(w.n_scans)]
        scans = [[pool[sidx]] + [pool[int(rng.integers(0, len(pool)))] for _ in times[sidx]]
                 for sidx in range(streams)]
        n_check = streams if a.serve_check < 0 else a.serve_check
    else:
        robots = [W.multi_robot_outdoor(sidx, len(times[sidx]) + 1) for sidx in range(streams)]
        w = robots[0]
        scans = [[r.scan(i) for i in range(r.n_scans)] for r in robots]
        n_check = 4 if a.serve_check < 0 else a.serve_check
    print("# serve: %s scans, %d requests generated" % (a.serve_scene, sum(len(x) for x in scans)), flush=True)
    arrivals = sorted((t, sidx, j + 1) for sidx in range(streams) for j, t in enumerate(times[sidx]))
    cfg = V.make_config(max_range=w.max_range, **W.PROBS)
    budget = int(a.serve_budget_gb * 1e9)
    srv = V.Server(streams, w.resolution, cfg, w.max_range, hbm_budget_bytes=budget, max_batch=32)
    # warm every map once (table sizing), then start the clock
    req = {}  # request id -> (stream, scan)
    for sidx in range(streams):
        req[srv.submit(sidx, scans[sidx][0].points, scans[sidx][0].origin)] = (sidx, 0)
    srv.drain()
    n_warm = len(srv.records())
    peak = srv.device_bytes()
    t0 = time.perf_counter()
    for t, sidx, j in arrivals:
        delay = t0 + t - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        req[srv.submit(sidx, scans[sidx][j].points, scans[sidx][j].origin)] = (sidx, j)
        peak = max(peak, srv.device_bytes())
    srv.drain()
    elapsed = time.perf_counter() - t0
    peak = max(peak, srv.device_bytes())
    all_recs = srv.records()
    applied = [[] for _ in range(streams)]  # scans of each map whose request succeeded, in order
    for r in sorted(all_recs, key=lambda r: r["request_id"]):
        if r["status"] == 0:
            sidx, j = req[r["request_id"]]
            applied[sidx].append(j)
    status_counts = {}
    for r in all_recs:
        name = V.status_string(r["status"])
        status_counts[name] = status_counts.get(name, 0) + 1
    recs = all_recs[n_warm:]
    lat = np.array([r["latency_ms"] for r in recs])
    rays = sum(r["points"] for r in recs)
    map_gb = []
    for sidx in range(streams):
        st = srv.map(sidx).stats()
        map_gb.append((st["device_bytes"] - st["scratch_bytes"]) / 1e9)
    # alone vs in the mix: a fresh map gets the requests of the served map that
    # succeeded (a refused request leaves its map unchanged), one by one in order
    same = True
    for sidx in range(n_check):
        ref = V.GpuMap(w.resolution)
        ref.set_config(cfg)
        ref.add_input_source("s", w.max_range)
        for j in applied[sidx]:
            ref.insert("s", scans[sidx][j].points, scans[sidx][j].origin, check=True)
        a_o, a_v, a_m = srv.map(sidx).export_leaves()
        b_o, b_v, b_m = ref.export_leaves()
        same &= bool(np.array_equal(a_o, b_o) and np.array_equal(a_v.view(np.uint32), b_v.view(np.uint32))
                     and np.array_equal(a_m, b_m))
        del a_o, a_v, a_m, b_o, b_v, b_m
        ref.close()
    print(json.dumps({"metric": "multi-robot serving: p50/p99 scan latency", "value": round(float(np.percentile(lat, 99)), 3),
                      "unit": "ms (p99)", "higher_is_better": False,
                      "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                      "max_ms": round(float(lat.max()), 3), "requests": len(recs), "streams": streams,
                      "offered_hz": rate * streams, "rays_per_s": round(rays / elapsed, 1),
                      "mean_batch": round(float(np.mean([r["batch_size"] for r in recs])), 3),
                      "max_batch": int(max(r["batch_size"] for r in recs)),
                      "all_ok": all(r["status"] == 0 for r in recs), "status_counts": status_counts,
                      "mem_wait_requests": int(sum(1 for r in recs if r["mem_waits"] > 0)),
                      "device_gb": round(srv.device_bytes() / 1e9, 2), "peak_device_gb": round(peak / 1e9, 2),
                      "budget_gb": a.serve_budget_gb,
                      "map_gb": {"min": round(min(map_gb), 2), "mean": round(float(np.mean(map_gb)), 2),
                                 "max": round(max(map_gb), 2)},
                      "maps_checked": n_check, "maps_equal_sequential": same,
                      "config": {"workload": "multi_robot_16_streams_0.05m_" + a.serve_scene, "seconds": a.seconds,
                                 "rate_hz_per_stream": rate}}), flush=True)
    srv.close()


def mode_copy(a, V, W):
    """Plain device-to-device copies of --copy-mib MiB: the reference kernel for
    calibrating the FETCH_SIZE / WRITE_SIZE counters."""
    nbytes = a.copy_mib << 20
    src, dst = V.DeviceBuffer(nbytes), V.DeviceBuffer(nbytes)
    dst.copy_from(src, nbytes)  # warm
    t0 = time.perf_counter()
    for _ in range(3):
        dst.copy_from(src, nbytes)
    dt = (time.perf_counter() - t0) / 3
    print(json.dumps({"metric": "device copy", "value": round(2 * nbytes / dt / 1e9, 1), "unit": "GB/s",
                      "bytes_per_copy": nbytes, "copies": 4}), flush=True)


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
    if a.mode != "plain":
        {"batch": mode_batch, "mixed": mode_mixed, "rgbd": mode_rgbd, "serve": mode_serve,
         "copy": mode_copy, "fast": mode_fast}[a.mode](a, V, W)
        return
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
        parts = a.tuning.split(",")
        b, spl, wpc, lf = parts[:4]
        trav = int(parts[4]) if len(parts) > 4 else 0
        defer = int(parts[5]) if len(parts) > 5 else 8
        assert g.set_tuning(int(b), int(spl), int(wpc), float(lf), trav, defer) == V.OK
    defer = g.get_tuning()["defer"]

    def step(i):
        j = i % a.scans
        g.insert_device("lidar", dev.address + offs[j], n_pts, stride, scans[j].origin, check=True)

    for i in range(a.warmup):
        step(i)
    g.synchronize()
    g.set_profiling(not a.no_profiling)
    st0 = g.stats()
    g.synchronize()
    t0 = time.perf_counter()
    for k in range(a.steps):
        step(a.warmup + k)  # scans stream in: the next scan's cast overlaps this one's map update
    g.synchronize()
    elapsed = time.perf_counter() - t0
    st1 = g.stats()
    d = {k: st1[k] - st0[k] for k in st1 if isinstance(st1[k], (int, float))}
    rays, visits, upd = d["rays"], d["visits"], d["voxel_updates"]
    # scan latency, outside the timed region: one scan at a time, insert to
    # map updated (synchronous), over the next scans of the stream
    g.set_profiling(False)
    lat = []
    for k in range(min(a.steps, a.latency_scans)):
        ts = time.perf_counter()
        step(a.warmup + a.steps + k)
        g.synchronize()
        lat.append(time.perf_counter() - ts)
    lat_ms = np.array(lat) * 1e3

    # PCIe-inclusive rate: the next scans of the stream handed over from host
    # memory (pageable arrays, as insertPointCloud receives a host cloud), the
    # host-to-device copy of every cloud inside the timed region
    host = None
    if a.host_scans > 0:
        g.synchronize()
        sh0 = g.stats()
        first = a.warmup + a.steps + min(a.steps, a.latency_scans)
        th = time.perf_counter()
        for k in range(a.host_scans):
            j = (first + k) % a.scans
            g.insert("lidar", scans[j].points, scans[j].origin, check=True)
        g.synchronize()
        th = time.perf_counter() - th
        hr = g.stats()["rays"] - sh0["rays"]
        host = {"rays_per_s": round(hr / th, 1), "ms_per_scan": round(th / a.host_scans * 1e3, 4),
                "scans": a.host_scans, "bytes_per_scan": int(scans[0].points.nbytes)}

    # Roofline of the dominant kernel, with its
    # algorithmic bytes: the voxel walk moves 12 B of point data per ray and
    # touches one 8-byte key per voxel visit; the log-odds apply reads a
    # key/flag and reads and writes a float per voxel update (16 B). With
    # deferred scans the apply is one group pass per `defer` scans.
    candidates = []
    if d["n_walk"] > 0 and d["ms_walk"] > 0:
        candidates.append(("k_brick_walk", d["ms_walk"], d["n_walk"], 12.0 * rays + 8.0 * visits))
    elif d["n_traverse"] > 0 and d["ms_traverse"] > 0:
        candidates.append(("k_traverse", d["ms_traverse"], d["n_traverse"], 12.0 * rays + 8.0 * visits))
    if d["n_apply"] > 0 and d["ms_apply"] > 0:
        candidates.append(("k_apply_groups" if defer > 1 else "k_apply", d["ms_apply"], d["n_apply"], 16.0 * upd))
    rooflines = []
    for name, ms, n, alg_bytes in candidates:
        achieved = alg_bytes / (ms * 1e-3) / 1e9
        traffic = None
        pmc = os.path.join(ROOT, "profiles", f"{name}_outdoor_stream_pmc.json")
        if os.path.exists(pmc):
            with open(pmc) as f:
                traffic = json.load(f).get("hbm_bytes_per_launch")
        rooflines.append({"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": name,
                          "launches": n, "avg_launch_ms": round(ms / n, 4),
                          "alg_bytes_per_launch": round(alg_bytes / n)})
    # the dominant kernel: the longest launch (the group pass; it and the walk
    # take about the same time per scan, the walk's under contention)
    roofline = max(rooflines, key=lambda r: r["avg_launch_ms"]) if rooflines else None

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
                   "input": "device-resident", "path": "plain per-scan insertPointCloud",
                   "defer": defer},
        "voxel_updates_per_s": round(upd / elapsed, 1),
        "visits_per_s": round(visits / elapsed, 1),
        "timed_scans": a.steps,
        "host_input": host,  # PCIe-inclusive (host clouds); `value` is device-resident input
        "p50_scan_ms": round(float(np.percentile(lat_ms, 50)), 4) if len(lat_ms) else None,  # synchronous, one scan at a time
        "p99_scan_ms": round(float(np.percentile(lat_ms, 99)), 4) if len(lat_ms) else None,
        "kernel_ms_per_step": {k: round(d["ms_" + k] / max(a.steps, 1), 4)
                               for k in ("prep", "traverse", "walk", "compact", "apply")},
        "update_leaves_per_step": round(d["update_leaves"] / max(a.steps, 1), 1),
        "visits_per_step": round(visits / max(a.steps, 1), 1),
        "map_leaves": st1["leaf_count"],
        "device_gb": round(st1["device_bytes"] / 1e9, 2),
        "roofline": roofline,
        "roofline_other": [r for r in rooflines if r is not roofline],
        "cpu_baseline": cpu,
        "bit_exact_check": checked,
    }
    print(json.dumps(out), flush=True)
    g.close()
    dev.free()


if __name__ == "__main__":
    main()
