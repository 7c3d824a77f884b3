"""CPU-side checks of the C ABI library and the host helpers (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import numpy as np

import oracle as O
import vdb_mapping_amd as V
from vdb_mapping_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "vdbgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vdbgpu_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    decl = declared_symbols()
    assert sorted(V.ABI_SYMBOLS) == decl
    out = subprocess.run(["nm", "-D", "--defined-only", V.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (vdbgpu_\w+)", out))
    missing = [s for s in decl if s not in exported]
    assert not missing, missing
    L = V.lib()
    for s in decl:
        getattr(L, s)


def test_library_is_gfx950_code():
    out = subprocess.run(["strings", V.LIB_PATH], capture_output=True, text=True, check=True).stdout
    assert "gfx950" in out


def test_status_strings_and_version():
    assert V.lib().vdbgpu_version() >= 100
    assert V.status_string(V.OK) == "ok"
    assert V.status_string(V.OUT_OF_RANGE) == "voxel index out of range"


def test_create_without_gpu_fails_cleanly():
    if V.device_count() > 0:
        return
    h = ctypes.c_void_p()
    assert V.lib().vdbgpu_create(0.1, ctypes.byref(h)) == V.DEVICE_ERROR
    assert not h.value
    assert V.lib().vdbgpu_create(-1.0, ctypes.byref(h)) == V.BAD_ARGUMENT


def test_config_struct_layout_matches_oracle():
    assert ctypes.sizeof(V.Config) == ctypes.sizeof(O.Config) == 56
    for (a, _), (b, _) in zip(V.Config._fields_, O.Config._fields_):
        assert a == b


def test_null_handles_are_rejected():
    L = V.lib()
    assert L.vdbgpu_integrate(None) == V.BAD_ARGUMENT
    assert L.vdbgpu_leaf_count(None) == -1
    assert L.vdbgpu_last_error(None) == b""


def test_workloads_are_deterministic():
    a = W.room_scan_workload().scan(0).points
    b = W.room_scan_workload().scan(0).points
    assert a.shape == (32768, 4) and a.dtype == np.float32
    assert a.tobytes() == b.tobytes()
    r = np.linalg.norm(a[:, :3], axis=1)
    assert r.min() > 1.0 and r.max() < 15.0  # inside the 20 x 20 x 5 m room


def test_outdoor_trajectory_shape():
    w = W.outdoor_stream_workload(3)
    steps = np.linalg.norm(np.diff(w.poses[:, :2], axis=0), axis=1)
    assert np.allclose(steps, 1.0)
    yaw = np.diff(w.poses[:, 5])
    assert np.abs(yaw).max() <= np.radians(5.0) + 1e-12
    s = w.scan(0)
    assert s.points.shape == (262144, 4)
