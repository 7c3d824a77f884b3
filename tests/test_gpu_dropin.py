"""The drop-in C++ API (include/vdb_mapping/*.hpp) on the GPU: the reference's
five unit tests (tests/mapping.cpp) re-expressed in tests/cpp/mapping_dropin.cpp,
plus live-grid, raycastPointCloud and raytrace checks."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "mapping_dropin")


@pytest.mark.gpu
def test_cpp_dropin_api():
    assert os.path.exists(BIN), "build with __graft_entry__.build() first"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


def test_cpp_dropin_compiles_against_headers_only():
    """The headers need no HIP, PCL, Eigen or OpenVDB headers (stand-ins)."""
    src = "#include <vdb_mapping/OccupancyVDBMapping.hpp>\nint main(){return 0;}\n"
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-DVDBGPU_NO_EXTERNAL_TYPES",
                        "-I" + os.path.join(ROOT, "include"), "-x", "c++", "-"],
                       input=src, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
