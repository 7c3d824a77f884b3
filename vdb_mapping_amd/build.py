"""Build recipe of the native libraries (in-tree, for gfx950).

    python -m vdb_mapping_amd.build      # or __graft_entry__.build()

libvdbgpu.so : csrc/*.hip -> hipcc --offload-arch=gfx950 (the product)
libworkloads.so : workloads.c -> gcc -fopenmp (synthetic scan generator)
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")

HIP_SOURCES = ["integrate_kernels.hip", "brick_kernels.hip", "batch_kernels.hip", "raytrace_kernels.hip",
               "edit_kernels.hip", "morph_kernels.hip", "tile_kernels.hip", "vdbgpu.hip", "server.hip", "ingest.hip"]
# -ffp-contract=off: no FMA contraction, every double op rounds on its own (bit-exact DDA).
# -packed-fp32-ops off: no v_pk_*_f32 generation.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
               "-shared", "-pthread", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(ROOT, "include")]


def hipcc_cmd(out=None):
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    return ["hipcc", *HIPCC_FLAGS, *srcs, "-o", out or os.path.join(HERE, "libvdbgpu.so")]


def build_vdbgpu(verbose=False):
    out = os.path.join(HERE, "libvdbgpu.so")
    tmp = out + ".tmp"
    cmd = hipcc_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, out)


def build_workloads(verbose=False):
    from . import workloads
    if verbose:
        print(" ".join(workloads.BUILD_CMD), flush=True)
    workloads.build()


def build_all(verbose=False):
    build_vdbgpu(verbose)
    build_workloads(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
    sys.exit(0)
