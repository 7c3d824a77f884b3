"""The .vdb / byte-array archive code of the drop-in headers
(include/vdb_mapping/vdb_io.hpp: the layout of saveMap / loadMap,
VDBMapping.hpp:193-211, :263-284, and of gridToByteArray / byteArrayToGrid,
:1310-1338) on the host: float and bool grids written as file and stream
archives with active-mask, zlib and no compression read back identical, every
inactive-value encoding of a leaf included; the 8^3 tiles of a pruned map
(createMapFromPointCloud's pruneGrid, OccupancyVDBMapping.hpp:173) read back as
leaves of their value and state; truncated and foreign input is refused. Compatibility with OpenVDB itself is not checked here (no OpenVDB)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "vdb_io_host")


def test_archive_round_trips():
    if not os.path.exists(BIN):
        import __graft_entry__ as G
        G.build_cpp_tests()
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout
    assert r.stdout.count("[  OK  ]") == 20
