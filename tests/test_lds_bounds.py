"""Every LDS index the solve kernels form stays inside the array it addresses (CPU only).

tests/cpp/lds_bounds.cpp includes the kernels' own layout (Lds<>), LU index functions (LuIdx<>)
and network-image builders, walks the kernels' loop nests over every lane, LU block, register
slot and network entry for net14/net86/net87 in FP64 and mixed precision, and checks each index
(panel publish/step, packed rows and multipliers, the triangular solves, the refinement residual
and row 0, the FP32 factor aliasing, the image offsets, the matrix fill).  Lane-conditional loads
select their index before the load, so an index formed for an excluded lane never leaves the
array either: an out-of-range LDS address is dropped by a ds_* instruction but faults when the
same access is issued as a flat (generic-address) instruction."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_every_lds_index_in_range(tmp_path):
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path / "lds_bounds")
    subprocess.run([cxx, "-O1", "-std=c++17", "-o", exe, os.path.join(ROOT, "tests", "cpp", "lds_bounds.cpp")],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "0 out of range" in r.stdout
