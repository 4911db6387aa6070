"""Launch grids of more than 2^32 work-items are split (no GPU needed).

HIP caps gridDim.x * blockDim.x below 2^32, while the C API accepts up to
2^31 - 1 systems; a batch of that many short systems (for example rows of
length 2 along x, 128 systems per 256-thread workgroup) needs 2^32
work-items. tests/c/grid_split.cpp drives the library's splitter with such
geometries on the host and checks that every sub-batch fits and that the
sub-batches cover every system once, in order.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def test_grid_split(tmp_path):
    exe = str(tmp_path / "grid_split")
    r = subprocess.run([HIPCC, "-std=c++17", "-O1", os.path.join(ROOT, "tests", "c", "grid_split.cpp"),
                        "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
