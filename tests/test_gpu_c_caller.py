"""The library called from plain C through the reference's header
(tests/c/drop_in.c, built by tridsolver_amd/build.py): x, y and z solves
(plain, plain, Inc) of a FP64 batch, an FP32 solve, a halo-padded
allocation and an invalid argument, each checked by the program itself
(residuals 1e-13 FP64, 1e-5 FP32). The program runs as a child process."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "c", "drop_in")


def test_c_caller():
    if not os.path.exists(BIN):
        pytest.fail("tests/c/drop_in is not built (python -m tridsolver_amd.build)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "drop_in: OK" in r.stdout
    assert r.stdout.count("residual") == 8
