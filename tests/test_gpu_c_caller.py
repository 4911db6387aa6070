"""The library called from plain C through this library's drop-in headers
(include/, which declare the reference's symbols and types; built by
tridsolver_amd/build.py):

- tests/c/drop_in.c: x, y and z solves (plain, plain, Inc) of a FP64 batch,
  an FP32 solve, a halo-padded allocation and an invalid argument;
- tests/c/fp_caller.c, built at FPPREC 0 and 1: the reference ADI app's
  style (FP arrays, F literals, the solver chosen by FPPREC), x, y and z-Inc
  sweeps of ADI coefficients.

Each program checks its own residuals (1e-13 FP64, 1e-5 FP32) and runs as a
child process."""
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


@pytest.mark.parametrize("name,size", [("fp_caller_f32", 4), ("fp_caller_f64", 8)])
def test_fp_caller(name, size):
    path = os.path.join(ROOT, "tests", "c", name)
    if not os.path.exists(path):
        pytest.fail("tests/c/%s is not built (python -m tridsolver_amd.build)" % name)
    r = subprocess.run([path], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fp_caller: OK" in r.stdout
    assert r.stdout.count("sizeof(FP) %d" % size) == 3
