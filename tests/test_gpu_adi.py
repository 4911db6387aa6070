"""ADI timestep on the GPU (preproc -> x -> y -> z Inc) against the oracle's
restatement of the reference's sequential ADI (apps/adi/src/adi_orig.c:147-242).

Tolerances: preproc outputs within 4 ulp-scale absolute error of the oracle
(the stencil sum may be contracted differently); u after the steps normwise
relative to the oracle <= 1e-12 (FP64) and <= 1e-5 (FP32); hipGraph replays
bit-identical to the eager steps.
"""
import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import adi  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _np(t):
    return t.cpu().numpy()


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("lam", [1.0, 0.37])
def test_preproc_matches_oracle(dtype, lam):
    nx, ny, nz = 45, 33, 37
    u = adi.initial_u(nx, ny, nz, dtype, "random", seed=3)
    ws = adi.Workspace(nx, ny, nz, dtype)
    adi.preproc(lam, u, ws)
    torch.cuda.synchronize()
    du, a, b, c = oracle.adi_preproc(lam, _np(u), nx, ny, nz)
    eps = 1e-15 if dtype == torch.float64 else 1e-6
    np.testing.assert_allclose(_np(ws.du), du, rtol=0, atol=16 * eps)
    for got, want in ((ws.a, a), (ws.b, b), (ws.c, c)):
        assert np.array_equal(_np(got), want)


def test_preproc_nine_array_form():
    """The reference's 9-coefficient-array call writes identical copies."""
    nx, ny, nz = 20, 18, 16
    dtype = torch.float64
    u = adi.initial_u(nx, ny, nz, dtype, "random", seed=4)
    ws = adi.Workspace(nx, ny, nz, dtype)
    nine = [torch.full_like(ws.du, 7.0) for _ in range(6)]
    adi.preproc(1.0, u, ws, nine=nine)
    torch.cuda.synchronize()
    for k, t in enumerate(nine):
        assert torch.equal(t, (ws.a, ws.b, ws.c)[k % 3])


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_adi_steps_match_oracle(dtype):
    nx, ny, nz, iters = 64, 48, 40, 5
    u = adi.initial_u(nx, ny, nz, dtype, "random", seed=5)
    u_cpu = _np(u).copy()
    ws = adi.Workspace(nx, ny, nz, dtype)
    adi.run(1.0, u, ws, iters, graph=False, sync=True)
    oracle.adi(1.0, u_cpu, nx, ny, nz, iters)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    H.assert_norm_close(_np(u), u_cpu, tol=tol)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_adi_graph_replay_matches_eager(dtype):
    """One captured step replayed 4 times, then again after eager work, equals
    the eager steps bit for bit."""
    nx, ny, nz = 40, 36, 32
    u0 = adi.initial_u(nx, ny, nz, dtype, "random", seed=6)
    ws1 = adi.Workspace(nx, ny, nz, dtype)
    ws2 = adi.Workspace(nx, ny, nz, dtype)
    ue = u0.clone()
    for _ in range(4):
        adi.step(1.0, ue, ws1)
    ug = u0.clone()
    adi.run(1.0, ug, ws2, 4, graph=True, sync=True)
    torch.cuda.synchronize()
    assert torch.equal(ue, ug)
    # more eager work, then another replay series on both
    ue.mul_(0.5)
    ug.mul_(0.5)
    for _ in range(3):
        adi.step(1.0, ue, ws1)
    adi.run(1.0, ug, ws2, 3, graph=True, sync=True)
    torch.cuda.synchronize()
    assert torch.equal(ue, ug)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("opts", [(0, 1, 1), (1, 2, 1)])
def test_adi_graph_run_lane_path(dtype, opts):
    """The reference app's -opty/-optz (and -optx) choices that select the
    per-lane solve, which needs a workspace, in a hipGraph-captured run: the
    library reserves the capture stream's workspace before capturing. The
    replayed steps equal the eager steps and the oracle ADI."""
    nx, ny, nz, iters = 40, 36, 32, 3
    u0 = adi.initial_u(nx, ny, nz, dtype, "random", seed=16)
    u_cpu = _np(u0).copy()
    ws1 = adi.Workspace(nx, ny, nz, dtype)
    ws2 = adi.Workspace(nx, ny, nz, dtype)
    ue = u0.clone()
    for _ in range(iters):
        adi.step(1.0, ue, ws1, opts=opts)
    ug = u0.clone()
    adi.run(1.0, ug, ws2, iters, graph=True, opts=opts, sync=True)
    torch.cuda.synchronize()
    assert torch.equal(ue, ug)
    oracle.adi(1.0, u_cpu, nx, ny, nz, iters)
    H.assert_norm_close(_np(ug), u_cpu, tol=1e-12 if dtype == torch.float64 else 1e-5)
    # a second captured run reuses nothing of the first one's stream
    adi.run(1.0, ug, ws2, 1, graph=True, opts=opts, sync=True)
    adi.step(1.0, ue, ws1, opts=opts)
    torch.cuda.synchronize()
    assert torch.equal(ue, ug)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_config_adi_256_20_steps(dtype):
    """ADI 256^3, lambda = 1, u0 ~ U(0,1) inside, Dirichlet 0, 20 steps; the
    GPU run (hipGraph replay) against the oracle ADI on the CPU."""
    n, iters = 256, 20
    u = adi.initial_u(n, n, n, dtype, "random", seed=256)
    u_cpu = _np(u).copy()
    ws = adi.Workspace(n, n, n, dtype)
    adi.run(1.0, u, ws, iters, graph=True, sync=True)
    oracle.adi(1.0, u_cpu, n, n, n, iters)
    got = _np(u)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    err = H.assert_norm_close(got, u_cpu, tol=tol)
    print("ADI 256^3 x 20 %s: normwise rel err vs oracle %.2e, max abs %.2e"
          % (dtype, err, np.abs(got.astype(np.float64) - u_cpu).max()))


# --- matrix-free step (tridAdiStepFused*, tridAdiRunFused*) -----------------
# Same sweeps with a, b, c computed from the grid position: results must equal
# the array-based step bit for bit (same kernel shapes, same arithmetic).

def _fused_step(lam, u, du, shape):
    import ctypes as C
    from tridsolver_amd import _lib
    from tridsolver_amd.api import use_torch_stream
    sfx = "D" if u.dtype == torch.float64 else "S"
    use_torch_stream()
    nx, ny, nz = shape
    _lib.check(getattr(_lib.lib(), "tridAdiStepFused" + sfx)(
        lam, C.c_void_p(u.data_ptr()), C.c_void_p(du.data_ptr()), nx, ny, nz),
        "tridAdiStepFused")


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("shape", [(64, 48, 40), (37, 29, 23), (5, 3, 4), (1, 6, 7),
                                   (130, 3, 70), (2, 2, 2)])
def test_fused_step_equals_array_step(dtype, shape):
    nx, ny, nz = shape
    u0 = adi.initial_u(nx, ny, nz, dtype, "random", seed=11)
    ws = adi.Workspace(nx, ny, nz, dtype)
    ua, uf = u0.clone(), u0.clone()
    du = torch.empty_like(u0)
    for _ in range(3):
        adi.step(0.8, ua, ws)
        _fused_step(0.8, uf, du, shape)
    torch.cuda.synchronize()
    assert torch.equal(ua, uf)
    assert torch.equal(ws.du, du)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_run_matches_oracle_and_graph(dtype):
    nx, ny, nz, iters = 64, 48, 40, 5
    u = adi.initial_u(nx, ny, nz, dtype, "random", seed=12)
    u_cpu = _np(u).copy()
    ug = u.clone()
    du = torch.empty_like(u)
    adi.run_fused(1.0, u, du, iters, (nx, ny, nz), graph=False, sync=True)
    adi.run_fused(1.0, ug, torch.empty_like(u), iters, (nx, ny, nz), graph=True, sync=True)
    oracle.adi(1.0, u_cpu, nx, ny, nz, iters)
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    H.assert_norm_close(_np(u), u_cpu, tol=tol)
    assert torch.equal(u, ug)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_fused_256_20_steps_equals_array(dtype):
    """ADI 256^3 x 20 steps: matrix-free graph run equals the array-based
    graph run bit for bit (the array run is checked against the oracle by
    test_config_adi_256_20_steps)."""
    n, iters = 256, 20
    u = adi.initial_u(n, n, n, dtype, "random", seed=256)
    uf = u.clone()
    ws = adi.Workspace(n, n, n, dtype)
    adi.run(1.0, u, ws, iters, graph=True, sync=True)
    du = ws.du
    del ws
    adi.run_fused(1.0, uf, du, iters, (n, n, n), graph=True, sync=True)
    torch.cuda.synchronize()
    assert torch.equal(u, uf)


def test_fused_rejects_bad_arguments():
    from tridsolver_amd import _lib
    L = _lib.lib()
    assert L.tridAdiStepFusedD(1.0, None, None, 4, 4, 4) == 3
    u = torch.zeros(8, dtype=torch.float64, device="cuda")
    import ctypes as C
    p = C.c_void_p(u.data_ptr())
    assert L.tridAdiStepFusedD(1.0, p, p, 0, 2, 4) == 3
    # a partial coefficient set stays an error on the array API
    assert L.tridAdiRunD(None, 1.0, p, p, None, p, p, 2, 2, 2, 1, 0) == 3


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("prec", [64, 32])
def test_cli_dump_matches_oracle(tmp_path, fused, prec):
    """The ADI command line with the reference's start (1 on the boundary),
    its raw dump read back and compared with the oracle ADI."""
    from tridsolver_amd import adi_io
    nx, ny, nz, iters = 30, 26, 22, 4
    path = tmp_path / "u.dat"
    adi.main(["-nx", str(nx), "-ny", str(ny), "-nz", str(nz), "-iter", str(iters),
              "-prec", str(prec), "-prof", "0", "-fused", str(fused), "-dump", str(path)])
    dt = np.float64 if prec == 64 else np.float32
    u = np.zeros((nz, ny, nx), dt)
    u[0], u[-1] = 1, 1
    u[:, 0], u[:, -1] = 1, 1
    u[:, :, 0], u[:, :, -1] = 1, 1
    u = u.reshape(-1)
    oracle.adi(1.0, u, nx, ny, nz, iters)
    ref = tmp_path / "ref.dat"
    adi_io.dump_u(ref, u, nx, ny, nz)
    r = adi_io.compare(ref, path, nx, ny, nz, prec)
    assert r["nan_points"] == 0
    assert r["normwise_rel"] <= (1e-12 if prec == 64 else 1e-5)
