"""Autotuner: tuned shapes are used by later calls, results stay correct
(normwise <= 1e-12 FP64 vs the oracle), the cache survives save/load, and
tune-on-first-use mode tunes new geometries."""
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import _lib, api  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    yield
    _lib.lib().tridHipAutotuneClear()
    _lib.lib().tridHipSetAutotuneMode(0)


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def solve_and_check(dims, sd, dtype=np.float64, seed=0):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    td = dev(d)
    api.trid_strided_batch(dev(a), dev(b), dev(c), td, solvedim=sd, dims=dims, sync=True)
    od = d.copy()
    oracle.batch(a, b, c, od, ndim=len(dims), solvedim=sd, dims=dims)
    H.assert_norm_close(td.cpu().numpy(), od, dtype)
    return api.last_config()


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_autotune_then_solve(sd, tmp_path):
    dims = [96, 80, 72]
    _lib.lib().tridHipAutotuneClear()
    ms = api.autotune(torch.float64, sd, dims)
    assert ms > 0
    cfg = solve_and_check(dims, sd)
    path = str(tmp_path / "tune.txt").encode()
    assert _lib.lib().tridHipAutotuneSave(path) == 0
    assert _lib.lib().tridHipAutotuneClear() == 0
    assert _lib.lib().tridHipAutotuneLoad(path) == 0
    assert solve_and_check(dims, sd) == cfg
    with open(path.decode()) as f:
        assert len([ln for ln in f if not ln.startswith("#")]) >= 1


def test_autotune_on_first_use():
    _lib.lib().tridHipAutotuneClear()
    assert _lib.lib().tridHipSetAutotuneMode(1) == 0
    try:
        for sd in range(3):
            solve_and_check([50, 44, 38], sd, np.float32, seed=sd)
    finally:
        _lib.lib().tridHipSetAutotuneMode(0)


def test_autotune_mixed_f32():
    """Tuning the FP32-stored solve picks a valid shape per direction for one
    refinement and one for more (the solve then reports it), and the results
    at 1 and 2 refinements stay within 1e-12 of the long-double oracle's
    solve of the promoted values."""
    dims = [128, 64, 32]
    rng = np.random.default_rng(21)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, np.float32)
    t = [dev(v) for v in (a, b, c, d)]
    for sd in range(3):
        assert api.autotune_mixed_f32(sd, dims) > 0
        xl = oracle.batch_ld(a, b, c, d, ndim=3, solvedim=sd, dims=dims)
        for refine in (1, 2):
            x = api.mixed_f32_strided_batch(*t, solvedim=sd, dims=dims, refinements=refine,
                                            sync=True)
            assert api.last_config()[0] == 10
            assert H.norm_rel_err(x.cpu().numpy(), xl) <= 1e-12, (sd, refine)


def test_autotune_mixed_f32_save_load(tmp_path):
    """The FP32-stored solve's tuned shapes survive save, clear and load: the
    solve after loading reports the shape it reported after tuning."""
    dims = [96, 48, 40]
    rng = np.random.default_rng(22)
    a, b, c, d = H.diag_dominant(rng, int(np.prod(dims)), np.float32)
    t = [dev(v) for v in (a, b, c, d)]
    _lib.lib().tridHipAutotuneClear()
    tuned = {}
    for sd in range(3):
        assert api.autotune_mixed_f32(sd, dims) > 0
        for refine in (1, 2):
            api.mixed_f32_strided_batch(*t, solvedim=sd, dims=dims, refinements=refine,
                                        sync=True)
            tuned[sd, refine] = api.last_config()
    path = str(tmp_path / "tune_f32.txt").encode()
    assert _lib.lib().tridHipAutotuneSave(path) == 0
    assert _lib.lib().tridHipAutotuneClear() == 0
    assert _lib.lib().tridHipAutotuneLoad(path) == 0
    for sd in range(3):
        xl = oracle.batch_ld(a, b, c, d, ndim=3, solvedim=sd, dims=dims)
        for refine in (1, 2):
            x = api.mixed_f32_strided_batch(*t, solvedim=sd, dims=dims, refinements=refine,
                                            sync=True)
            assert api.last_config() == tuned[sd, refine]
            assert H.norm_rel_err(x.cpu().numpy(), xl) <= 1e-12
    _lib.lib().tridHipAutotuneClear()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_autotune_ragged_save_load(dtype, tmp_path):
    """A tuned ragged shape (rows per lane, workgroups per CU) is cached per
    precision and maximum length, and survives save, clear and load."""
    rng = np.random.default_rng(23)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(2048), 1500))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]), dtype)
    toff = dev(off)
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    want = d.copy()
    oracle.ragged(a, b, c, want, off)
    _lib.lib().tridHipAutotuneClear()
    api.autotune_ragged(tdt, toff, 2048)
    td = dev(d)
    api.ragged_batch(dev(a), dev(b), dev(c), td, toff, max_len=2048, sync=True)
    tuned = api.last_config()
    H.assert_norm_close(td.cpu().numpy(), want, dtype)
    path = str(tmp_path / "tune_ragged.txt").encode()
    assert _lib.lib().tridHipAutotuneSave(path) == 0
    assert _lib.lib().tridHipAutotuneClear() == 0
    assert _lib.lib().tridHipAutotuneLoad(path) == 0
    td = dev(d)
    api.ragged_batch(dev(a), dev(b), dev(c), td, toff, max_len=2048, sync=True)
    assert api.last_config() == tuned
    H.assert_norm_close(td.cpu().numpy(), want, dtype)
    _lib.lib().tridHipAutotuneClear()


def test_autotune_mixed_and_ragged():
    dims = [128, 64, 32]
    for sd in range(3):
        assert api.autotune_mixed(sd, dims) > 0
        rng = np.random.default_rng(sd)
        n = int(np.prod(dims))
        a, b, c, d = H.diag_dominant(rng, n)
        td = dev(d)
        api.mixed_strided_batch(dev(a), dev(b), dev(c), td, solvedim=sd, dims=dims, sync=True)
        xl = oracle.batch_ld(a, b, c, d, ndim=3, solvedim=sd, dims=dims)
        assert H.norm_rel_err(td.cpu().numpy(), xl) <= 1e-12
    rng = np.random.default_rng(12)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(4096), 2000))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]))
    toff = dev(off)
    api.autotune_ragged(torch.float64, toff, 4096)
    td = dev(d)
    api.ragged_batch(dev(a), dev(b), dev(c), td, toff, max_len=4096, sync=True)
    want = d.copy()
    oracle.ragged(a, b, c, want, off)
    H.assert_norm_close(td.cpu().numpy(), want, np.float64)


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_autotune_adjoint(sd):
    """The adjoint's tuned shapes (with and without the coefficient
    gradients) are used by later calls and keep the results of the CPU
    adjoint (normwise <= 1e-12 FP64)."""
    dims = [72, 40, 48]
    rng = np.random.default_rng(70 + sd)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n)
    w = rng.standard_normal(n)
    x = d.copy()
    oracle.batch(a, b, c, x, ndim=3, solvedim=sd, dims=dims)
    ref = oracle.adjoint(a, b, c, x, w, ndim=3, solvedim=sd, dims=dims)
    _lib.lib().tridHipAutotuneClear()
    for grads in (True, False):
        assert api.autotune_adjoint(torch.float64, sd, dims, grads=grads) > 0
        want = True if grads else None
        out = api.adjoint_strided_batch(dev(a), dev(b), dev(c), dev(x), dev(w), solvedim=sd,
                                        dims=dims, grad_a=want, grad_b=want, grad_c=want,
                                        sync=True)
        assert api.last_config()[0] == 4
        got = [t.cpu().numpy() for t in out if t is not None]
        assert len(got) == (4 if grads else 1)
        for g, r in zip(got, ref):
            H.assert_norm_close(g, r, tol=1e-12)
