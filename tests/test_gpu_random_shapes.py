"""Seeded random sweep of the C API's input space against the oracle
(plain and Inc solves, short and long systems, both mixed-precision solves,
ragged batches).

Each case draws ndim 1-3, extents 1..70 (one dimension sometimes up to 600),
pads of up to 3 extra elements per dimension, a base offset into the
allocation, the solve direction, FP32 or FP64, plain or Inc, and the
algorithm hint (auto, per-lane, hybrid). The GPU result must match the
oracle's solve of the same arrays:
  normwise  <= 1e-12 (FP64), 1e-5 (FP32), and the reference tests'
  elementwise formula (tests/helpers.py);
  elements outside the solved region, and d for Inc, stay untouched.
"""
import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import api  # noqa: E402

CASES = 60
HINTS = [(0, 0, 0), (1, 1, 1), (3, 3, 3)]


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def draw(rng):
    ndim = int(rng.integers(1, 4))
    dims = [int(rng.integers(1, 71)) for _ in range(ndim)]
    if rng.random() < 0.3:
        dims[int(rng.integers(0, ndim))] = int(rng.integers(100, 601))
    pads = [d + int(rng.integers(0, 4)) for d in dims]
    solvedim = int(rng.integers(0, ndim))
    dtype = np.float64 if rng.random() < 0.6 else np.float32
    inc = bool(rng.random() < 0.4)
    hint = HINTS[int(rng.integers(0, len(HINTS)))]
    offset = int(rng.integers(0, 5))
    return ndim, dims, pads, solvedim, dtype, inc, hint, offset


@pytest.mark.parametrize("case", range(CASES))
def test_random_shape(case):
    rng = np.random.default_rng(1000 + case)
    ndim, dims, pads, sd, dtype, inc, hint, offset = draw(rng)
    total = offset + int(np.prod(pads)) + 3  # a tail past the last element
    a, b, c, d = H.diag_dominant(rng, total, dtype)
    u = rng.uniform(-1, 1, total).astype(dtype) if inc else None
    ta, tb, tc, td = (torch.from_numpy(t.copy()).cuda() for t in (a, b, c, d))
    tu = torch.from_numpy(u.copy()).cuda() if inc else None
    api.trid_strided_batch(ta, tb, tc, td, solvedim=sd, dims=dims, pads=pads, u=tu,
                           offset=offset, opts=hint, sync=True)
    od, ou = d.copy(), (u.copy() if inc else None)
    oracle.batch(a, b, c, od, ou, ndim=ndim, solvedim=sd, dims=dims, pads=pads, inc=inc,
                 offset=offset)
    got_d = td.cpu().numpy()
    what = "case %d: ndim %d dims %s pads %s sd %d %s inc %s hint %s offset %d config %s" % (
        case, ndim, dims, pads, sd, np.dtype(dtype).name, inc, hint, offset, api.last_config())
    if inc:
        got_u = tu.cpu().numpy()
        assert np.array_equal(got_d, d), what + ": Inc wrote d"
        H.assert_norm_close(got_u - u, ou - u, dtype)
        H.ref_allclose(ou, got_u, dtype)
    else:
        H.assert_norm_close(got_d, od, dtype)
        H.ref_allclose(od, got_d, dtype)
        # outside the solved region nothing changes (the oracle leaves it as is)
        solved = np.zeros(total, dtype=bool)
        for idx in H.system_slices(ndim, sd, dims, pads, offset):
            solved[idx] = True
        assert np.array_equal(got_d[~solved], d[~solved]), what + ": wrote outside the systems"


def draw_long(rng):
    """A random layout whose solve dimension is long: 4097..40000 rows
    (log-uniform), the other extents 1..4, so that contiguous systems take
    the one-workgroup solves of up to 16384 rows or the long-system path, and
    strided ones the long-system path."""
    ndim = int(rng.integers(1, 4))
    sd = int(rng.integers(0, ndim))
    dims = [int(rng.integers(1, 5)) for _ in range(ndim)]
    dims[sd] = int(np.round(np.exp(rng.uniform(np.log(4097), np.log(40000)))))
    pads = [d + int(rng.integers(0, 4)) for d in dims]
    dtype = np.float64 if rng.random() < 0.6 else np.float32
    inc = bool(rng.random() < 0.4)
    offset = int(rng.integers(0, 5))
    return ndim, dims, pads, sd, dtype, inc, offset


@pytest.mark.parametrize("case", range(24))
def test_random_long_shape(case):
    """Plain and Inc solves of long systems on random layouts: as
    test_random_shape, against the oracle; the configuration the call ran is
    named in the failure message."""
    rng = np.random.default_rng(5000 + case)
    ndim, dims, pads, sd, dtype, inc, offset = draw_long(rng)
    total = offset + int(np.prod(pads)) + 3
    a, b, c, d = H.diag_dominant(rng, total, dtype)
    u = rng.uniform(-1, 1, total).astype(dtype) if inc else None
    ta, tb, tc, td = (torch.from_numpy(t.copy()).cuda() for t in (a, b, c, d))
    tu = torch.from_numpy(u.copy()).cuda() if inc else None
    api.trid_strided_batch(ta, tb, tc, td, solvedim=sd, dims=dims, pads=pads, u=tu,
                           offset=offset, sync=True)
    od, ou = d.copy(), (u.copy() if inc else None)
    oracle.batch(a, b, c, od, ou, ndim=ndim, solvedim=sd, dims=dims, pads=pads, inc=inc,
                 offset=offset)
    what = "case %d: dims %s pads %s sd %d %s inc %s offset %d config %s" % (
        case, dims, pads, sd, np.dtype(dtype).name, inc, offset, api.last_config())
    got_d = td.cpu().numpy()
    if inc:
        assert np.array_equal(got_d, d), what + ": Inc wrote d"
        H.assert_norm_close(tu.cpu().numpy() - u, ou - u, dtype)
    else:
        H.assert_norm_close(got_d, od, dtype)
        solved = np.zeros(total, dtype=bool)
        for idx in H.system_slices(ndim, sd, dims, pads, offset):
            solved[idx] = True
        assert np.array_equal(got_d[~solved], d[~solved]), what + ": wrote outside the systems"


@pytest.mark.parametrize("case", range(30))
def test_random_shape_mixed(case):
    """FP64 by FP32 sweeps + 2 FP64 refinements on random shapes (default and
    tuned-candidate kernels): <= 1e-12 of the long-double oracle."""
    rng = np.random.default_rng(2000 + case)
    ndim, dims, pads, sd, _, _, _, offset = draw(rng)
    total = offset + int(np.prod(pads)) + 3
    a, b, c, d = H.diag_dominant(rng, total)
    td = torch.from_numpy(d.copy()).cuda()
    api.mixed_strided_batch(*(torch.from_numpy(t.copy()).cuda() for t in (a, b, c)), td,
                            solvedim=sd, dims=dims, pads=pads, offset=offset, refinements=2,
                            sync=True)
    xl = d.copy()
    sl = list(H.system_slices(ndim, sd, dims, pads, offset))
    ref = oracle.batch_ld(a, b, c, d, ndim=ndim, solvedim=sd, dims=dims, pads=pads) \
        if offset == 0 else None
    got = td.cpu().numpy()
    if ref is None:  # the long-double oracle takes unshifted arrays: shift the view
        sh = lambda x: np.ascontiguousarray(x[offset:offset + int(np.prod(pads))])
        ref = oracle.batch_ld(sh(a), sh(b), sh(c), sh(d), ndim=ndim, solvedim=sd, dims=dims,
                              pads=pads)
        xl[offset:offset + int(np.prod(pads))] = ref
    else:
        xl = ref
    idx = np.concatenate(sl)
    H.assert_norm_close(got[idx], xl[idx], np.float64, tol=1e-12)
    mask = np.ones(total, dtype=bool)
    mask[idx] = False
    assert np.array_equal(got[mask], d[mask])


@pytest.mark.parametrize("case", range(30))
def test_random_shape_mixed_f32(case):
    """FP64 solution of FP32-stored systems on random shapes, pads and
    offsets: <= 1e-12 of the long-double oracle's solution of the FP32 values
    promoted; x keeps what it held outside the systems, and d is not
    written."""
    rng = np.random.default_rng(4000 + case)
    ndim, dims, pads, sd, _, _, _, offset = draw(rng)
    total = offset + int(np.prod(pads)) + 3
    a, b, c, d = H.diag_dominant(rng, total, np.float32)
    t = [torch.from_numpy(v.copy()).cuda() for v in (a, b, c, d)]
    x0 = torch.full((total,), -5.0, dtype=torch.float64, device="cuda")
    x = api.mixed_f32_strided_batch(*t, x0, solvedim=sd, dims=dims, pads=pads, offset=offset,
                                    refinements=1, sync=True).cpu().numpy()
    sh = lambda v: np.ascontiguousarray(v[offset:offset + int(np.prod(pads))].astype(np.float64))
    ref = oracle.batch_ld(sh(a), sh(b), sh(c), sh(d), ndim=ndim, solvedim=sd, dims=dims,
                          pads=pads)
    want = np.full(total, -5.0)
    want[offset:offset + int(np.prod(pads))] = ref
    idx = np.concatenate(list(H.system_slices(ndim, sd, dims, pads, offset)))
    what = "case %d: ndim %d dims %s pads %s sd %d offset %d config %s" % (
        case, ndim, dims, pads, sd, offset, api.last_config())
    H.assert_norm_close(x[idx], want[idx], np.float64, tol=1e-12)
    mask = np.ones(total, dtype=bool)
    mask[idx] = False
    assert np.all(x[mask] == -5.0), what + ": wrote outside the systems"
    assert np.array_equal(t[3].cpu().numpy(), d), what + ": wrote d"


@pytest.mark.parametrize("case", range(30))
def test_random_ragged(case):
    """Random ragged batches: 1..300 systems of 0..8192 rows (log-uniform, with
    empty ones), FP32/FP64, plain/Inc, default or forced rows per lane in one,
    two or three solve launches, and the mixed-precision ragged solves
    (FP64-stored and FP32-stored)."""
    rng = np.random.default_rng(3000 + case)
    S = int(rng.integers(1, 301))
    lens = np.round(np.exp(rng.uniform(0, np.log(8192), S))).astype(np.int64)
    lens[rng.random(S) < 0.05] = 0
    dtype = np.float64 if rng.random() < 0.6 else np.float32
    inc = bool(rng.random() < 0.3)
    M = [0, 4, 8, 16][int(rng.integers(0, 4))]
    launches = int(rng.integers(1, 3))
    if M == 16 and launches == 2 and rng.random() < 0.5:
        launches = 3  # long systems at 32 rows per lane
    max_len = int(lens.max())
    limit = {4: 4096, 8: 8192, 16: 8192}.get(M, 8192)
    if M == 4:  # 4 rows per lane hold up to 4096 rows (1024 lanes)
        lens = np.minimum(lens, limit)
        max_len = int(lens.max())
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    n = int(off[-1]) + 1
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    u = rng.uniform(-1, 1, n).astype(dtype) if inc else None
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    td, tu = dev(d), (dev(u) if inc else None)
    ctx = api.forced_ragged(M, launches=launches) if M else None
    if ctx:
        ctx.__enter__()
    try:
        api.ragged_batch(dev(a), dev(b), dev(c), td, dev(off), max_len=max(max_len, 1), u=tu,
                         sync=True)
    finally:
        if ctx:
            ctx.__exit__(None, None, None)
    want, wu = d.copy(), (u.copy() if inc else None)
    oracle.ragged(a, b, c, want, off, wu, inc=inc)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    if inc:
        assert np.array_equal(td.cpu().numpy(), d)
        H.assert_norm_close(tu.cpu().numpy() - u, wu - u, dtype, tol=tol)
    else:
        H.assert_norm_close(td.cpu().numpy(), want, dtype, tol=tol)
    if dtype == np.float64 and not inc and max_len <= 8192:
        tm = dev(d)
        api.ragged_batch_mixed(dev(a), dev(b), dev(c), tm, dev(off), max_len=max(max_len, 1),
                               refinements=2, sync=True)
        H.assert_norm_close(tm.cpu().numpy(), want, dtype, tol=1e-12)
    if dtype == np.float32 and not inc and max_len <= 8192:
        # FP64 solution of the FP32-stored batch against the oracle's FP64 solve
        # of the promoted values
        w64 = d.astype(np.float64)
        oracle.ragged(*(v.astype(np.float64) for v in (a, b, c)), w64, off)
        x = api.ragged_batch_mixed_f32(dev(a), dev(b), dev(c), dev(d), dev(off),
                                       max_len=max(max_len, 1), refinements=1, sync=True)
        n = int(off[-1])
        H.assert_norm_close(x.cpu().numpy()[:n], w64[:n], np.float64, tol=1e-12)
