"""Pins the CPU oracle (oracle/trid_oracle.c) before it is trusted.

Checks, per oracle function:
  * the reference-format fixtures (scipy SuperLU answers, the reference test
    generator's method) with the reference tests' tolerance formula;
  * bit-exact agreement with the reference CPU library compiled from its own
    sources (oracle/_ref, live when built; committed outputs otherwise);
  * mathematics: dense LU brute force, N=1 / N=2 closed forms, a=c=0,
    manufactured solutions, adjoint identities, finite differences.
A planted mistake (a dropped term, a transposed operand) must fail them.
"""
import json
import os

import numpy as np
import pytest

import oracle
from tests import helpers as H
from tests.golden import gen_fixtures as G

GOLD = os.path.join(os.path.dirname(__file__), "golden")
SMALL = sorted(G.COMMITTED)


@pytest.fixture(scope="module")
def big_cases(tmp_path_factory):
    out = tmp_path_factory.mktemp("fixtures")
    return G.reference_cases(str(out), G.ON_DEMAND)


def _load(name, dtype):
    return G.load_case(os.path.join(GOLD, "files", name), dtype)


def test_committed_fixtures_match_generator(tmp_path):
    with open(os.path.join(GOLD, "manifest.json")) as f:
        manifest = json.load(f)
    for name in SMALL:
        assert G.file_digest(os.path.join(GOLD, "files", name)) == manifest[name]
    # Regenerating reproduces the same bytes (numpy legacy seed(1) stream).
    regen = G.reference_cases(str(tmp_path), ["one_dim_small", "two_dim_small_solve1"])
    for name, path in regen.items():
        assert G.file_digest(path) == manifest[name]


def test_generated_3d_fixtures_match_manifest(big_cases):
    with open(os.path.join(GOLD, "manifest.json")) as f:
        manifest = json.load(f)
    for name, path in big_cases.items():
        assert G.file_digest(path) == manifest[name], name


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", SMALL)
def test_oracle_on_reference_fixtures(name, dtype):
    m = _load(name, dtype)
    d = m["d"].copy()
    oracle.batch(m["a"], m["b"], m["c"], d, ndim=m["ndim"], solvedim=m["solvedim"],
                 dims=m["dims"])
    H.ref_allclose(m["u"], d, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", sorted(G.ON_DEMAND))
def test_oracle_on_reference_fixtures_3d(big_cases, name, dtype):
    m = G.load_case(big_cases[name], dtype)
    d = m["d"].copy()
    oracle.batch(m["a"], m["b"], m["c"], d, ndim=m["ndim"], solvedim=m["solvedim"],
                 dims=m["dims"])
    H.ref_allclose(m["u"], d, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", SMALL)
def test_oracle_on_reference_fixtures_padded(name, dtype):
    """Halo-padded layout with offset pointers (scalar/test/test_cpu.cpp:86-128)."""
    m = _load(name, dtype)
    dims = m["dims"]
    arrs = [H.pad_halo(m[k], dims) for k in "abcd"]
    off = H.halo_offset(dims)
    pads = [t + 2 for t in dims]
    oracle.batch(*arrs, ndim=m["ndim"], solvedim=m["solvedim"], dims=dims, pads=pads,
                 offset=off)
    H.ref_allclose(H.pad_halo(m["u"], dims), arrs[3], dtype)


def _ref_store():
    return np.load(os.path.join(GOLD, "ref_outputs.npz"), allow_pickle=False)


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_oracle_bitexact_vs_committed_reference_outputs(dtype):
    """The reference library's own outputs (tests/golden/ref_outputs.npz)."""
    z = _ref_store()
    isa = str(z["simd_isa"])
    simd_vec = (64 if isa == "avx512" else 32) // np.dtype(dtype).itemsize
    for si, dims in enumerate(z["shapes"].tolist()):
        key = "s%d_%s" % (si, dtype)
        a, b, c, d, u = (z["%s_%s" % (key, k)] for k in "abcdu")
        for sd in range(3):
            for inc in (0, 1):
                for padded in (0, 1):
                    if padded:
                        A, B, Cc, D, U = (H.pad_halo(t, dims) for t in (a, b, c, d, u))
                        pads, off = [t + 2 for t in dims], H.halo_offset(dims)
                    else:
                        A, B, Cc, D, U = (t.copy() for t in (a, b, c, d, u))
                        pads, off = dims, 0
                    oracle.batch(A, B, Cc, D, U, ndim=3, solvedim=sd, dims=dims, pads=pads,
                                 inc=bool(inc), simd_vec=simd_vec, offset=off)
                    got = U if inc else D
                    want = z["%s_sd%d_inc%d_pad%d" % (key, sd, inc, padded)]
                    assert np.array_equal(got, want), (key, sd, inc, padded)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims", [[37, 21, 19], [64, 16, 24], [33], [40, 17], [8, 9, 10]])
def test_oracle_bitexact_vs_reference_library(dims, dtype):
    rng = np.random.default_rng(sum(dims))
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    u = rng.uniform(-1, 1, n).astype(dtype)
    sv = oracle.ref_simd_vec(dtype)
    for sd in range(len(dims)):
        for inc in (False, True):
            d1, u1, d2, u2 = d.copy(), u.copy(), d.copy(), u.copy()
            oracle.batch(a, b, c, d1, u1, ndim=len(dims), solvedim=sd, dims=dims, inc=inc,
                         simd_vec=sv)
            assert oracle.ref_batch(a, b, c, d2, u2, ndim=len(dims), solvedim=sd,
                                    dims=dims, inc=inc) == 0
            assert np.array_equal(u1 if inc else d1, u2 if inc else d2)
            if inc:  # Inc leaves d alone (scalar/src/cpu/trid_cpu.cpp:341-351)
                assert np.array_equal(d2, d)


@pytest.mark.parametrize("N", [1, 2, 3, 7, 33, 64])
def test_oracle_dense_lu_bruteforce(N):
    rng = np.random.default_rng(N)
    for dtype, tol in ((np.float64, 1e-13), (np.float32, 1e-5)):
        a, b, c, d = H.diag_dominant(rng, N, dtype)
        x = d.copy()
        oracle.thomas(a, b, c, x, N=N)
        assert H.norm_rel_err(x, H.dense_solve(a, b, c, d)) < tol


def test_oracle_closed_forms():
    # N = 1: x = d / b, whatever a and c hold.
    a, b, c, d = (np.array([v]) for v in (5.0, 4.0, -3.0, 2.0))
    oracle.thomas(a, b, c, d, N=1)
    assert d[0] == 0.5
    # N = 2: Cramer's rule.
    a = np.array([9.0, 1.5]); b = np.array([3.0, 4.0]); c = np.array([2.0, 7.0])
    d = np.array([5.0, 6.0])
    det = b[0] * b[1] - c[0] * a[1]
    x0 = (d[0] * b[1] - c[0] * d[1]) / det
    x1 = (b[0] * d[1] - a[1] * d[0]) / det
    oracle.thomas(a, b, c, d, N=2)
    np.testing.assert_allclose(d, [x0, x1], rtol=1e-15)
    # a = c = 0: x = (1/b) * d, the reference's reciprocal-then-multiply
    # (trid_cpu.cpp:321-334), within one rounding of d / b.
    rng = np.random.default_rng(3)
    b = rng.uniform(1, 2, 50); d = rng.uniform(-1, 1, 50)
    x = d.copy()
    oracle.thomas(np.zeros(50), b, np.zeros(50), x, N=50)
    assert np.array_equal(x, (1.0 / b) * d)
    np.testing.assert_allclose(x, d / b, rtol=2.3e-16, atol=0)


def test_oracle_manufactured_solution_and_residual():
    """d = A x* for a known x*; solve recovers x*; long-double residual ~ eps."""
    rng = np.random.default_rng(11)
    dims = [24, 20, 16]
    n = int(np.prod(dims))
    a, b, c, _ = H.diag_dominant(rng, n)
    xs = rng.uniform(-1, 1, n)
    for sd in range(3):
        d = np.zeros(n)
        for idx in H.system_slices(3, sd, dims):
            N = len(idx)
            v = b[idx] * xs[idx]
            v[1:] += a[idx][1:] * xs[idx][:-1]
            v[:-1] += c[idx][:-1] * xs[idx][1:]
            d[idx] = v
        x = d.copy()
        oracle.batch(a, b, c, x, ndim=3, solvedim=sd, dims=dims)
        assert H.norm_rel_err(x, xs) < 1e-14
        r, dm = oracle.residual(a, b, c, d, x, ndim=3, solvedim=sd, dims=dims)
        assert r / dm < 1e-15
        xl = oracle.batch_ld(a, b, c, d, ndim=3, solvedim=sd, dims=dims)
        assert H.norm_rel_err(xl, xs) < 1e-15


def test_planted_mistakes_are_caught():
    """A dropped term or a transposed operand fails the dense-LU check."""
    rng = np.random.default_rng(5)
    N = 16
    a, b, c, d = H.diag_dominant(rng, N)
    ref = H.dense_solve(a, b, c, d)

    def thomas_py(a, b, c, d, drop=False, swap=False):
        if swap:
            a, c = c, a
        cp = np.zeros(N); dp = np.zeros(N)
        cp[0] = c[0] / b[0]; dp[0] = d[0] / b[0]
        for i in range(1, N):
            m = b[i] - (0 if drop else a[i] * cp[i - 1])
            cp[i] = c[i] / m
            dp[i] = (d[i] - a[i] * dp[i - 1]) / m
        x = np.zeros(N); x[-1] = dp[-1]
        for i in range(N - 2, -1, -1):
            x[i] = dp[i] - cp[i] * x[i + 1]
        return x

    assert H.norm_rel_err(thomas_py(a, b, c, d), ref) < 1e-13
    assert H.norm_rel_err(thomas_py(a, b, c, d, drop=True), ref) > 1e-3
    assert H.norm_rel_err(thomas_py(a, b, c, d, swap=True), ref) > 1e-3


def test_oracle_ragged_matches_dense():
    rng = np.random.default_rng(8)
    lens = [1, 2, 5, 64, 3, 130]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]))
    x = d.copy()
    oracle.ragged(a, b, c, x, off)
    for s in range(len(lens)):
        sl = slice(off[s], off[s + 1])
        assert H.norm_rel_err(x[sl], H.dense_solve(a[sl], b[sl], c[sl], d[sl])) < 1e-13
    u0 = rng.uniform(-1, 1, int(off[-1]))
    u = u0.copy()
    oracle.ragged(a, b, c, d.copy(), off, u, inc=True)
    np.testing.assert_allclose(u - u0, x, rtol=1e-12, atol=1e-14)


def test_oracle_ragged_s_matches_dense_and_batch():
    """orc_ragged_s: each FP32 system against a dense FP64 solve of the same
    values (FP32 tolerance), bit for bit against orc_batch_s run on that slice
    alone (the same trid_scalar arithmetic), and Inc = plain + u."""
    rng = np.random.default_rng(18)
    lens = [1, 2, 7, 64, 0, 3, 200, 33]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]), np.float32)
    x = d.copy()
    oracle.ragged(a, b, c, x, off)
    for s in range(len(lens)):
        sl = slice(off[s], off[s + 1])
        if lens[s] == 0:
            continue
        want = H.dense_solve(*(v[sl].astype(np.float64) for v in (a, b, c, d)))
        assert H.norm_rel_err(x[sl], want) < 1e-5, s
        one = d[sl].copy()
        oracle.batch(a[sl].copy(), b[sl].copy(), c[sl].copy(), one, ndim=1, solvedim=0,
                     dims=[lens[s]])
        assert np.array_equal(one, x[sl]), s
    u0 = rng.uniform(-1, 1, int(off[-1])).astype(np.float32)
    u = u0.copy()
    dd = d.copy()
    oracle.ragged(a, b, c, dd, off, u, inc=True)
    assert np.array_equal(dd, d)
    np.testing.assert_allclose(u - u0, x, rtol=1e-5, atol=1e-6)


def test_oracle_ragged_ld_matches_dense():
    """orc_ragged_ld (the long-double error reference of ragged batches)
    against dense FP64 LU per system; entries past the last system stay 0."""
    rng = np.random.default_rng(19)
    lens = [1, 2, 5, 64, 0, 130]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]) + 3)
    x = oracle.ragged_ld(a, b, c, d, off)
    for s in range(len(lens)):
        sl = slice(off[s], off[s + 1])
        if lens[s]:
            assert H.norm_rel_err(x[sl], H.dense_solve(a[sl], b[sl], c[sl], d[sl])) < 1e-14
    assert np.all(x[off[-1]:] == 0)
    # a dropped sub-diagonal term in the oracle would show: compare against a
    # dense solve of the system without it
    sl = slice(off[3], off[4])
    a0 = a[sl].copy()
    a0[1:] = 0
    assert H.norm_rel_err(x[sl], H.dense_solve(a0, b[sl], c[sl], d[sl])) > 1e-3


def test_oracle_adjoint_identities():
    """A^T lam = w; <w, A^-1 d> = <lam, d>; gradients by central differences."""
    rng = np.random.default_rng(9)
    dims = [12, 10, 8]
    n = int(np.prod(dims))
    for sd in range(3):
        a, b, c, d = H.diag_dominant(rng, n)
        w = rng.standard_normal(n)
        x = d.copy()
        oracle.batch(a, b, c, x, ndim=3, solvedim=sd, dims=dims)
        lam, ga, gb, gc = oracle.adjoint(a, b, c, x, w, ndim=3, solvedim=sd, dims=dims)
        assert abs(np.dot(w, x) - np.dot(lam, d)) < 1e-12 * np.abs(w).sum()
        # dense transpose solve on each system
        for idx in list(H.system_slices(3, sd, dims))[:5]:
            N = len(idx)
            A = np.diag(b[idx]) + np.diag(a[idx][1:], -1) + np.diag(c[idx][:-1], 1)
            np.testing.assert_allclose(lam[idx], np.linalg.solve(A.T, w[idx]), rtol=1e-12,
                                       atol=1e-14)
            assert ga[idx][0] == 0 and gc[idx][-1] == 0

        def loss(a_, b_, c_, d_):
            xx = d_.copy()
            oracle.batch(a_, b_, c_, xx, ndim=3, solvedim=sd, dims=dims)
            return float(np.dot(w, xx))

        h = 1e-6
        for arr, g in ((a, ga), (b, gb), (c, gc), (d, lam)):
            for i in rng.choice(n, 8, replace=False):
                old = arr[i]
                arr[i] = old + h; lp = loss(a, b, c, d)
                arr[i] = old - h; lm = loss(a, b, c, d)
                arr[i] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_oracle_adi_preproc_closed_form():
    rng = np.random.default_rng(4)
    nx, ny, nz = 7, 6, 5
    u = rng.uniform(0, 1, nx * ny * nz)
    for dt, lam in ((np.float64, 1.0), (np.float64, 0.3), (np.float32, 1.0)):
        uu = u.astype(dt)
        du, a, b, c = oracle.adi_preproc(lam, uu, nx, ny, nz)
        U = uu.astype(np.float64).reshape(nz, ny, nx)
        lap = np.zeros_like(U)
        lap[1:-1, 1:-1, 1:-1] = (U[1:-1, 1:-1, :-2] + U[1:-1, 1:-1, 2:] + U[1:-1, :-2, 1:-1]
                                 + U[1:-1, 2:, 1:-1] + U[:-2, 1:-1, 1:-1]
                                 + U[2:, 1:-1, 1:-1] - 6 * U[1:-1, 1:-1, 1:-1])
        inner = np.zeros_like(U, dtype=bool)
        inner[1:-1, 1:-1, 1:-1] = True
        inner = inner.ravel()
        tol = 1e-14 if dt == np.float64 else 1e-5
        np.testing.assert_allclose(du, (lam * lap).ravel(), atol=tol)
        assert np.all(a[inner] == dt(-0.5 * lam)) and np.all(c[inner] == dt(-0.5 * lam))
        assert np.all(b[inner] == dt(1 + lam))
        assert np.all(a[~inner] == 0) and np.all(c[~inner] == 0) and np.all(b[~inner] == 1)
        assert np.all(du[~inner] == 0)


@pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built here")
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_oracle_adi_matches_reference_library_driven_adi(dtype):
    """ADI steps: the oracle's adi_orig restatement vs the reference library
    called in the ADI pattern (apps/adi/src/adi_cpu.cpp:223-269)."""
    rng = np.random.default_rng(6)
    nx, ny, nz = 24, 20, 18
    u0 = rng.uniform(0, 1, nx * ny * nz).astype(dtype)
    u1 = u0.copy()
    oracle.adi(1.0, u1, nx, ny, nz, 3)
    u2 = u0.copy()
    dims = [nx, ny, nz]
    for _ in range(3):
        du, a, b, c = oracle.adi_preproc(1.0, u2, nx, ny, nz)
        for sd in (0, 1):
            oracle.ref_batch(a, b, c, du, ndim=3, solvedim=sd, dims=dims)
        oracle.ref_batch(a, b, c, du, u2, ndim=3, solvedim=2, dims=dims, inc=True)
    # x rows take the fused SIMD path in the library when nx % SIMD_VEC == 0;
    # the oracle ADI uses the scalar form everywhere.
    H.assert_norm_close(u1, u2, dtype, tol=1e-14 if dtype == np.float64 else 1e-6)
