"""Shared generators and tolerance checks for the parity tests."""
from __future__ import annotations

import numpy as np

# The reference tests' tolerance, |e - a| <= atol + rtol * min(|e|, |a|)
# (scalar/test/utils.hpp:14-17, scalar/test/catch_utils.hpp:62-84).
REF_TOL = {np.float64: (1e-11, 1e-11), np.float32: (1e-6, 1e-5)}
# Normwise relative tolerance against the oracle on identical inputs:
# max|x - x_oracle| / max|x_oracle|.
NORM_TOL = {np.float64: 1e-12, np.float32: 1e-5}


def dt_key(dtype):
    return np.float64 if np.dtype(dtype) == np.float64 else np.float32


def ref_allclose(expected, actual, dtype=None):
    """Elementwise check with the reference tests' formula; returns the worst
    (diff - tolerance) and asserts it is <= 0."""
    dtype = dt_key(dtype or expected.dtype)
    atol, rtol = REF_TOL[dtype]
    e = np.asarray(expected, dtype=np.float64)
    a = np.asarray(actual, dtype=np.float64)
    tol = atol + rtol * np.minimum(np.abs(e), np.abs(a))
    diff = np.abs(e - a)
    bad = diff > tol
    if bad.any():
        i = int(np.argmax(diff - tol))
        raise AssertionError("mismatch at %d: expected %r actual %r diff %.3e tol %.3e (%d bad)"
                             % (i, e.flat[i], a.flat[i], diff.flat[i], tol.flat[i], bad.sum()))


def norm_rel_err(x, ref):
    x = np.asarray(x, dtype=np.float64)
    ref = np.asarray(ref, dtype=np.float64)
    den = np.abs(ref).max()
    return float(np.abs(x - ref).max() / (den if den > 0 else 1.0))


def assert_norm_close(x, ref, dtype=None, tol=None):
    dtype = dt_key(dtype or ref.dtype)
    tol = NORM_TOL[dtype] if tol is None else tol
    err = norm_rel_err(x, ref)
    assert err <= tol, "normwise relative error %.3e > %.1e" % (err, tol)
    return err


def diag_dominant(rng, n, dtype=np.float64):
    """a,c ~ U(-1,1), b = 2+|a|+|c|+U(0,1), d ~ U(-1,1) (the benchmark
    generator: strict dominance margin >= 2)."""
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 2 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, n)
    d = rng.uniform(-1, 1, n)
    return [np.ascontiguousarray(t.astype(dtype)) for t in (a, b, c, d)]


def dense_solve(a, b, c, d):
    """Dense LU solve of one system (numpy/LAPACK); a[0], c[-1] ignored."""
    n = len(b)
    A = np.diag(np.asarray(b, dtype=np.float64))
    if n > 1:
        A += np.diag(np.asarray(a[1:], dtype=np.float64), -1)
        A += np.diag(np.asarray(c[:-1], dtype=np.float64), 1)
    return np.linalg.solve(A, np.asarray(d, dtype=np.float64))


def system_slices(ndim, solvedim, dims, pads=None, offset=0):
    """Yields the flat index arrays of every system (x fastest)."""
    pads = dims if pads is None else pads
    d3 = list(dims) + [1] * (3 - ndim)
    p3 = list(pads) + [1] * (3 - ndim)
    strides = [1, p3[0], p3[0] * p3[1]]
    others = [k for k in range(3) if k != solvedim]
    N = d3[solvedim]
    for i1 in range(d3[others[1]]):
        for i0 in range(d3[others[0]]):
            base = offset + i0 * strides[others[0]] + i1 * strides[others[1]]
            yield base + np.arange(N) * strides[solvedim]


def pad_halo(x, dims):
    """One zero layer around every dimension (scalar/test/catch_utils.hpp:86-140)."""
    arr = np.asarray(x).reshape(list(dims)[::-1])
    return np.ascontiguousarray(np.pad(arr, 1)).ravel()


def halo_offset(dims):
    """Offset of the first interior element (scalar/test/test_cpu.cpp:109-114)."""
    p = [t + 2 for t in dims]
    off, stride = 0, 1
    for i in range(len(dims)):
        off += stride
        stride *= p[i]
    return off
