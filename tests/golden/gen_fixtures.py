"""Writes the golden fixtures used by the parity tests.

Two kinds of fixture live in tests/golden/:

1. ``files/<case>`` in the reference's text fixture format
   (scalar/test/gen_testcases.py:44-74): line 1 ``ndim solvedim``, line 2 the
   dims, then a, b, c, d, u (FP64 solution), u_float (solution of the inputs
   rounded to 6 decimals), each in column-major (x fastest) order.
   The generator follows gen_testcases.py:76-109: numpy legacy ``seed(1)``,
   the same case order and shapes, a = -1+0.1U, b = 2+U, c = -1+0.1U, d = U,
   and scipy's sparse ``spsolve`` per system as the solver (numpy and scipy are
   libraries, not the reference). Only the cases up to 32x32 are committed as
   text; the 32^3 cases are regenerated on demand by ``reference_cases(out)``
   and pinned by the SHA-256 digests in ``manifest.json``.

2. ``ref_outputs.npz``: small inputs and the outputs of the reference CPU
   library compiled from its own sources (oracle/_ref), for every solvedim,
   FP32/FP64, plain and Inc, unpadded and halo-padded. This pins the oracle
   against the reference itself where oracle/_ref cannot be built.

Run from the repository root:  python tests/golden/gen_fixtures.py
"""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PRECISION = 15
PRECISION_FLOAT = 6
# (name, shape, solvedim) in the order gen_testcases.py:96-109 draws them.
CASES = [
    ("one_dim_small", [5], 0),
    ("one_dim_large", [200], 0),
    ("two_dim_small_solve0", [8, 8], 0),
    ("two_dim_small_solve1", [8, 8], 1),
    ("two_dim_large_solve0", [32, 32], 0),
    ("two_dim_large_solve1", [32, 32], 1),
    ("three_dim_large_solve0", [32, 32, 32], 0),
    ("three_dim_large_solve1", [32, 32, 32], 1),
    ("three_dim_large_solve2", [32, 32, 32], 2),
    ("four_dim_large_solve0", [32, 32, 32, 32], 0),
    ("four_dim_large_solve1", [32, 32, 32, 32], 1),
    ("four_dim_large_solve2", [32, 32, 32, 32], 2),
    ("four_dim_large_solve3", [32, 32, 32, 32], 3),
]
COMMITTED = {"one_dim_small", "one_dim_large", "two_dim_small_solve0",
             "two_dim_small_solve1", "two_dim_large_solve0", "two_dim_large_solve1"}
ON_DEMAND = {"three_dim_large_solve0", "three_dim_large_solve1", "three_dim_large_solve2"}


def solve_axis(a, b, c, d, axis, precision=0):
    """Solution of every system along ``axis`` (numpy C-order array whose axis
    ``axis`` is the solve axis), by scipy's sparse direct solver, optionally on
    inputs rounded to ``precision`` decimals (gen_testcases.py:9-42)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as slg

    N = a.shape[axis]
    move = lambda t: np.moveaxis(t, axis, -1).reshape(-1, N)
    aa, bb, cc, dd = move(a), move(b), move(c), move(d)
    if precision > 0:
        aa, bb, cc, dd = (t.round(precision) for t in (aa, bb, cc, dd))
    uu = np.zeros(dd.shape)
    for i in range(aa.shape[0]):
        m = sp.diags([aa[i, 1:], bb[i, :], cc[i, :-1]], [-1, 0, 1], format="csc")
        uu[i, :] = slg.spsolve(m, dd[i, :])
    shp = list(np.moveaxis(d, axis, -1).shape)
    return np.moveaxis(uu.reshape(shp), -1, axis)


def _fmt(arr, prec):
    return " ".join(str(round(v, prec)) for v in arr.flatten(order="F"))


def write_case(path, a, b, c, d, u, u_float, solvedim):
    with open(path, "w") as f:
        f.write("{} {}\n".format(a.ndim, solvedim))
        f.write(" ".join(str(s) for s in a.shape) + "\n")
        for arr in (a, b, c, d, u):
            f.write(_fmt(arr, PRECISION) + "\n")
        f.write(_fmt(u_float, PRECISION_FLOAT) + "\n")


def reference_cases(out_dir, names=None):
    """Regenerates the reference-format cases into ``out_dir``.

    The numpy array of ``shape`` is indexed [x, y, z, ...] as in the reference
    generator, where ``flatten(order='F')`` makes x fastest. The 4-D cases are
    drawn (to keep the random stream identical) but never solved or written.
    """
    names = set(n for n, _, _ in CASES) if names is None else set(names)
    os.makedirs(out_dir, exist_ok=True)
    np.random.seed(1)
    written = {}
    for name, shape, sd in CASES:
        a = -1 + np.random.rand(*shape) * 0.1
        b = 2 + np.random.rand(*shape)
        c = -1 + np.random.rand(*shape) * 0.1
        d = np.random.rand(*shape)
        if len(shape) > 3 or name not in names:
            continue
        u = solve_axis(a, b, c, d, sd)
        u_float = solve_axis(a, b, c, d, sd, precision=PRECISION_FLOAT)
        path = os.path.join(out_dir, name)
        write_case(path, a, b, c, d, u, u_float, sd)
        written[name] = path
    return written


def load_case(path, dtype=np.float64):
    """Reads a reference-format fixture like MeshLoader (scalar/test/utils.hpp:207-247).

    Returns dict(ndim, solvedim, dims, a, b, c, d, u) with arrays flat in x-fastest
    order; ``u`` is the FP64 line for float64 and the rounded line for float32.
    """
    with open(path) as f:
        lines = f.read().split("\n")
    ndim, sd = (int(t) for t in lines[0].split())
    dims = [int(t) for t in lines[1].split()]
    arr = [np.array(lines[i].split(), dtype=np.float64) for i in range(2, 8)]
    u = arr[4] if dtype == np.float64 else arr[5]
    cast = lambda t: np.ascontiguousarray(t.astype(dtype))
    return dict(ndim=ndim, solvedim=sd, dims=dims, a=cast(arr[0]), b=cast(arr[1]),
                c=cast(arr[2]), d=cast(arr[3]), u=cast(u))


def file_digest(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


# ---------------------------------------------------------------------------
# Reference-library golden outputs
# ---------------------------------------------------------------------------
REF_SHAPES = [[16, 6, 5], [13, 7, 5]]


def baseline_generator(rng, n, dtype):
    """The diagonally dominant generator: a,c ~ U(-1,1), b = 2+|a|+|c|+U(0,1),
    d ~ U(-1,1)."""
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 2 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, n)
    d = rng.uniform(-1, 1, n)
    return [np.ascontiguousarray(t.astype(dtype)) for t in (a, b, c, d)]


def pad_halo(x, dims):
    """Adds one zero layer on every side (scalar/test/catch_utils.hpp:86-140)."""
    arr = x.reshape(dims[::-1])
    return np.ascontiguousarray(np.pad(arr, 1)).ravel()


def halo_offset(dims):
    p = [t + 2 for t in dims]
    off, stride = 0, 1
    for i in range(len(dims)):
        off += stride
        stride *= p[i]
    return off


def ref_outputs(path):
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import oracle

    if not oracle.ref_available():
        raise SystemExit("oracle/_ref is not built; run make -C oracle")
    rng = np.random.default_rng(20251202)
    store = {}
    for si, dims in enumerate(REF_SHAPES):
        n = int(np.prod(dims))
        for dt in (np.float64, np.float32):
            key = "s%d_%s" % (si, np.dtype(dt).name)
            a, b, c, d = baseline_generator(rng, n, dt)
            u = np.ascontiguousarray(rng.uniform(-1, 1, n).astype(dt))
            for nm, arr in zip("abcdu", (a, b, c, d, u)):
                store["%s_%s" % (key, nm)] = arr
            for sd in range(3):
                for inc in (0, 1):
                    for padded in (0, 1):
                        if padded:
                            A, B, Cc, D, U = (pad_halo(t, dims) for t in (a, b, c, d, u))
                            pads = [t + 2 for t in dims]
                            off = halo_offset(dims)
                        else:
                            A, B, Cc, D, U = (t.copy() for t in (a, b, c, d, u))
                            pads, off = dims, 0
                        st = oracle.ref_batch(A, B, Cc, D, U, ndim=3, solvedim=sd,
                                              dims=dims, pads=pads, inc=bool(inc), offset=off)
                        assert st == 0
                        res = U if inc else D
                        store["%s_sd%d_inc%d_pad%d" % (key, sd, inc, padded)] = res
    store["shapes"] = np.array(REF_SHAPES, dtype=np.int64)
    store["simd_isa"] = np.array(oracle.ref_isa())
    np.savez_compressed(path, **store)


def main():
    files = os.path.join(HERE, "files")
    written = reference_cases(files, COMMITTED)
    manifest = {name: file_digest(p) for name, p in sorted(written.items())}
    import tempfile
    with tempfile.TemporaryDirectory() as tmp:
        big = reference_cases(tmp, ON_DEMAND)
        for name, p in sorted(big.items()):
            manifest[name] = file_digest(p)
    with open(os.path.join(HERE, "manifest.json"), "w") as f:
        json.dump(manifest, f, indent=1, sort_keys=True)
    ref_outputs(os.path.join(HERE, "ref_outputs.npz"))
    print("wrote", sorted(manifest))


if __name__ == "__main__":
    main()
