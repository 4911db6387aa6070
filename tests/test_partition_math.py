"""The partitioned solve's algebra, checked on the CPU against dense solves.

A system of N rows is split into G contiguous shards of n_k >= 2 rows (as G
GPUs would hold it). Per shard, with F = its first unknown and E = its last:
  * interior rows 1..n-2 are written as x_i = dd_i - aa_i F - cc_i E by a
    top-down sweep (gives row n-2) and a bottom-up sweep (gives row 1);
  * rows 0 and n-1 then read, normalised by their diagonal,
        alpha E_prev + F + gamma E       = delta
        alpha' F     + E + gamma' F_next = delta'
    -- 6 numbers per system and shard (the reference's distributed solver
    exchanges 6 as well, scalar/src/cuda/trid_cuda_mpi.cu:454-519);
  * the 2G unknowns (F_0, E_0, F_1, E_1, ...) form a tridiagonal system with
    unit diagonal; shard k needs only E_{k-1} and F_{k+1}, each obtained by a
    twisted (top-down meets bottom-up) elimination without storage;
  * shard k then solves its own rows with d_0 -= a_0 E_{k-1} and
    d_{n-1} -= c_{n-1} F_{k+1}.
This file restates that algebra in numpy (test code only) and pins it with
dense solves; the GPU kernels are checked against the oracle's global solve.
"""
import numpy as np
import pytest


def dense(a, b, c, d):
    n = len(b)
    A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    return np.linalg.solve(A, d)


def boundary(a, b, c, d, first, last):
    """6 numbers of one shard (a[0] / c[-1] are couplings to the neighbours;
    ignored for the first / last shard)."""
    n = len(b)
    assert n >= 2
    a0 = 0.0 if first else a[0]
    cn = 0.0 if last else c[-1]
    if n == 2:
        dd1, aa1, cc1 = 0.0, 0.0, -1.0  # row 1 is E itself
        ddm, aam, ccm = 0.0, -1.0, 0.0  # row n-2 is F itself
    else:
        # top-down over rows 1..n-2: x_i = D - A F - C x_{i+1}
        D, A, C = d[1] / b[1], a[1] / b[1], c[1] / b[1]
        for i in range(2, n - 1):
            beta = b[i] - a[i] * C
            D, A, C = (d[i] - a[i] * D) / beta, -a[i] * A / beta, c[i] / beta
        ddm, aam, ccm = D, A, C
        # bottom-up over rows n-2..1: x_i = D' - C' E - A' x_{i-1}
        D, C, A = d[n - 2] / b[n - 2], c[n - 2] / b[n - 2], a[n - 2] / b[n - 2]
        for i in range(n - 3, 0, -1):
            gam = b[i] - c[i] * A
            D, C, A = (d[i] - c[i] * D) / gam, -c[i] * C / gam, a[i] / gam
        dd1, aa1, cc1 = D, A, C
    b0 = b[0] - c[0] * aa1
    bE = b[n - 1] - a[n - 1] * ccm
    return (a0 / b0, -c[0] * cc1 / b0, (d[0] - c[0] * dd1) / b0,
            -a[n - 1] * aam / bE, cn / bE, (d[n - 1] - a[n - 1] * ddm) / bE)


def row1_forward(a, b, c, d):
    """Row 1 of a shard as x_1 = dd1 - aa1 F - cc1 E from the top-down sweep
    alone: each new row x_i = D_i - A_i F - C_i x_{i+1} is substituted into
    x_1 = P + Q F + R x_i (the GPU boundary step's one-pass form)."""
    n = len(b)
    D, A, C = d[1] / b[1], a[1] / b[1], c[1] / b[1]
    P, Q, R = D, -A, -C
    for i in range(2, n - 1):
        beta = b[i] - a[i] * C
        D, A, C = (d[i] - a[i] * D) / beta, -a[i] * A / beta, c[i] / beta
        P, Q, R = P + R * D, Q - R * A, -R * C
    return P, -Q, -R


def reduced_rows(recs):
    """(sub, sup, rhs) of the 2G-row system, unit diagonal."""
    sub, sup, rhs = [], [], []
    for al, ga, de, al2, ga2, de2 in recs:
        sub += [al, al2]
        sup += [ga, ga2]
        rhs += [de, de2]
    return np.array(sub), np.array(sup), np.array(rhs)


def twisted(sub, sup, rhs, j):
    R = len(rhs)
    cf, df = 0.0, 0.0
    for i in range(j):
        den = 1.0 - sub[i] * cf
        cf, df = sup[i] / den, (rhs[i] - sub[i] * df) / den
    ar, dr = 0.0, 0.0
    for i in range(R - 1, j, -1):
        den = 1.0 - sup[i] * ar
        ar, dr = sub[i] / den, (rhs[i] - sup[i] * dr) / den
    return (rhs[j] - sub[j] * df - sup[j] * dr) / (1.0 - sub[j] * cf - sup[j] * ar)


def partitioned(a, b, c, d, sizes):
    G = len(sizes)
    off = np.concatenate([[0], np.cumsum(sizes)])
    recs = [boundary(a[off[k]:off[k + 1]], b[off[k]:off[k + 1]], c[off[k]:off[k + 1]],
                     d[off[k]:off[k + 1]], k == 0, k == G - 1) for k in range(G)]
    sub, sup, rhs = reduced_rows(recs)
    x = np.empty_like(d)
    for k in range(G):
        sl = slice(off[k], off[k + 1])
        aa, bb, cc, dl = a[sl].copy(), b[sl], c[sl].copy(), d[sl].copy()
        if k > 0:
            dl[0] -= aa[0] * twisted(sub, sup, rhs, 2 * k - 1)
        if k < G - 1:
            dl[-1] -= cc[-1] * twisted(sub, sup, rhs, 2 * k + 2)
        aa[0], cc[-1] = 0.0, 0.0
        x[sl] = dense(aa, bb, cc, dl)
    return x


def system(rng, n):
    a = rng.uniform(-1, 1, n)
    c = rng.uniform(-1, 1, n)
    b = 2 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, n)
    d = rng.uniform(-1, 1, n)
    a[0] = 7.0  # never used
    c[-1] = -7.0
    return a, b, c, d


@pytest.mark.parametrize("sizes", [[9], [2, 2], [2, 3, 2], [5, 8, 2, 13], [30] * 7,
                                   [2] * 16, [3, 40, 3, 40, 2]])
def test_partitioned_equals_dense(sizes):
    rng = np.random.default_rng(len(sizes) * 31 + sum(sizes))
    a, b, c, d = system(rng, sum(sizes))
    aa, cc = a.copy(), c.copy()
    aa[0], cc[-1] = 0.0, 0.0
    want = dense(aa, b, cc, d)
    got = partitioned(a, b, c, d, sizes)
    assert np.abs(got - want).max() <= 1e-13 * np.abs(want).max()


def test_boundary_of_two_and_three_rows():
    """n = 2: the rows are the shard's equations themselves; n = 3: both
    sweeps reduce to the single interior row."""
    a, b, c, d = (np.array(v, float) for v in ([9, 2, 3], [4, 5, 6], [1, 2, 9], [7, 8, 9]))
    r2 = boundary(a[:2], b[:2], np.array([1.0, 9.0]), d[:2], False, False)
    assert r2 == pytest.approx((9 / 4, 1 / 4, 7 / 4, 2 / 5, 9 / 5, 8 / 5))
    r3 = boundary(a, b, c, d, True, True)
    # row 1: 2 F + 5 x1 + 2 E = 8, so x1 = 8/5 - (2/5) F - (2/5) E;
    # row 0: 4 F + x1 = 7        ->  (4 - 2/5) F - (2/5) E = 7 - 8/5
    # row 2: 3 x1 + 6 E = 9      ->  -(6/5) F + (6 - 6/5) E = 9 - 24/5
    b0 = 4 - 2 / 5
    bE = 6 - 6 / 5
    assert r3 == pytest.approx((0.0, -(2 / 5) / b0, (7 - 8 / 5) / b0,
                                -(6 / 5) / bE, 0.0, (9 - 24 / 5) / bE))


@pytest.mark.parametrize("n", [3, 4, 5, 17, 64, 200])
def test_row1_from_the_top_down_sweep_alone(n):
    """The one-pass form of row 1 (the GPU boundary step's) equals the
    bottom-up sweep's and the dense truth: with the interior matrix A_int of
    rows 1..n-2, x_int = A_int^-1 (d_int - a_1 F e_first - c_{n-2} E e_last),
    so dd1 = (A_int^-1 d_int)_first, aa1 = a_1 (A_int^-1 e_first)_first,
    cc1 = c_{n-2} (A_int^-1 e_last)_first."""
    rng = np.random.default_rng(n)
    a, b, c, d = system(rng, n)
    fwd = row1_forward(a, b, c, d)
    ai, bi, ci = a[1:n - 1], b[1:n - 1], c[1:n - 1]
    A = np.diag(bi) + np.diag(ai[1:], -1) + np.diag(ci[:-1], 1)
    m = n - 2
    e_first, e_last = np.eye(m)[0], np.eye(m)[m - 1]
    truth = (np.linalg.solve(A, d[1:n - 1])[0], a[1] * np.linalg.solve(A, e_first)[0],
             c[n - 2] * np.linalg.solve(A, e_last)[0])
    assert fwd == pytest.approx(truth, rel=1e-12, abs=1e-14)
    # the bottom-up sweep of the two-sweep form gives the same numbers
    D, C, Aa = d[n - 2] / b[n - 2], c[n - 2] / b[n - 2], a[n - 2] / b[n - 2]
    for i in range(n - 3, 0, -1):
        gam = b[i] - c[i] * Aa
        D, C, Aa = (d[i] - c[i] * D) / gam, -c[i] * C / gam, a[i] / gam
    assert fwd == pytest.approx((D, Aa, C), rel=1e-12, abs=1e-14)


def test_twisted_matches_dense_reduced():
    rng = np.random.default_rng(3)
    R = 12
    sub = rng.uniform(-0.3, 0.3, R)
    sup = rng.uniform(-0.3, 0.3, R)
    rhs = rng.uniform(-1, 1, R)
    A = np.eye(R) + np.diag(sub[1:], -1) + np.diag(sup[:-1], 1)
    sub[0], sup[-1] = 0.0, 0.0
    x = np.linalg.solve(A, rhs)
    for j in range(R):
        assert twisted(sub, sup, rhs, j) == pytest.approx(x[j], rel=1e-13, abs=1e-14)
