"""Full-size (512^3) checks through size-independent properties, where the
CPU oracle would take too long: the residual max|d - A x| / max|d| of the GPU
solution, computed in FP64 by torch on the GPU, stays at rounding level for
every direction, for FP64 and FP32 (the FP32 residual evaluated in FP64 on the
FP32 data), for the mixed-precision solve, and Inc = plain + u.

Tolerances: FP64 and mixed residual <= 1e-14 relative; FP32 <= 1e-5.
"""
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import api  # noqa: E402

N = 512


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    yield
    torch.cuda.empty_cache()


def gen(dtype, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    n3 = N ** 3
    a = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    c = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    b = 2 + a.abs() + c.abs() + torch.rand(n3, dtype=dtype, device="cuda", generator=g)
    d = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    return a, b, c, d


def residual(a, b, c, d, x, sd):
    """max |d - A x| / max |d| along torch dim (2 - sd) of [z, y, x] views, FP64."""
    shp = (N, N, N)
    A, B, C, D, X = (t.double().view(shp) for t in (a, b, c, d, x))
    dim = 2 - sd
    r = D - B * X
    n = N
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[dim] = slice(1, n)
    hi[dim] = slice(0, n - 1)
    r[tuple(lo)] -= A[tuple(lo)] * X[tuple(hi)]
    r[tuple(hi)] -= C[tuple(hi)] * X[tuple(lo)]
    return float(r.abs().max() / D.abs().max())


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-14), (torch.float32, 1e-5)])
def test_residual_512cube(dtype, tol):
    a, b, c, d = gen(dtype, 5120)
    for sd in range(3):
        x = d.clone()
        api.trid_strided_batch(a, b, c, x, solvedim=sd, dims=[N, N, N], sync=True)
        res = residual(a, b, c, d, x, sd)
        print("512^3 %s sd=%d residual %.2e config %s" % (dtype, sd, res, api.last_config()))
        assert res <= tol
        del x


def test_mixed_residual_512cube():
    a, b, c, d = gen(torch.float64, 5121)
    for sd in range(3):
        x = d.clone()
        api.mixed_strided_batch(a, b, c, x, solvedim=sd, dims=[N, N, N], refinements=2, sync=True)
        res = residual(a, b, c, d, x, sd)
        print("512^3 mixed sd=%d residual %.2e" % (sd, res))
        assert res <= 1e-14
        del x


def test_inc_equals_plain_plus_u_512cube():
    a, b, c, d = gen(torch.float64, 5122)
    u0 = torch.rand_like(d)
    for sd in range(3):
        x = d.clone()
        api.trid_strided_batch(a, b, c, x, solvedim=sd, dims=[N, N, N])
        u = u0.clone()
        api.trid_strided_batch(a, b, c, d, solvedim=sd, dims=[N, N, N], u=u, sync=True)
        assert float((u - (u0 + x)).abs().max()) <= 1e-15 * float(u.abs().max())
        del x, u


def test_more_than_2_pow_32_work_items():
    """2^31 - 1 systems of 2 rows along x (FP32, 16 GiB per array): the
    default shape (one row per lane, 128 systems per 256-thread workgroup)
    needs 2^32 work-items, more than one HIP grid holds, so the launch is
    split into sub-batches. Every system is checked against the 2 x 2
    closed form on a sample of 200,000 systems spread over the whole batch
    (first and last included), and the residual on all of them."""
    ns = 2 ** 31 - 1
    n = 2 * ns
    g = torch.Generator(device="cuda")
    g.manual_seed(2)
    a = torch.rand(n, dtype=torch.float32, device="cuda", generator=g) * 2 - 1
    c = torch.rand(n, dtype=torch.float32, device="cuda", generator=g) * 2 - 1
    b = 2 + a.abs() + c.abs()
    d = torch.rand(n, dtype=torch.float32, device="cuda", generator=g) * 2 - 1
    x = d.clone()
    api.trid_strided_batch(a, b, c, x, solvedim=0, dims=[2, ns], sync=True)
    assert api.last_config()[:4] == (3, 1, 2, 128)
    # residual of both rows of every system (a[0] and c[1] unused), in FP64 by chunks
    worst = 0.0
    step = 1 << 28
    for s0 in range(0, ns, step):
        s1 = min(ns, s0 + step)
        sl = slice(2 * s0, 2 * s1)
        av, bv, cv, dv, xv = (t[sl].view(-1, 2).double() for t in (a, b, c, d, x))
        r0 = dv[:, 0] - bv[:, 0] * xv[:, 0] - cv[:, 0] * xv[:, 1]
        r1 = dv[:, 1] - av[:, 1] * xv[:, 0] - bv[:, 1] * xv[:, 1]
        worst = max(worst, float(torch.maximum(r0.abs(), r1.abs()).max()))
        del av, bv, cv, dv, xv, r0, r1
    assert worst <= 1e-5, worst
    # closed form on a sample
    idx = torch.cat([torch.tensor([0, ns - 1], device="cuda"),
                     torch.randint(0, ns, (200000,), device="cuda", generator=g)])
    a1, b0, b1, c0, d0, d1 = (t[2 * idx + k].double() for t, k in
                              ((a, 1), (b, 0), (b, 1), (c, 0), (d, 0), (d, 1)))
    det = b0 * b1 - c0 * a1
    x0 = (d0 * b1 - c0 * d1) / det
    x1 = (b0 * d1 - a1 * d0) / det
    got0, got1 = x[2 * idx].double(), x[2 * idx + 1].double()
    err = float(torch.maximum((got0 - x0).abs(), (got1 - x1).abs()).max())
    assert err <= 1e-5 * float(torch.maximum(x0.abs(), x1.abs()).max()), err
    del a, b, c, d, x
    torch.cuda.empty_cache()
