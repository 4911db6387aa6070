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
