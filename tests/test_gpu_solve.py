"""GPU parity of the batched solve through the C ABI (libtridhip.so) against
the oracle on identical seeded inputs, and against the reference-format
fixtures.

Tolerances (written here, used everywhere below):
  normwise  max|x - x_oracle| / max|x_oracle| <= 1e-12 (FP64), 1e-5 (FP32);
  elementwise, the reference tests' formula |e-a| <= atol + rtol*min(|e|,|a|)
  with atol = rtol = 1e-11 (FP64), 1e-6 / 1e-5 (FP32)
  (scalar/test/utils.hpp:14-17).
"""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from tests.golden import gen_fixtures as G

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import api  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)
    yield
    torch.cuda.synchronize()


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def gpu_solve(a, b, c, d, u=None, *, solvedim, dims, pads=None, offset=0,
              opts=(0, 0, 0), sync=True):
    pads = dims if pads is None else pads
    ta, tb, tc, td = (dev(t) for t in (a, b, c, d))
    tu = dev(u) if u is not None else None
    api.trid_strided_batch(ta, tb, tc, td, solvedim=solvedim, dims=dims, pads=pads, u=tu,
                           offset=offset, opts=opts, sync=sync)
    if not sync:
        torch.cuda.synchronize()
    return td.cpu().numpy(), (tu.cpu().numpy() if tu is not None else None)


ALGOS = {"auto": (0, 0, 0), "lane": (1, 1, 1), "hybrid": (3, 3, 3)}


def run_case(dims, solvedim, dtype, inc, opts, seed, garbage_ends=True):
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    if garbage_ends:
        # a[0] and c[N-1] of every system must be ignored; give them values that
        # would visibly corrupt any result that used them.
        for idx in H.system_slices(len(dims), solvedim, dims):
            a[idx[0]] = 1e3
            c[idx[-1]] = -1e3
    u = rng.uniform(-1, 1, n).astype(dtype) if inc else None
    gd, gu = gpu_solve(a, b, c, d, u, solvedim=solvedim, dims=dims, opts=opts)
    od, ou = d.copy(), (u.copy() if inc else None)
    oracle.batch(a, b, c, od, ou, ndim=len(dims), solvedim=solvedim, dims=dims, inc=inc)
    if inc:
        assert np.array_equal(gd, d), "Inc must not write d"
        H.assert_norm_close(gu - u, ou - u, dtype)
        H.ref_allclose(ou, gu, dtype)
    else:
        H.assert_norm_close(gd, od, dtype)
        H.ref_allclose(od, gd, dtype)
    return api.last_config()


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", sorted(G.COMMITTED))
def test_reference_fixtures(name, dtype):
    m = G.load_case("tests/golden/files/" + name, dtype)
    for opts in ALGOS.values():
        gd, _ = gpu_solve(m["a"], m["b"], m["c"], m["d"], solvedim=m["solvedim"],
                          dims=m["dims"], opts=opts)
        H.ref_allclose(m["u"], gd, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reference_fixtures_3d(tmp_path, dtype):
    cases = G.reference_cases(str(tmp_path), G.ON_DEMAND)
    for name, path in sorted(cases.items()):
        m = G.load_case(path, dtype)
        for opts in ALGOS.values():
            gd, _ = gpu_solve(m["a"], m["b"], m["c"], m["d"], solvedim=m["solvedim"],
                              dims=m["dims"], opts=opts)
            H.ref_allclose(m["u"], gd, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("name", sorted(G.COMMITTED))
def test_reference_fixtures_padded(name, dtype):
    m = G.load_case("tests/golden/files/" + name, dtype)
    dims = m["dims"]
    arrs = [H.pad_halo(m[k], dims) for k in "abcd"]
    off, pads = H.halo_offset(dims), [t + 2 for t in dims]
    for opts in ALGOS.values():
        gd, _ = gpu_solve(*arrs, solvedim=m["solvedim"], dims=dims, pads=pads, offset=off,
                          opts=opts)
        want = H.pad_halo(m["u"], dims)
        H.ref_allclose(want, gd, dtype)


@pytest.mark.parametrize("algo", sorted(ALGOS))
@pytest.mark.parametrize("inc", [False, True])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("dims", [[37, 21, 19], [64, 64, 64], [128, 24, 40], [33],
                                  [700], [40, 17], [5, 300]])
def test_random_vs_oracle(dims, dtype, inc, algo):
    for sd in range(len(dims)):
        run_case(dims, sd, dtype, inc, ALGOS[algo], seed=len(dims) * 100 + sd)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("M", [1, 2, 4, 8, 16])
@pytest.mark.parametrize("W,G", [(1, 1), (3, 8), (16, 8), (16, 3)])
def test_forced_hybrid_shapes(M, W, G, dtype):
    """Every compiled M, several W and XCD groupings, lengths that are not
    multiples of M, block counts that are not multiples of G."""
    for dims, sd in (([61, 5, 7], 0), ([19, 45, 3], 1), ([17, 4, 51], 2), ([40, 37, 9], 1)):
        N = dims[sd]
        P = -(-N // M)
        if W * P > (512 if np.dtype(dtype).itemsize * M > 64 else 1024):
            continue
        with api.forced_hybrid(M, P, W, G):
            cfg = run_case(dims, sd, dtype, False, (0, 0, 0), seed=M * 7 + W)
        assert cfg == (3, M, P, W, G)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("W,G", [(1, 1), (16, 8), (32, 3), (5, 16)])
@pytest.mark.parametrize("inc", [False, True])
@pytest.mark.parametrize("variant", [1, 2])
def test_forced_streamed_variant(W, G, inc, variant, dtype):
    """The streamed strided variants (loads a few rows ahead of the sweep; 2:
    one sweep value per row in LDS): the compiled M (8 FP64, 16 FP32), y and
    z, lengths not multiples of M, Inc; twice that M is refused."""
    Ms = (8,) if dtype == np.float64 else (16,)
    dims, sd = [19, 45, 3], 1
    with api.forced_hybrid(2 * Ms[0], -(-dims[sd] // (2 * Ms[0])), W, G, variant):
        with pytest.raises(Exception):
            run_case(dims, sd, dtype, inc, (0, 0, 0), seed=1)
    for M in Ms:
        for dims, sd in (([19, 45, 3], 1), ([17, 4, 51], 2), ([40, 37, 9], 1), ([33, 8, 130], 2)):
            N = dims[sd]
            P = -(-N // M)
            if W * P > 1024:
                continue
            with api.forced_hybrid(M, P, W, G, variant):
                cfg = run_case(dims, sd, dtype, inc, (0, 0, 0), seed=M * 13 + W)
            assert cfg == (5 + variant, M, P, W, G)


def _wide_shape(dtype, M, N, seg=256):
    """P for the wide variant (whole waves: W*P a multiple of 64), or None if
    the shape does not fit (threads, LDS)."""
    esize = np.dtype(dtype).itemsize
    W = seg // esize
    per_wave = max(1, 64 // W)
    P = -(-N // M)
    P = -(-P // per_wave) * per_wave
    threads = 512 if (dtype == np.float64 and M == 32) else 1024
    lds = P * M * seg + 3 * W * P * esize  # the tile and a one-slot PCR ring
    if W * P > threads or lds > 160 * 1024:
        return None, W
    return P, W


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("seg", [128, 256])
@pytest.mark.parametrize("G", [1, 8])
@pytest.mark.parametrize("inc", [False, True])
def test_forced_wide_variant(G, inc, seg, dtype):
    """The strided variant with a staged in LDS by global->LDS loads (128- or
    256-byte rows): both compiled M, y and z, partial workgroups (n0 not a
    multiple of W, n0 < W), lengths not multiples of M, N up to 512, Inc."""
    cases = (([40, 37, 9], 1), ([17, 4, 51], 2), ([64, 3, 100], 2), ([33, 8, 130], 2),
             ([70, 256, 2], 1), ([32, 2, 512], 2), ([96, 300, 1], 1), ([20, 3, 61], 2))
    ran = 0
    for M in (16, 32):
        for dims, sd in cases:
            P, W = _wide_shape(dtype, M, dims[sd], seg)
            if P is None:
                continue
            with api.forced_hybrid(M, P, W, G, 3):
                cfg = run_case(dims, sd, dtype, inc, (0, 0, 0), seed=M * 7 + dims[sd])
            assert cfg == (8, M, P, W, G)
            ran += 1
    assert ran >= 8


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
@pytest.mark.parametrize("G", [1, 8])
def test_forced_long32_contiguous(G, inc, dtype):
    """Contiguous systems of 4097-8192 rows, one per 256-thread workgroup at
    32 rows per lane with `a` moved into the LDS tile by global->LDS loads
    (variant 4; FP32: two half-wave loads per 64 rows): lengths at both ends
    of the range and between, Inc; refused for shorter and strided systems."""
    for dims in ([4097, 3, 2], [5000, 7], [8192, 2, 2], [6001, 5]):
        with api.forced_hybrid(32, 256, 1, G, 4):
            cfg = run_case(dims, 0, dtype, inc, (0, 0, 0), seed=dims[0] + G)
        assert cfg == (13, 32, 256, 1, G)
    with api.forced_hybrid(32, 256, 1, G, 4):
        with pytest.raises(Exception):
            run_case([4096, 3], 0, dtype, inc, (0, 0, 0), seed=1)
        with pytest.raises(Exception):
            run_case([3, 5000], 1, dtype, inc, (0, 0, 0), seed=2)
        with pytest.raises(Exception):
            run_case([8193, 2], 0, dtype, inc, (0, 0, 0), seed=3)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
@pytest.mark.parametrize("G", [1, 8])
def test_forced_long32_512_contiguous(G, inc, dtype):
    """Contiguous systems of 8193-16384 rows, one per 512-thread workgroup at
    32 rows per lane (variant 4 at P = 512, 156 KiB of LDS at FP64): lengths
    at both ends of the range and between, Inc; refused for 8192 rows and
    fewer and for strided systems."""
    for dims in ([8193, 3], [10001, 2, 2], [16384, 3], [12345, 2]):
        with api.forced_hybrid(32, 512, 1, G, 4):
            cfg = run_case(dims, 0, dtype, inc, (0, 0, 0), seed=dims[0] + G)
        assert cfg == (13, 32, 512, 1, G)
    with api.forced_hybrid(32, 512, 1, G, 4):
        with pytest.raises(Exception):
            run_case([8192, 3], 0, dtype, inc, (0, 0, 0), seed=1)
        with pytest.raises(Exception):
            run_case([16385, 2], 0, dtype, inc, (0, 0, 0), seed=2)
        with pytest.raises(Exception):
            run_case([3, 9000], 1, dtype, inc, (0, 0, 0), seed=3)


@pytest.mark.parametrize("P,n0", [(256, 4500), (512, 9001)])
def test_forced_long32_padded(P, n0):
    """Variant 4 on the interior of a halo-padded allocation (row pitch and
    plane pitch not multiples of anything): the halo is untouched."""
    rng = np.random.default_rng(29)
    dims = [n0, 3, 2]
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n)
    arrs = [H.pad_halo(t, dims) for t in (a, b, c, d)]
    off, pads = H.halo_offset(dims), [t + 2 for t in dims]
    with api.forced_hybrid(32, P, 1, 8, 4):
        gd, _ = gpu_solve(*arrs, solvedim=0, dims=dims, pads=pads, offset=off)
        assert api.last_config()[0] == 13
    od = arrs[3].copy()
    oracle.batch(arrs[0], arrs[1], arrs[2], od, ndim=3, solvedim=0, dims=dims, pads=pads,
                 offset=off)
    H.assert_norm_close(gd, od, np.float64)
    mask = H.pad_halo(np.ones(n, dtype=bool), dims)
    assert np.array_equal(gd[~mask], arrs[3][~mask])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_forced_wide_padded(dtype):
    """Wide variant on the interior of a halo-padded allocation."""
    rng = np.random.default_rng(23)
    dims = [70, 22, 40]
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    arrs = [H.pad_halo(t, dims) for t in (a, b, c, d)]
    off, pads = H.halo_offset(dims), [t + 2 for t in dims]
    for sd, seg in ((1, 256), (2, 256), (1, 128), (2, 128)):
        P, W = _wide_shape(dtype, 16, dims[sd], seg)
        with api.forced_hybrid(16, P, W, 8, 3):
            gd, _ = gpu_solve(*arrs, solvedim=sd, dims=dims, pads=pads, offset=off)
            assert api.last_config()[0] == 8
        od = arrs[3].copy()
        oracle.batch(arrs[0], arrs[1], arrs[2], od, ndim=3, solvedim=sd, dims=dims,
                     pads=pads, offset=off)
        H.assert_norm_close(gd, od, dtype)
        H.ref_allclose(od, gd, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_wide_rejects_unfit_shapes(dtype):
    """Shapes the wide variant cannot run are refused before any launch:
    contiguous (x) systems, rows of other than 128 or 256 bytes, partial
    waves, too few rows."""
    from tridsolver_amd._lib import TridError
    rng = np.random.default_rng(5)
    dims = [64, 32, 32]
    a, b, c, d = H.diag_dominant(rng, int(np.prod(dims)), dtype)
    W = 32 if dtype == np.float64 else 64
    for sd, cfg in ((0, (16, 4, W, 1, 3)), (1, (16, 8, W // 4, 1, 3)),
                    (2, (16, 1, W, 1, 3))):
        with api.forced_hybrid(*cfg):
            with pytest.raises(TridError) as ei:
                gpu_solve(a, b, c, d, solvedim=sd, dims=dims)
        assert ei.value.status == 3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_edge_lengths(dtype):
    for dims in ([1], [2], [1, 9], [9, 1], [2, 3, 1], [3000], [2500, 2]):
        for sd in range(len(dims)):
            for algo in ("lane", "hybrid", "auto"):
                run_case(dims, sd, dtype, False, ALGOS[algo], seed=11)
                run_case(dims, sd, dtype, True, ALGOS[algo], seed=12)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_padded_random(dtype):
    """Interior of a halo-padded allocation; the halo must stay untouched."""
    rng = np.random.default_rng(21)
    dims = [30, 22, 14]
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    arrs = [H.pad_halo(t, dims) for t in (a, b, c, d)]
    off, pads = H.halo_offset(dims), [t + 2 for t in dims]
    for sd in range(3):
        for algo in ALGOS.values():
            gd, _ = gpu_solve(*arrs, solvedim=sd, dims=dims, pads=pads, offset=off,
                              opts=algo)
            od = arrs[3].copy()
            oracle.batch(arrs[0], arrs[1], arrs[2], od, ndim=3, solvedim=sd, dims=dims,
                         pads=pads, offset=off)
            H.assert_norm_close(gd, od, dtype)
            mask = H.pad_halo(np.ones(n, dtype=bool), dims)
            assert np.array_equal(gd[~mask], arrs[3][~mask])


def test_async_on_torch_stream():
    rng = np.random.default_rng(5)
    dims = [64, 32, 16]
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        ta, tb, tc, td = (dev(t) for t in (a, b, c, d))
        api.trid_strided_batch(ta, tb, tc, td, solvedim=1, dims=dims, sync=False)
        api.trid_strided_batch(ta, tb, tc, td, solvedim=2, dims=dims, sync=False)
    s.synchronize()
    od = d.copy()
    oracle.batch(a, b, c, od, ndim=3, solvedim=1, dims=dims)
    oracle.batch(a, b, c, od, ndim=3, solvedim=2, dims=dims)
    H.assert_norm_close(td.cpu().numpy(), od)
    api.use_torch_stream(torch.cuda.default_stream())


@pytest.mark.parametrize("solvedim", [0, 1, 2])
def test_config_64cube_fp64(solvedim):
    """64^3 FP64 along x, y, z: the GPU matches the reference CPU library
    (or the oracle where it is not built) to 1e-12; errors vs the
    long-double solve are reported."""
    rng = np.random.default_rng(64)
    dims = [64, 64, 64]
    n = 64 ** 3
    a, b, c, d = H.diag_dominant(rng, n)
    gd, _ = gpu_solve(a, b, c, d, solvedim=solvedim, dims=dims)
    ref = d.copy()
    if oracle.ref_available():
        oracle.ref_batch(a, b, c, ref, ndim=3, solvedim=solvedim, dims=dims)
    else:
        oracle.batch(a, b, c, ref, ndim=3, solvedim=solvedim, dims=dims)
    err = H.assert_norm_close(gd, ref, np.float64)
    xl = oracle.batch_ld(a, b, c, d, ndim=3, solvedim=solvedim, dims=dims)
    print("64^3 sd=%d: vs CPU %.2e, GPU vs long double %.2e, CPU vs long double %.2e"
          % (solvedim, err, H.norm_rel_err(gd, xl), H.norm_rel_err(ref, xl)))


def test_invalid_arguments_on_device():
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    with pytest.raises(Exception):
        api.trid_strided_batch(t, t, t, t, solvedim=1, dims=[8], pads=[8])
    with pytest.raises(Exception):
        api.trid_strided_batch(t, t, t, t, solvedim=0, dims=[4, 3], pads=[4, 3])
