"""Systems longer than one workgroup holds (above 16384 contiguous / 2048
strided FP64 rows): the hybrid with its reduced system solved by a second
batched solve (kernels_long.hip), against the oracle on the same inputs --
normwise <= 1e-12 FP64 / 1e-5 FP32 and the reference tests' elementwise
formula, with a[0] and c[N-1] planted at +-1e3 -- along x, y and z, plain and
Inc, halo-padded, at lengths that need a second level of reduction, and in a
hipGraph capture with the workspace reserved beforehand. The reference's
Thomas paths stop at N_MAX = 1024 rows (scalar/CMakeLists.txt:58-61); its
per-system semantics are those of trid_scalar (scalar/src/cpu/trid_cpu.cpp:
312-353)."""
import numpy as np
import pytest

import oracle
from tests import helpers as H
from tests.test_gpu_solve import run_case

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import _lib, api  # noqa: E402

LONG = 12  # tridHipLastConfig algorithm of the long-system path


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


# (dims, solvedim): around the longest system one workgroup holds, several
# segments, and lengths whose reduced systems are themselves too long (a
# second level: more than 16384 reduced rows, 33 segments of 8192 rows)
CASES = [([8193, 3], 0), ([16385, 3], 0), ([20000, 5, 2], 0), ([140000, 2], 0),
         ([270000, 2], 0),
         ([40, 2049, 3], 1), ([17, 5000, 2], 1), ([8, 40000], 1),
         ([20, 3, 3000], 2), ([33, 2, 9000], 2)]


def _expected_algo(n, sd, fp64):
    """tridHipLastConfig's algorithm for an untuned solve of n-row systems:
    contiguous, one workgroup holds 16384 rows (FP64 above 4096 rows at 32
    rows per lane, algorithm 13; FP32 at 16 rows per lane on up to 1024
    lanes, 3); strided, 16 rows per lane on 128 (FP64) or 256 (FP32) lanes."""
    if sd == 0:
        if n > 16384:
            return LONG
        return 13 if fp64 and n > 4096 else 3
    return LONG if n > (2048 if fp64 else 4096) else 3


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("inc", [False, True])
@pytest.mark.parametrize("dims,sd", CASES)
def test_long_vs_oracle(dims, sd, inc, dtype):
    cfg = run_case(dims, sd, dtype, inc, (0, 0, 0), seed=sum(dims) + sd)
    assert cfg[0] == _expected_algo(dims[sd], sd, dtype == np.float64), cfg


@pytest.mark.parametrize("sd", [0, 1, 2])
def test_long_padded(sd):
    """Halo-padded arrays, pointers offset by one halo layer: the padding is
    untouched and the interior matches the oracle."""
    dims = [[17000, 3, 2], [5, 3000, 2], [5, 2, 3000]][sd]
    rng = np.random.default_rng(90 + sd)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n)
    arrs = [H.pad_halo(t, dims) for t in (a, b, c, d)]
    off, pads = H.halo_offset(dims), [t + 2 for t in dims]
    td = torch.from_numpy(arrs[3].copy()).cuda()
    ta, tb, tc = (torch.from_numpy(t).cuda() for t in arrs[:3])
    api.trid_strided_batch(ta, tb, tc, td, solvedim=sd, dims=dims, pads=pads, offset=off,
                           sync=True)
    assert api.last_config()[0] == LONG
    od = arrs[3].copy()
    oracle.batch(arrs[0], arrs[1], arrs[2], od, ndim=3, solvedim=sd, dims=dims, pads=pads,
                 offset=off)
    got = td.cpu().numpy()
    H.assert_norm_close(got, od, np.float64)
    mask = H.pad_halo(np.ones(n, dtype=bool), dims)
    assert np.array_equal(got[~mask], arrs[3][~mask])


def test_long_opts_keep_the_lane_path():
    """The reference's per-lane options (opts 1 and 2) still select the
    per-lane Thomas kernel for long systems."""
    cfg = run_case([9000, 3], 0, np.float64, False, (1, 0, 0), seed=7)
    assert cfg[0] == 1


@pytest.mark.parametrize("sd", [0, 1])
def test_long_graph_replay(sd):
    """Captured once after reserving tridHipWorkspaceBytes for the capture
    stream; replays with eager work in between equal the eager solve."""
    dims = [[20000, 4], [16, 6000]][sd]
    rng = np.random.default_rng(60 + sd)
    n = int(np.prod(dims))
    a, b, c, d = (torch.from_numpy(t).cuda() for t in H.diag_dominant(rng, n))
    eager = d.clone()
    api.trid_strided_batch(a, b, c, eager, solvedim=sd, dims=dims, sync=True)
    need = _lib.lib().tridHipWorkspaceBytes(1, 0, len(dims), sd, _lib.int3(dims))
    assert need > 0
    s = torch.cuda.Stream()
    work = torch.empty_like(d)
    with torch.cuda.stream(s):
        api.reserve_workspace(need, stream=s)
        work.copy_(d)
        api.trid_strided_batch(a, b, c, work, solvedim=sd, dims=dims, stream=s)  # warm-up
    s.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        api.trid_strided_batch(a, b, c, work, solvedim=sd, dims=dims, stream=s)
    try:
        for k in range(3):
            work.copy_(d)
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(work, eager), k
            d.mul_(1.0)  # eager work between replays
    finally:
        api.use_torch_stream(torch.cuda.default_stream())
