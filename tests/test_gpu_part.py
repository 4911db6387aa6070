"""Systems split across shards (tridXmtsvStridedBatchPartBoundary /
[Inc]PartSolve), the shards emulated in one process on one GPU: each shard
gets its own arrays holding its rows of every system; the gathered boundary
blocks stand in for the allgather of a distributed run. The reassembled
solution is compared with the oracle's solve of the whole systems
(normwise 1e-12 FP64, 1e-5 FP32, and the reference tests' elementwise
formula)."""
import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import api  # noqa: E402
from tridsolver_amd._lib import TridError  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def split(arr, dims, sd, sizes):
    """The shards' own arrays: rows [o_k, o_k + n_k) along sd, contiguous."""
    grid = arr.reshape(dims[::-1])  # numpy order: last dim is x
    axis = len(dims) - 1 - sd
    out, o = [], 0
    for n in sizes:
        sl = [slice(None)] * len(dims)
        sl[axis] = slice(o, o + n)
        out.append(np.ascontiguousarray(grid[tuple(sl)]).reshape(-1))
        o += n
    return out


def join(parts, dims, sd, sizes):
    axis = len(dims) - 1 - sd
    shaped = []
    for p, n in zip(parts, sizes):
        dk = list(dims)
        dk[sd] = n
        shaped.append(p.reshape(dk[::-1]))
    return np.concatenate(shaped, axis=axis).reshape(-1)


def run_partitioned(a, b, c, d, dims, sd, sizes, u=None, dev_dtype=None):
    G = len(sizes)
    pa, pb, pc, pd = (split(t, dims, sd, sizes) for t in (a, b, c, d))
    pu = split(u, dims, sd, sizes) if u is not None else [None] * G
    nsys = int(np.prod([dm for i, dm in enumerate(dims) if i != sd]))
    dt = torch.float64 if a.dtype == np.float64 else torch.float32
    tens = []
    bnds = []
    for k in range(G):
        dk = list(dims)
        dk[sd] = sizes[k]
        t = [torch.from_numpy(x).cuda() for x in (pa[k], pb[k], pc[k], pd[k])]
        tu = torch.from_numpy(pu[k]).cuda() if u is not None else None
        bnd = torch.empty(6 * nsys, dtype=dt, device="cuda")
        api.part_boundary(*t, bnd, solvedim=sd, shard=k, nshards=G, dims=dk, sync=True)
        tens.append((t, tu, dk))
        bnds.append(bnd)
    all_ = torch.cat(bnds)  # the allgather, shard-major
    outs, d_after = [], []
    for k in range(G):
        t, tu, dk = tens[k]
        api.part_solve(*t, all_, solvedim=sd, shard=k, nshards=G, dims=dk, u=tu, sync=True)
        outs.append((tu if tu is not None else t[3]).cpu().numpy())
        d_after.append(t[3].cpu().numpy())
    return join(outs, dims, sd, sizes), join(d_after, dims, sd, sizes)


def check(got, want, dtype):
    H.assert_norm_close(got, want, dtype, tol=1e-12 if dtype == np.float64 else 1e-5)
    H.ref_allclose(want, got, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("sd,dims,sizes", [
    (0, [60, 7, 5], [20, 20, 20]),
    (0, [45, 9], [2, 30, 13]),
    (1, [33, 50, 6], [7, 2, 16, 25]),
    (2, [40, 12, 64], [8] * 8),
    (2, [17, 5, 300], [100, 3, 197]),
    (1, [70, 9], [5, 4]),
    (2, [12, 10, 30], [30]),
])
def test_partitioned_vs_whole_solve(dtype, sd, dims, sizes):
    rng = np.random.default_rng(sum(sizes) + sd)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    for idx in H.system_slices(len(dims), sd, dims):
        a[idx[0]] = 1e3  # never used by the whole-system solve
        c[idx[-1]] = -1e3
    got, _ = run_partitioned(a, b, c, d, dims, sd, sizes)
    want = d.copy()
    oracle.batch(a, b, c, want, ndim=len(dims), solvedim=sd, dims=dims)
    check(got, want, dtype)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_partitioned_inc_leaves_d(dtype):
    """Inc: u += x, and every shard's d comes back bit for bit."""
    rng = np.random.default_rng(8)
    dims, sd, sizes = [24, 20, 48], 2, [10, 2, 36]
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, dtype)
    u = rng.uniform(-1, 1, n).astype(dtype)
    got, d_after = run_partitioned(a, b, c, d, dims, sd, sizes, u=u)
    assert np.array_equal(d_after, d)
    ou = u.copy()
    od = d.copy()
    oracle.batch(a, b, c, od, ou, ndim=3, solvedim=sd, dims=dims, inc=True)
    check(got - u, ou - u, dtype)


def test_partitioned_512_slab_shards():
    """z solves of a 128 x 128 x 512 FP64 grid held as 8 z-slabs of 64 planes
    (the share of each of 8 GPUs); against the whole-grid GPU solve."""
    dims, sd, G = [128, 128, 512], 2, 8
    rng = np.random.default_rng(512)
    n = int(np.prod(dims))
    a, b, c, d = H.diag_dominant(rng, n, np.float64)
    got, _ = run_partitioned(a, b, c, d, dims, sd, [64] * G)
    td = torch.from_numpy(d.copy()).cuda()
    api.trid_strided_batch(*(torch.from_numpy(x).cuda() for x in (a, b, c)), td, solvedim=sd,
                           dims=dims, sync=True)
    H.assert_norm_close(got, td.cpu().numpy(), np.float64, tol=1e-12)


def test_partitioned_rejects():
    rng = np.random.default_rng(2)
    dims = [8, 6, 1]
    a, b, c, d = (torch.from_numpy(x).cuda() for x in H.diag_dominant(rng, 48, np.float64))
    bnd = torch.empty(6 * 48, dtype=torch.float64, device="cuda")
    with pytest.raises(TridError):  # a one-row shard
        api.part_boundary(a, b, c, d, bnd, solvedim=2, shard=0, nshards=2, dims=dims)
    with pytest.raises(TridError):  # shard out of range
        api.part_boundary(a, b, c, d, bnd, solvedim=1, shard=2, nshards=2, dims=dims)
    with pytest.raises(TridError):
        api.part_solve(a, b, c, d, torch.zeros(2 * 6 * 8, dtype=torch.float64, device="cuda"),
                       solvedim=1, shard=0, nshards=0, dims=dims)
