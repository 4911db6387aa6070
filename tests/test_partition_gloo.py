"""The partitioned solve's exchange over torch.distributed (gloo, CPU, 3
ranks): each rank holds its shard of a batch of systems, computes its
boundary records (the numpy restatement of the per-shard step in
test_partition_math.py), gathers every rank's block shard-major with
all_gather -- the allgather a multi-GPU run makes over RCCL -- and finishes
its rows. Rank 0 compares the reassembled batch with dense solves."""
import os
import socket

import numpy as np
import pytest

torch = pytest.importorskip("torch")
dist = pytest.importorskip("torch.distributed")
mp = pytest.importorskip("torch.multiprocessing")

from tests.test_partition_math import boundary, dense, system, twisted  # noqa: E402

SIZES = [5, 2, 9]  # rows of each rank's shard
NSYS = 7


def _free_port():
    s = socket.socket()
    s.bind(("127.0.0.1", 0))
    port = s.getsockname()[1]
    s.close()
    return port


def _rank(rank, world, port, out):
    os.environ["MASTER_ADDR"] = "127.0.0.1"
    os.environ["MASTER_PORT"] = str(port)
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        rng = np.random.default_rng(99)
        systems = [system(rng, sum(SIZES)) for _ in range(NSYS)]  # same on every rank
        o = sum(SIZES[:rank])
        sl = slice(o, o + SIZES[rank])
        mine = [tuple(v[sl].copy() for v in s) for s in systems]
        # step 1: 6 reals per system, laid out [6][nsys]
        rec = np.array([boundary(*m, rank == 0, rank == world - 1) for m in mine]).T
        block = torch.from_numpy(np.ascontiguousarray(rec).reshape(-1))
        # step 2: the exchange, shard-major
        parts = [torch.empty_like(block) for _ in range(world)]
        dist.all_gather(parts, block)
        all_ = torch.cat(parts).numpy().reshape(world, 6, NSYS)
        # step 3: E_{k-1}, F_{k+1} from the reduced rows, then the shard's solve
        xs = []
        for s, (a, b, c, d) in enumerate(mine):
            sub = np.ravel(all_[:, [0, 3], s])
            sup = np.ravel(all_[:, [1, 4], s])
            rhs = np.ravel(all_[:, [2, 5], s])
            d = d.copy()
            if rank > 0:
                d[0] -= a[0] * twisted(sub, sup, rhs, 2 * rank - 1)
            if rank < world - 1:
                d[-1] -= c[-1] * twisted(sub, sup, rhs, 2 * rank + 2)
            a, c = a.copy(), c.copy()
            a[0], c[-1] = 0.0, 0.0
            xs.append(dense(a, b, c, d))
        # collect the results on rank 0 (padded: gloo gathers equal sizes)
        mine_x = torch.zeros(max(SIZES) * NSYS, dtype=torch.float64)
        mine_x[:SIZES[rank] * NSYS] = torch.from_numpy(np.concatenate(xs))
        got = [torch.empty_like(mine_x) for _ in range(world)]
        dist.all_gather(got, mine_x)
        got = [g[:SIZES[r] * NSYS] for r, g in enumerate(got)]
        if rank == 0:
            err = 0.0
            for s, (a, b, c, d) in enumerate(systems):
                x = np.concatenate([got[r].numpy().reshape(NSYS, SIZES[r])[s]
                                    for r in range(world)])
                a, c = a.copy(), c.copy()
                a[0], c[-1] = 0.0, 0.0
                want = dense(a, b, c, d)
                err = max(err, np.abs(x - want).max() / np.abs(want).max())
            out.put(err)
    finally:
        dist.destroy_process_group()


def test_exchange_over_gloo():
    world = len(SIZES)
    ctx = mp.get_context("spawn")
    out = ctx.SimpleQueue()
    port = _free_port()
    procs = [ctx.Process(target=_rank, args=(r, world, port, out)) for r in range(world)]
    for p in procs:
        p.start()
    for p in procs:
        p.join(120)
    assert all(p.exitcode == 0 for p in procs), [p.exitcode for p in procs]
    assert out.get() <= 1e-13
