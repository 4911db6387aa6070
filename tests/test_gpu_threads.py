"""Host threads calling the library at once, each on its own stream.

The library keeps the stream, the forced shapes and the last error per
calling thread, and guards its shared state (device checks, workspaces,
tuning cache, dynamic-LDS grants) with locks or atomics. Five threads solve
different batches concurrently -- hybrid solves along x, y and z, the
per-lane path (which uses a per-stream workspace), ragged batches -- and one
of them forces a hybrid shape; every result must match the oracle, and the
forced shape must not leak into the other threads.
"""
import threading

import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import api  # noqa: E402


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _solve_case(seed, dims, sd, dtype):
    rng = np.random.default_rng(seed)
    a, b, c, d = H.diag_dominant(rng, int(np.prod(dims)), dtype)
    want = d.copy()
    oracle.batch(a, b, c, want, ndim=len(dims), solvedim=sd, dims=dims)
    return (a, b, c, d), want


def _ragged_case(seed, dtype):
    rng = np.random.default_rng(seed)
    lens = np.round(np.exp(rng.uniform(0, np.log(2048), 400))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]) + 1, dtype)
    want = d.copy()
    oracle.ragged(a, b, c, want, off)
    return (a, b, c, d, off), want


def test_threads_on_own_streams():
    jobs = [
        ("solve", _solve_case(1, [96, 80, 64], 0, np.float64), dict(sd=0, dims=[96, 80, 64])),
        ("solve", _solve_case(2, [64, 72, 90], 2, np.float32), dict(sd=2, dims=[64, 72, 90])),
        ("lane", _solve_case(3, [40, 300, 20], 1, np.float64), dict(sd=1, dims=[40, 300, 20])),
        ("ragged", _ragged_case(4, np.float64), {}),
        ("forced", _solve_case(5, [256, 30, 20], 0, np.float64), dict(sd=0, dims=[256, 30, 20])),
    ]
    results, shapes, errors = {}, {}, []
    start = threading.Barrier(len(jobs))

    def work(k, kind, case, kw):
        try:
            torch.cuda.set_device(0)
            arrays, _ = case
            dev = [torch.from_numpy(np.ascontiguousarray(v)).cuda() for v in arrays]
            stream = torch.cuda.Stream()
            with torch.cuda.stream(stream):
                start.wait()
                for _ in range(4):  # repeated calls interleave with the other threads'
                    td = dev[3].clone()
                    if kind == "ragged":
                        api.ragged_batch(dev[0], dev[1], dev[2], td, dev[4], max_len=2048,
                                         stream=stream)
                    elif kind == "lane":  # hint 1: the per-lane path and its workspace
                        api.trid_strided_batch(*dev[:3], td, solvedim=kw["sd"], dims=kw["dims"],
                                               opts=(1, 1, 1), stream=stream)
                    elif kind == "forced":
                        with api.forced_hybrid(4, 64, 4, 8, 0):
                            api.trid_strided_batch(*dev[:3], td, solvedim=kw["sd"],
                                                   dims=kw["dims"], stream=stream)
                    else:
                        api.trid_strided_batch(*dev[:3], td, solvedim=kw["sd"], dims=kw["dims"],
                                               stream=stream)
                    shapes.setdefault(k, []).append(api.last_config())
                stream.synchronize()
            results[k] = td.cpu().numpy()
        except Exception as e:  # reported by the main thread
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=work, args=(k, *job)) for k, job in enumerate(jobs)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads), "a thread did not finish"
    assert not errors, errors
    for k, (kind, case, _) in enumerate(jobs):
        arrays, want = case
        dtype = arrays[3].dtype
        H.assert_norm_close(results[k], want, dtype)
        algos = {cfg[0] for cfg in shapes[k]}
        if kind == "lane":
            assert algos == {1}, shapes[k]
        elif kind == "ragged":
            assert algos == {5}, shapes[k]
        elif kind == "forced":
            assert {cfg[:4] for cfg in shapes[k]} == {(3, 4, 64, 4)}, shapes[k]
        else:  # the other thread's forced shape is not seen here
            assert all(cfg[1:4] != (4, 64, 4) for cfg in shapes[k]), shapes[k]
