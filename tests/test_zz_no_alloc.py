"""tridDmtsvRaggedBatchNoAlloc (the ragged FP64 solve that allocates
nothing: its plan lives in the library's fixed per-device scratch array)
against tridDmtsvRaggedBatch, through the C ABI: bit-identical results on two
calls in a row, the second with a smaller batch; no device memory taken by the
call (free memory unchanged by a call on a stream that has no workspace);
capturable in a hipGraph without reserving anything; batches beyond the
scratch's stated maximum rejected before any launch."""
import ctypes as C

import numpy as np
import pytest

import oracle
from tests import helpers as H

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

from tridsolver_amd import _lib, api  # noqa: E402

MAX_SYSTEMS = 1 << 20  # TRID_HIP_NOALLOC_MAX_SYSTEMS in include/tridsolver_hip.h


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    torch.cuda.set_device(0)


def _case(seed, nsys, max_len):
    rng = np.random.default_rng(seed)
    lens = np.round(np.exp(rng.uniform(np.log(1), np.log(max_len), nsys))).astype(np.int64)
    lens[:4] = [0, 1, 2, max_len]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    a, b, c, d = H.diag_dominant(rng, int(off[-1]))
    return off, a, b, c, d


def _dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).cuda()


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _solve(name, ta, tb, tc, td, toff, nsys, max_len, sync=1):
    p = api.make_params((0, 0, 0), sync)
    return getattr(_lib.lib(), name)(C.byref(p), _ptr(ta), _ptr(tb), _ptr(tc), _ptr(td),
                                     _ptr(toff), nsys, max_len)


def test_no_alloc_matches_ragged_batch_twice():
    api.use_torch_stream()
    for k, (seed, nsys, max_len) in enumerate([(71, 3000, 8192), (72, 700, 2048)]):
        off, a, b, c, d = _case(seed, nsys, max_len)
        ta, tb, tc, toff = _dev(a), _dev(b), _dev(c), _dev(off)
        want, got = _dev(d), _dev(d)
        assert _solve("tridDmtsvRaggedBatch", ta, tb, tc, want, toff, nsys, max_len) == 0
        assert _solve("tridDmtsvRaggedBatchNoAlloc", ta, tb, tc, got, toff, nsys, max_len) == 0
        torch.cuda.synchronize()
        assert torch.equal(got, want), k
        # and against the oracle, on the first 300 systems
        m = 300
        ref = d[:off[m]].copy()
        oracle.ragged(a[:off[m]], b[:off[m]], c[:off[m]], ref, off[:m + 1])
        H.assert_norm_close(got[:off[m]].cpu().numpy(), ref, np.float64)


def test_no_alloc_takes_no_device_memory_and_captures():
    off, a, b, c, d = _case(73, 2000, 4096)
    ta, tb, tc, toff = _dev(a), _dev(b), _dev(c), _dev(off)
    want = _dev(d)
    api.use_torch_stream()
    assert _solve("tridDmtsvRaggedBatch", ta, tb, tc, want, toff, 2000, 4096) == 0
    s = torch.cuda.Stream()  # a stream the library has no workspace for
    work = _dev(d)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    with torch.cuda.stream(s):
        api.use_torch_stream(s)
        assert _solve("tridDmtsvRaggedBatchNoAlloc", ta, tb, tc, work, toff, 2000, 4096) == 0
    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    # the workspace path would have taken at least 64 MiB here
    assert free0 - free1 < (16 << 20), (free0, free1)
    assert torch.equal(work, want)
    # captured without reserving a workspace for the capture stream
    g = torch.cuda.CUDAGraph()
    s2 = torch.cuda.Stream()
    with torch.cuda.graph(g, stream=s2):
        api.use_torch_stream(s2)
        assert _solve("tridDmtsvRaggedBatchNoAlloc", ta, tb, tc, work, toff, 2000, 4096,
                      sync=0) == 0
    try:
        for k in range(3):
            work.copy_(_dev(d))
            torch.cuda.synchronize()
            g.replay()
            torch.cuda.synchronize()
            assert torch.equal(work, want), k
    finally:
        api.use_torch_stream(torch.cuda.default_stream())


def test_no_alloc_rejects_too_many_systems():
    api.use_torch_stream()
    t = torch.zeros(8, dtype=torch.float64, device="cuda")
    toff = torch.zeros(8, dtype=torch.int64, device="cuda")
    # rejected on the host, before any launch: nothing reads past these arrays
    st = _solve("tridDmtsvRaggedBatchNoAlloc", t, t, t, t, toff, MAX_SYSTEMS + 1, 8)
    assert _lib.STATUS[st] == "TRID_STATUS_INVALID_VALUE"
    st = _solve("tridDmtsvRaggedBatchNoAlloc", t, t, t, t, toff, 7, 9000)  # longer than 8192
    assert _lib.STATUS[st] == "TRID_STATUS_INVALID_VALUE"
