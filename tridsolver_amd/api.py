"""Python mirror of the tridsolver C API on torch device tensors.

``trid_strided_batch`` has the argument meaning of the reference's
``trid{S,D}mtsvStridedBatch[Inc]`` (scalar/include/tridsolver.h:38-68): arrays
of extent ``pads`` (x fastest), systems along ``solvedim`` of length
``dims[solvedim]``, result in ``d`` (or added into ``u`` for Inc). Every call
goes through libtridhip.so; torch only provides device memory and the stream.
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import TridParams, check, int3

_FN = {
    (torch.float64, False): "tridDmtsvStridedBatch",
    (torch.float32, False): "tridSmtsvStridedBatch",
    (torch.float64, True): "tridDmtsvStridedBatchInc",
    (torch.float32, True): "tridSmtsvStridedBatchInc",
}


def _ptr(t, offset=0):
    return C.c_void_p(t.data_ptr() + offset * t.element_size())


def use_torch_stream(stream=None):
    """Makes this thread's library calls run on a torch stream (default: the
    current one), so they are ordered with the torch work around them."""
    s = torch.cuda.current_stream() if stream is None else stream
    check(_lib.lib().tridHipSetStream(C.c_void_p(s.cuda_stream)), "tridHipSetStream")


def make_params(opts=(0, 0, 0), sync=1):
    p = TridParams()
    p.mpi_params = None
    for i in range(3):
        p.opts[i] = int(opts[i]) if i < len(opts) else 0
    p.sync = int(sync)
    return p


def _validate(tensors, dims, pads, offset):
    t0 = tensors[0]
    if t0.dtype not in (torch.float32, torch.float64):
        raise TypeError("tridsolver: float32 or float64 tensors required")
    need = offset
    stride = 1
    for dm, pd in zip(dims, pads):
        need += (dm - 1) * stride
        stride *= pd
    for t in tensors:
        if t.dtype != t0.dtype or t.device != t0.device:
            raise TypeError("tridsolver: all arrays must share dtype and device")
        if not t.is_contiguous():
            raise ValueError("tridsolver: arrays must be contiguous")
        if t.numel() < need + 1:
            raise ValueError("tridsolver: array smaller than dims/pads/offset describe")
        if t.device.type != "cuda":
            raise ValueError("tridsolver: arrays must be on a GPU")


def trid_strided_batch(a, b, c, d, *, solvedim, dims=None, pads=None, u=None,
                       offset=0, opts=(0, 0, 0), sync=False, stream=None):
    """Solves along ``solvedim``; d is overwritten, or u += x when u is given.

    dims default to the tensor shape reversed (x fastest); pads default to the
    storage extents (the tensor shape reversed). ``offset`` is the element
    offset of the first interior point inside each array. With sync=False the
    call is enqueued on the torch stream; with sync=True it returns after the
    results are complete (the reference's behaviour).
    """
    if dims is None:
        dims = list(d.shape[::-1])
    if pads is None:
        pads = list(d.shape[::-1]) if d.dim() == len(dims) else list(dims)
    ndim = len(dims)
    inc = u is not None
    arrs = [a, b, c, d] + ([u] if inc else [])
    _validate(arrs, dims, pads, offset)
    use_torch_stream(stream)
    fn = getattr(_lib.lib(), _FN[(d.dtype, inc)])
    params = make_params(opts, 1 if sync else 0)
    ptrs = [_ptr(t, offset) for t in arrs]
    st = fn(C.byref(params), *ptrs, ndim, solvedim, int3(dims), int3(pads))
    check(st, _FN[(d.dtype, inc)])
    return u if inc else d


def last_config():
    """(algorithm, M, P, W, G) of this thread's last solve (1 lane, 3 hybrid)."""
    v = [C.c_int(0) for _ in range(5)]
    _lib.lib().tridHipLastConfig(*[C.byref(x) for x in v])
    return tuple(x.value for x in v)


class forced_hybrid:
    """Context manager forcing the hybrid shape (M rows/lane, P lanes/system,
    W systems/workgroup, G XCD grouping) for this thread's solves."""

    def __init__(self, M, P, W, G=8):
        self.cfg = (M, P, W, G)

    def __enter__(self):
        check(_lib.lib().tridHipForceHybridConfig(*self.cfg), "tridHipForceHybridConfig")
        return self

    def __exit__(self, *exc):
        _lib.lib().tridHipForceHybridConfig(0, 0, 0, 0)
        return False
