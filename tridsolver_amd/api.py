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


def part_boundary(a, b, c, d, bnd, *, solvedim, shard, nshards, dims=None, pads=None,
                  offset=0, opts=(0, 0, 0), sync=False, stream=None):
    """Step 1 of a solve whose systems are split across shards
    (tridXmtsvStridedBatchPartBoundary): 6 reals per system into bnd
    ([6][nsys]). This shard's arrays hold its rows of every system."""
    if dims is None:
        dims = list(d.shape[::-1])
    if pads is None:
        pads = list(d.shape[::-1]) if d.dim() == len(dims) else list(dims)
    _validate([a, b, c, d], dims, pads, offset)
    nsys = 1
    for i, dm in enumerate(dims):
        if i != solvedim:
            nsys *= dm
    if bnd.dtype != d.dtype or bnd.device != d.device or bnd.numel() < 6 * nsys:
        raise ValueError("bnd must hold 6 * nsys values of the arrays' dtype and device")
    use_torch_stream(stream)
    name = ("tridD" if d.dtype == torch.float64 else "tridS") + "mtsvStridedBatchPartBoundary"
    params = make_params(opts, 1 if sync else 0)
    st = getattr(_lib.lib(), name)(C.byref(params), *[_ptr(t, offset) for t in (a, b, c, d)],
                                   len(dims), solvedim, int3(dims), int3(pads), shard, nshards,
                                   _ptr(bnd))
    check(st, name)
    return bnd


def part_solve(a, b, c, d, bnd_all, *, solvedim, shard, nshards, dims=None, pads=None,
               u=None, offset=0, opts=(0, 0, 0), sync=False, stream=None):
    """Step 3 (tridXmtsvStridedBatch[Inc]PartSolve): this shard's rows of the
    solution, given every shard's boundary block gathered shard-major."""
    if dims is None:
        dims = list(d.shape[::-1])
    if pads is None:
        pads = list(d.shape[::-1]) if d.dim() == len(dims) else list(dims)
    inc = u is not None
    arrs = [a, b, c, d] + ([u] if inc else [])
    _validate(arrs, dims, pads, offset)
    nsys = 1
    for i, dm in enumerate(dims):
        if i != solvedim:
            nsys *= dm
    if (bnd_all.dtype != d.dtype or bnd_all.device != d.device
            or bnd_all.numel() < 6 * nsys * nshards):
        raise ValueError("bnd_all must hold nshards * 6 * nsys values")
    use_torch_stream(stream)
    name = (("tridD" if d.dtype == torch.float64 else "tridS")
            + ("mtsvStridedBatchIncPartSolve" if inc else "mtsvStridedBatchPartSolve"))
    params = make_params(opts, 1 if sync else 0)
    st = getattr(_lib.lib(), name)(C.byref(params), *[_ptr(t, offset) for t in arrs],
                                   len(dims), solvedim, int3(dims), int3(pads), shard, nshards,
                                   _ptr(bnd_all))
    check(st, name)
    return u if inc else d


def reserve_workspace(nbytes, stream=None):
    """Grows the library's workspace for (device, stream) to at least nbytes
    (tridHipReserveWorkspace): done before a hipGraph capture, which cannot
    allocate. Binds this thread's library calls to `stream` (default: the
    current torch stream)."""
    use_torch_stream(stream)
    check(_lib.lib().tridHipReserveWorkspace(int(nbytes)), "tridHipReserveWorkspace")


def ragged_workspace_bytes(nsys):
    """Workspace bytes a ragged batch of nsys systems plans in."""
    return int(_lib.lib().tridHipRaggedWorkspaceBytes(int(nsys)))


def last_config():
    """(algorithm, M, P, W, G) of this thread's last solve (1 lane, 3 hybrid)."""
    v = [C.c_int(0) for _ in range(5)]
    _lib.lib().tridHipLastConfig(*[C.byref(x) for x in v])
    return tuple(x.value for x in v)


class forced_hybrid:
    """Context manager forcing the hybrid shape (M rows/lane, P lanes/system,
    W systems/workgroup, G XCD grouping) for this thread's solves."""

    def __init__(self, M, P, W, G=8, variant=0):
        self.cfg = (M, P, W, G, variant)

    def __enter__(self):
        check(_lib.lib().tridHipForceHybridConfig(*self.cfg), "tridHipForceHybridConfig")
        return self

    def __exit__(self, *exc):
        _lib.lib().tridHipForceHybridConfig(0, 0, 0, 0, 0)
        return False


def mixed_strided_batch(a, b, c, d, *, solvedim, dims=None, pads=None, offset=0,
                        refinements=2, sync=False, stream=None):
    """FP64 solve by FP32 sweeps plus FP64 refinement (tridDmtsvStridedBatchMixed);
    d (float64) is overwritten with the solution."""
    if d.dtype != torch.float64:
        raise TypeError("mixed precision solve takes float64 operands")
    dims = list(d.shape[::-1]) if dims is None else dims
    pads = (list(d.shape[::-1]) if d.dim() == len(dims) else list(dims)) if pads is None else pads
    _validate([a, b, c, d], dims, pads, offset)
    use_torch_stream(stream)
    params = make_params((0, 0, 0), 1 if sync else 0)
    st = _lib.lib().tridDmtsvStridedBatchMixed(
        C.byref(params), *[_ptr(t, offset) for t in (a, b, c, d)], len(dims), solvedim,
        int3(dims), int3(pads), int(refinements))
    check(st, "tridDmtsvStridedBatchMixed")
    return d


def mixed_f32_strided_batch(a, b, c, d, x=None, *, solvedim, dims=None, pads=None, offset=0,
                            refinements=1, sync=False, stream=None):
    """FP64 solution of the system stored in FP32 (tridDmtsvStridedBatchMixedF32):
    a, b, c, d float32, x float64 of the same shape (allocated when None);
    returns x. d is not written."""
    if d.dtype != torch.float32:
        raise TypeError("the FP32-stored solve takes float32 operands")
    x = torch.empty(d.shape, dtype=torch.float64, device=d.device) if x is None else x
    if x.dtype != torch.float64 or x.numel() != d.numel() or x.device != d.device:
        raise TypeError("x must be a float64 tensor of d's size on d's device")
    dims = list(d.shape[::-1]) if dims is None else dims
    pads = (list(d.shape[::-1]) if d.dim() == len(dims) else list(dims)) if pads is None else pads
    _validate([a, b, c, d], dims, pads, offset)
    _validate([x], dims, pads, offset)
    use_torch_stream(stream)
    params = make_params((0, 0, 0), 1 if sync else 0)
    st = _lib.lib().tridDmtsvStridedBatchMixedF32(
        C.byref(params), *[_ptr(t, offset) for t in (a, b, c, d, x)], len(dims), solvedim,
        int3(dims), int3(pads), int(refinements))
    check(st, "tridDmtsvStridedBatchMixedF32")
    return x


def adjoint_strided_batch(a, b, c, x, w, *, solvedim, dims=None, pads=None, offset=0,
                          lam=None, grad_a=None, grad_b=None, grad_c=None, sync=False,
                          stream=None):
    """Adjoint of the solve x = A^-1 d for the incoming gradient w = dL/dx.

    Returns (lam, grad_a, grad_b, grad_c); lam = dL/dd is always computed,
    each gradient only if its output tensor is given (or True to allocate).
    x (the forward solution) is read only for those gradients and may be None
    when none is requested.
    """
    dims = list(w.shape[::-1]) if dims is None else dims
    pads = (list(w.shape[::-1]) if w.dim() == len(dims) else list(dims)) if pads is None else pads
    lam = torch.empty_like(w) if lam is None else lam
    outs = []
    for g in (grad_a, grad_b, grad_c):
        outs.append(torch.empty_like(w) if g is True else g)
    if x is None and any(g is not None for g in outs):
        raise ValueError("the coefficient gradients need the forward solution x")
    arrs = [t for t in (a, b, c, x, w, lam) if t is not None] + [g for g in outs if g is not None]
    _validate(arrs, dims, pads, offset)
    use_torch_stream(stream)
    params = make_params((0, 0, 0), 1 if sync else 0)
    name = "tridDmtsvStridedBatchAdjoint" if w.dtype == torch.float64 else \
        "tridSmtsvStridedBatchAdjoint"
    ptr = lambda t: None if t is None else _ptr(t, offset)
    st = getattr(_lib.lib(), name)(C.byref(params), ptr(a), ptr(b), ptr(c), ptr(x), ptr(w),
                                   ptr(lam), *[ptr(g) for g in outs], len(dims), solvedim,
                                   int3(dims), int3(pads))
    check(st, name)
    return (lam, *outs)


def ragged_batch(a, b, c, d, offsets, *, max_len, u=None, sync=False, stream=None):
    """Ragged batch (tridXmtsvRaggedBatch[Inc]): systems back to back,
    offsets (int64, device, nsys+1) delimit them."""
    if offsets.dtype != torch.int64 or offsets.device != d.device:
        raise TypeError("offsets must be an int64 tensor on the arrays' device")
    arrs = [a, b, c, d] + ([u] if u is not None else [])
    _validate(arrs, [d.numel()], [d.numel()], 0)
    use_torch_stream(stream)
    params = make_params((0, 0, 0), 1 if sync else 0)
    pre = "tridD" if d.dtype == torch.float64 else "tridS"
    name = pre + ("mtsvRaggedBatchInc" if u is not None else "mtsvRaggedBatch")
    args = [_ptr(t) for t in arrs] + [_ptr(offsets)]
    st = getattr(_lib.lib(), name)(C.byref(params), *args, offsets.numel() - 1, int(max_len))
    check(st, name)
    return u if u is not None else d


def ragged_batch_mixed(a, b, c, d, offsets, *, max_len, refinements=2, sync=False,
                       stream=None):
    """FP64 ragged batch by FP32 sweeps plus FP64 refinement
    (tridDmtsvRaggedBatchMixed); lengths up to 8192."""
    if d.dtype != torch.float64:
        raise TypeError("the mixed-precision ragged solve takes FP64 arrays")
    if offsets.dtype != torch.int64 or offsets.device != d.device:
        raise TypeError("offsets must be an int64 tensor on the arrays' device")
    arrs = [a, b, c, d]
    _validate(arrs, [d.numel()], [d.numel()], 0)
    use_torch_stream(stream)
    params = make_params((0, 0, 0), 1 if sync else 0)
    args = [_ptr(t) for t in arrs] + [_ptr(offsets)]
    st = _lib.lib().tridDmtsvRaggedBatchMixed(C.byref(params), *args, offsets.numel() - 1,
                                              int(max_len), int(refinements))
    check(st, "tridDmtsvRaggedBatchMixed")
    return d


def ragged_batch_mixed_f32(a, b, c, d, offsets, x=None, *, max_len, refinements=1, sync=False,
                           stream=None):
    """FP64 solution of a ragged batch stored in FP32 (tridDmtsvRaggedBatchMixedF32):
    a, b, c, d float32, x float64 of d's size (allocated when None); returns
    x. Lengths up to 8192."""
    if d.dtype != torch.float32:
        raise TypeError("the FP32-stored ragged solve takes float32 operands")
    if offsets.dtype != torch.int64 or offsets.device != d.device:
        raise TypeError("offsets must be an int64 tensor on the arrays' device")
    x = torch.empty(d.shape, dtype=torch.float64, device=d.device) if x is None else x
    if x.dtype != torch.float64 or x.numel() != d.numel() or x.device != d.device:
        raise TypeError("x must be a float64 tensor of d's size on d's device")
    _validate([a, b, c, d], [d.numel()], [d.numel()], 0)
    _validate([x], [d.numel()], [d.numel()], 0)
    use_torch_stream(stream)
    params = make_params((0, 0, 0), 1 if sync else 0)
    args = [_ptr(t) for t in (a, b, c, d, x)] + [_ptr(offsets)]
    st = _lib.lib().tridDmtsvRaggedBatchMixedF32(C.byref(params), *args, offsets.numel() - 1,
                                                 int(max_len), int(refinements))
    check(st, "tridDmtsvRaggedBatchMixedF32")
    return x


class forced_ragged:
    """Context manager forcing the ragged kernel's rows per lane, workgroups
    per CU and solve launches (1; 2 with the short systems apart; 3, at M = 16:
    one launch taking the systems over 4096 rows at 32 rows per lane, then the
    rest) for this thread's calls."""

    def __init__(self, M, per_cu=1, launches=1):
        self.cfg = (M, per_cu, launches)

    def __enter__(self):
        check(_lib.lib().tridHipForceRaggedConfig(*self.cfg), "tridHipForceRaggedConfig")
        return self

    def __exit__(self, *exc):
        _lib.lib().tridHipForceRaggedConfig(0, 1, 1)
        return False


def autotune(dtype, solvedim, dims, pads=None, inc=False):
    """Tunes the solve for one geometry (scratch data); returns the best time
    in ms. Later solves of that geometry use the winning shape."""
    pads = dims if pads is None else pads
    ms = C.c_float(0)
    st = _lib.lib().tridHipAutotune(1 if dtype == torch.float64 else 0, 1 if inc else 0,
                                    len(dims), solvedim, int3(dims), int3(pads), C.byref(ms))
    check(st, "tridHipAutotune")
    return ms.value


def autotune_mixed(solvedim, dims, pads=None):
    pads = dims if pads is None else pads
    ms = C.c_float(0)
    check(_lib.lib().tridHipAutotuneMixed(len(dims), solvedim, int3(dims), int3(pads),
                                          C.byref(ms)), "tridHipAutotuneMixed")
    return ms.value


def autotune_mixed_f32(solvedim, dims, pads=None):
    """Tunes the FP32-stored solve (one refinement) for one geometry; returns
    the best time in ms."""
    pads = dims if pads is None else pads
    ms = C.c_float(0)
    check(_lib.lib().tridHipAutotuneMixedF32(len(dims), solvedim, int3(dims), int3(pads),
                                             C.byref(ms)), "tridHipAutotuneMixedF32")
    return ms.value


def autotune_adjoint(dtype, solvedim, dims, pads=None, grads=True):
    """Tunes the adjoint solve (with or without the coefficient gradients)
    for one geometry; returns the best time in ms."""
    pads = dims if pads is None else pads
    ms = C.c_float(0)
    check(_lib.lib().tridHipAutotuneAdjoint(1 if dtype == torch.float64 else 0,
                                            1 if grads else 0, len(dims), solvedim, int3(dims),
                                            int3(pads), C.byref(ms)), "tridHipAutotuneAdjoint")
    return ms.value


def autotune_ragged(dtype, offsets, max_len):
    check(_lib.lib().tridHipAutotuneRagged(1 if dtype == torch.float64 else 0, _ptr(offsets),
                                           offsets.numel() - 1, int(max_len)),
          "tridHipAutotuneRagged")
