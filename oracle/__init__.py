"""CPU oracle for the tridsolver_amd parity tests.

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and bench.py's
cpu_baseline leg use this package to check and time against. The product
library never imports it.

Two libraries are wrapped:

* ``liboracle.so`` -- the C restatement of the reference CPU algorithm
  (``trid_oracle.c``; each function cites the reference lines it follows).
* ``_ref/libtridcpu_{avx512,avx2}.so`` -- the reference CPU solver itself,
  compiled from its sources (scalar/src/cpu/trid_cpu.cpp) by ``oracle/Makefile``.
  Present only where it was built; ``ref_available()`` says whether it is.

All functions take numpy arrays (C-contiguous, float64 or float32) and work in
place like the C API.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# The reference keeps SIMD scratch c2/d2[N_MAX] on each OpenMP thread's stack
# (scalar/src/cpu/trid_cpu.cpp:115-116); with N_MAX=8192 that is 1 MiB per
# array at FP64, beyond libgomp's default thread stack.
os.environ.setdefault("OMP_STACKSIZE", "64M")

_P = C.c_void_p
_I = C.c_int
_L = C.c_int64
_D = C.c_double
_F = C.c_float


def build(quiet: bool = True) -> None:
    """Compile liboracle.so and, where the reference sources exist, _ref."""
    out = subprocess.run(["make", "-C", _HERE, "all"], capture_output=True, text=True)
    if out.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + out.stdout + out.stderr)
    if not quiet:
        print(out.stdout)


def _load_oracle():
    path = os.path.join(_HERE, "liboracle.so")
    if not os.path.exists(path):
        build()
    lib = C.CDLL(path)
    for name in ("orc_batch_d", "orc_batch_s"):
        f = getattr(lib, name)
        f.argtypes = [_P, _P, _P, _P, _P, _I, _I, _P, _P, _I, _I, _I]
        f.restype = _I
    for name in ("orc_thomas_d", "orc_thomas_s"):
        f = getattr(lib, name)
        f.argtypes = [_P, _P, _P, _P, _P, _L, _L, _I]
        f.restype = None
    lib.orc_batch_ld.argtypes = [_P, _P, _P, _P, _P, _I, _I, _P, _P, _I]
    lib.orc_batch_ld.restype = _I
    for name in ("orc_ragged_d", "orc_ragged_s"):
        f = getattr(lib, name)
        f.argtypes = [_P, _P, _P, _P, _P, _P, _L, _I, _I]
        f.restype = None
    lib.orc_ragged_ld.argtypes = [_P, _P, _P, _P, _P, _P, _L, _I]
    lib.orc_ragged_ld.restype = None
    lib.orc_adjoint_d.argtypes = [_P] * 9 + [_I, _I, _P, _P, _I]
    lib.orc_adjoint_d.restype = _I
    lib.orc_adi_preproc_d.argtypes = [_D, _P, _P, _P, _P, _P, _I, _I, _I]
    lib.orc_adi_preproc_s.argtypes = [_F, _P, _P, _P, _P, _P, _I, _I, _I]
    lib.orc_adi_d.argtypes = [_D, _P, _I, _I, _I, _I, _I]
    lib.orc_adi_s.argtypes = [_F, _P, _I, _I, _I, _I, _I]
    lib.orc_residual_d.argtypes = [_P] * 5 + [_I, _I, _P, _P, _P]
    lib.orc_residual_d.restype = _D
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = _load_oracle()
    return _lib


def _ptr(x):
    return None if x is None else x.ctypes.data_as(_P)


def _ints(v):
    arr = (C.c_int * 3)(*([int(t) for t in v] + [1] * (3 - len(v))))
    return arr


def _check(x, dtype):
    if x is None:
        return
    assert isinstance(x, np.ndarray) and x.dtype == dtype and x.flags.c_contiguous, (
        "oracle arrays must be C-contiguous %s" % dtype)


def batch(a, b, c, d, u=None, *, ndim, solvedim, dims, pads=None, inc=False,
          simd_vec=0, nthreads=0, offset=0):
    """In-place batched solve with the reference's scalar arithmetic.

    Follows scalar/src/cpu/trid_cpu.cpp:358-455. ``offset`` is the element
    offset of the first interior point inside the (padded) arrays, as in the
    reference's padded tests (scalar/test/test_cpu.cpp:109-124).
    """
    pads = dims if pads is None else pads
    dt = d.dtype
    for x in (a, b, c, d, u):
        _check(x, dt)
    es = dt.itemsize
    base = lambda x: None if x is None else C.c_void_p(x.ctypes.data + offset * es)
    f = lib().orc_batch_d if dt == np.float64 else lib().orc_batch_s
    st = f(base(a), base(b), base(c), base(d), base(u), ndim, solvedim, _ints(dims),
           _ints(pads), int(bool(inc)), simd_vec, nthreads)
    if st != 0:
        raise ValueError("oracle batch: invalid arguments")


def thomas(a, b, c, d, u=None, *, N, stride=1, inc=False):
    """Single system, trid_scalar arithmetic (trid_cpu.cpp:312-353)."""
    f = lib().orc_thomas_d if d.dtype == np.float64 else lib().orc_thomas_s
    f(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(u), N, stride, int(bool(inc)))


def batch_ld(a, b, c, d, *, ndim, solvedim, dims, pads=None, nthreads=0):
    """Long-double Thomas of FP64 inputs, rounded to FP64 (error reference)."""
    pads = dims if pads is None else pads
    arrs = [np.ascontiguousarray(t, dtype=np.float64) for t in (a, b, c, d)]
    x = np.zeros_like(arrs[3])
    lib().orc_batch_ld(*[_ptr(t) for t in arrs], _ptr(x), ndim, solvedim, _ints(dims),
                       _ints(pads), nthreads)
    return x


def ragged(a, b, c, d, offsets, u=None, *, inc=False, nthreads=0):
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    f = lib().orc_ragged_d if d.dtype == np.float64 else lib().orc_ragged_s
    f(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(u), _ptr(offsets), len(offsets) - 1,
      int(bool(inc)), nthreads)


def ragged_ld(a, b, c, d, offsets, nthreads=0):
    """Long-double solve of each FP64 system of a ragged batch, rounded to
    FP64 (error reference); entries no system owns stay 0."""
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    arrs = [np.ascontiguousarray(t, dtype=np.float64) for t in (a, b, c, d)]
    x = np.zeros_like(arrs[3])
    lib().orc_ragged_ld(*[_ptr(t) for t in arrs], _ptr(x), _ptr(offsets), len(offsets) - 1,
                        nthreads)
    return x


def adjoint(a, b, c, x, w, *, ndim, solvedim, dims, pads=None, nthreads=0):
    """Returns (lam, grad_a, grad_b, grad_c) of L = sum(w * x), x = A^-1 d."""
    pads = dims if pads is None else pads
    lam, ga, gb, gc = (np.zeros_like(x) for _ in range(4))
    st = lib().orc_adjoint_d(_ptr(a), _ptr(b), _ptr(c), _ptr(x), _ptr(w), _ptr(lam),
                             _ptr(ga), _ptr(gb), _ptr(gc), ndim, solvedim, _ints(dims),
                             _ints(pads), nthreads)
    if st != 0:
        raise ValueError("oracle adjoint: invalid arguments")
    return lam, ga, gb, gc


def adi_preproc(lam, u, nx, ny, nz, out=None):
    """(du, a, b, c) of the ADI preproc; `out` may supply the four arrays."""
    dt = u.dtype
    du, a, b, c = out if out is not None else (np.empty_like(u) for _ in range(4))
    f = lib().orc_adi_preproc_d if dt == np.float64 else lib().orc_adi_preproc_s
    f(lam, _ptr(u), _ptr(du), _ptr(a), _ptr(b), _ptr(c), nx, ny, nz)
    return du, a, b, c


def adi(lam, u, nx, ny, nz, iters, nthreads=0):
    """In-place: ``iters`` ADI steps on u (apps/adi/src/adi_orig.c:147-242)."""
    f = lib().orc_adi_d if u.dtype == np.float64 else lib().orc_adi_s
    f(lam, _ptr(u), nx, ny, nz, iters, nthreads)


def residual(a, b, c, d, x, *, ndim, solvedim, dims, pads=None):
    """(max |d - A x| in long double, max |d|)."""
    pads = dims if pads is None else pads
    dm = C.c_double(0)
    r = lib().orc_residual_d(_ptr(a), _ptr(b), _ptr(c), _ptr(d), _ptr(x), ndim, solvedim,
                             _ints(dims), _ints(pads), C.byref(dm))
    return r, dm.value


# --------------------------------------------------------------------------
# The reference CPU library (oracle/_ref), compiled from its own sources.
# --------------------------------------------------------------------------
def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return all(k in fl for k in ("avx512f", "avx512dq", "avx512bw",
                                                 "avx512vl", "avx512cd"))
    except OSError:
        pass
    return False


_ref = None


def ref_path():
    isa = "avx512" if _cpu_has_avx512() else "avx2"
    return os.path.join(_HERE, "_ref", "libtridcpu_%s.so" % isa), isa


def ref_available() -> bool:
    return os.path.exists(ref_path()[0])


def ref_isa() -> str:
    return ref_path()[1]


def ref_simd_vec(dtype) -> int:
    """SIMD_VEC of the loaded reference build (scalar/include/trid_simd.h:39-67)."""
    width = 64 if ref_isa() == "avx512" else 32
    return width // np.dtype(dtype).itemsize


def ref():
    global _ref
    if _ref is None:
        path, _ = ref_path()
        if not os.path.exists(path):
            raise FileNotFoundError("reference CPU library not built: " + path)
        _ref = C.CDLL(path)
        for name in ("tridDmtsvStridedBatch", "tridSmtsvStridedBatch"):
            f = getattr(_ref, name)
            f.argtypes = [_P, _P, _P, _P, _P, _I, _I, _P, _P]
            f.restype = _I
        for name in ("tridDmtsvStridedBatchInc", "tridSmtsvStridedBatchInc"):
            f = getattr(_ref, name)
            f.argtypes = [_P, _P, _P, _P, _P, _P, _I, _I, _P, _P]
            f.restype = _I
        for name in ("trid_scalarD", "trid_scalarS"):
            f = getattr(_ref, name)
            f.argtypes = [_P, _P, _P, _P, _I, _I]
            f.restype = None
    return _ref


def ref_set_threads(n: int) -> None:
    """OpenMP threads of the reference library's later parallel regions
    (omp_set_num_threads of the libgomp instance it links)."""
    ref()
    C.CDLL("libgomp.so.1").omp_set_num_threads(int(n))


def ref_batch(a, b, c, d, u=None, *, ndim, solvedim, dims, pads=None, inc=False,
              offset=0):
    """Calls the reference's own tridDmtsvStridedBatch[Inc] on host arrays."""
    pads = dims if pads is None else pads
    dt = d.dtype
    es = dt.itemsize
    base = lambda x: None if x is None else C.c_void_p(x.ctypes.data + offset * es)
    pre = "tridD" if dt == np.float64 else "tridS"
    if inc:
        f = getattr(ref(), pre + "mtsvStridedBatchInc")
        st = f(None, base(a), base(b), base(c), base(d), base(u), ndim, solvedim,
               _ints(dims), _ints(pads))
    else:
        f = getattr(ref(), pre + "mtsvStridedBatch")
        st = f(None, base(a), base(b), base(c), base(d), ndim, solvedim, _ints(dims),
               _ints(pads))
    return st


def ref_thomas(a, b, c, d, *, N, stride=1):
    f = ref().trid_scalarD if d.dtype == np.float64 else ref().trid_scalarS
    f(_ptr(a), _ptr(b), _ptr(c), _ptr(d), N, stride)
