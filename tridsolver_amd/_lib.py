"""ctypes binding of libtridhip.so, the C ABI library built from csrc/.

The library is loaded from tridsolver_amd/lib/ (built in-tree by
``tridsolver_amd.build``). There is no fallback: if the library is missing,
``lib()`` raises, so nothing can silently run on another implementation.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "lib", "libtridhip.so")

P = C.c_void_p
I = C.c_int
L = C.c_int64
SZ = C.c_size_t
D = C.c_double
F = C.c_float
IP = C.POINTER(C.c_int)


class TridParams(C.Structure):
    """Layout of struct TridParams (include/trid_params.h; reference
    scalar/include/trid_params.h:31-35)."""
    _fields_ = [("mpi_params", C.c_void_p), ("opts", C.c_int * 3), ("sync", C.c_int)]


STATUS = {
    0: "TRID_STATUS_SUCCESS",
    1: "TRID_STATUS_NOT_INITIALIZED",
    2: "TRID_STATUS_ALLOC_FAILED",
    3: "TRID_STATUS_INVALID_VALUE",
    4: "TRID_STATUS_ARCH_MISMATCH",
    5: "TRID_STATUS_MAPPING_ERROR",
    6: "TRID_STATUS_EXECUTION_FAILED",
    7: "TRID_STATUS_INTERNAL_ERROR",
    8: "TRID_STATUS_MATRIX_TYPE_NOT_SUPPORTED",
    9: "TRID_STATUS_ZERO_PIVOT",
}


class TridError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        super().__init__("%s returned %s" % (what, STATUS.get(status, status)))


# name -> (argtypes, restype); every symbol declared in include/*.h.
SIGNATURES = {
    # include/tridsolver.h
    "tridDmtsvStridedBatch": ([P, P, P, P, P, I, I, P, P], I),
    "tridSmtsvStridedBatch": ([P, P, P, P, P, I, I, P, P], I),
    "tridDmtsvStridedBatchInc": ([P, P, P, P, P, P, I, I, P, P], I),
    "tridSmtsvStridedBatchInc": ([P, P, P, P, P, P, I, I, P, P], I),
    # include/tridsolver_hip.h
    "tridHipLastErrorString": ([], C.c_char_p),
    "tridHipSetStream": ([P], I),
    "tridHipGetStream": ([C.POINTER(C.c_void_p)], I),
    "tridHipReserveWorkspace": ([SZ], I),
    "tridHipReleaseWorkspaces": ([], I),
    "tridHipWorkspaceBytes": ([I, I, I, I, P], SZ),
    "tridHipLastConfig": ([IP, IP, IP, IP, IP], I),
    "tridHipForceHybridConfig": ([I, I, I, I, I], I),
    "tridHipStreamCopy": ([P, P, SZ], I),
    "tridHipStreamCopyShape": ([P, P, SZ, I, I, I], I),
    "tridHipAccessProbe": ([P, P, P, P, I, P, I, I, I], I),
    "tridDmtsvStridedBatchMixed": ([P, P, P, P, P, I, I, P, P, I], I),
    "tridDmtsvStridedBatchMixedF32": ([P, P, P, P, P, P, I, I, P, P, I], I),
    "tridDmtsvStridedBatchAdjoint": ([P] * 10 + [I, I, P, P], I),
    "tridSmtsvStridedBatchAdjoint": ([P] * 10 + [I, I, P, P], I),
    "tridDmtsvRaggedBatch": ([P] * 6 + [L, L], I),
    "tridSmtsvRaggedBatch": ([P] * 6 + [L, L], I),
    "tridDmtsvRaggedBatchInc": ([P] * 7 + [L, L], I),
    "tridDmtsvRaggedBatchMixed": ([P] * 6 + [L, L, I], I),
    "tridDmtsvRaggedBatchMixedF32": ([P] * 7 + [L, L, I], I),
    "tridDmtsvStridedBatchPartBoundary": ([P] * 5 + [I, I, P, P, I, I, P], I),
    "tridSmtsvStridedBatchPartBoundary": ([P] * 5 + [I, I, P, P, I, I, P], I),
    "tridDmtsvStridedBatchPartSolve": ([P] * 5 + [I, I, P, P, I, I, P], I),
    "tridSmtsvStridedBatchPartSolve": ([P] * 5 + [I, I, P, P, I, I, P], I),
    "tridDmtsvStridedBatchIncPartSolve": ([P] * 6 + [I, I, P, P, I, I, P], I),
    "tridSmtsvStridedBatchIncPartSolve": ([P] * 6 + [I, I, P, P, I, I, P], I),
    "tridSmtsvRaggedBatchInc": ([P] * 7 + [L, L], I),
    "tridHipRaggedWorkspaceBytes": ([L], SZ),
    "tridDmtsvRaggedBatchNoAlloc": ([P] * 6 + [L, L], I),
    "tridHipSetAutotuneMode": ([I], I),
    "tridHipAutotune": ([I, I, I, I, P, P, C.POINTER(C.c_float)], I),
    "tridHipAutotuneMixed": ([I, I, P, P, C.POINTER(C.c_float)], I),
    "tridHipAutotuneMixedF32": ([I, I, P, P, C.POINTER(C.c_float)], I),
    "tridHipAutotuneRagged": ([I, P, L, L], I),
    "tridHipAutotuneAdjoint": ([I, I, I, I, P, P, C.POINTER(C.c_float)], I),
    "tridHipAutotuneSave": ([C.c_char_p], I),
    "tridHipAutotuneLoad": ([C.c_char_p], I),
    "tridHipAutotuneClear": ([], I),
    "tridHipForceRaggedConfig": ([I, I, I], I),
    # include/tridsolver_adi.h
    "tridAdiPreprocD": ([D] + [P] * 11 + [I, I, I], I),
    "tridAdiPreprocS": ([F] + [P] * 11 + [I, I, I], I),
    "tridAdiStepD": ([P, D] + [P] * 5 + [I, I, I], I),
    "tridAdiStepS": ([P, F] + [P] * 5 + [I, I, I], I),
    "tridAdiRunD": ([P, D] + [P] * 5 + [I, I, I, I, I], I),
    "tridAdiRunS": ([P, F] + [P] * 5 + [I, I, I, I, I], I),
    "tridAdiStepFusedD": ([D, P, P, I, I, I], I),
    "tridAdiStepFusedS": ([F, P, P, I, I, I], I),
    "tridAdiRunFusedD": ([P, D, P, P, I, I, I, I, I], I),
    "tridAdiRunFusedS": ([P, F, P, P, I, I, I, I, I], I),
}

_lib = None


def lib():
    """Loads libtridhip.so (once) and binds every declared symbol."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(
                "libtridhip.so is not built (%s); run `python -m tridsolver_amd.build`"
                % LIBPATH)
        lib_ = C.CDLL(LIBPATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(lib_, name)
            f.argtypes = args
            f.restype = res
        _lib = lib_
    return _lib


def exported_symbols():
    return sorted(SIGNATURES)


def check(status, what=""):
    if status != 0:
        detail = ""
        if status == 6 and _lib is not None:
            detail = " (%s)" % _lib.tridHipLastErrorString().decode()
        raise TridError(status, what + detail)
    return status


def int3(v):
    vals = [int(t) for t in v]
    return (C.c_int * 3)(*(vals + [1] * (3 - len(vals))))
