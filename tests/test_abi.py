"""C-ABI library checks that need no GPU: it loads, exports every symbol the
headers declare, keeps the reference's struct/enum layout, and rejects bad
arguments before touching a device."""
import ctypes as C
import os
import re
import subprocess

import pytest

from tridsolver_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def _declared_symbols():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"EXTERN_C\s+[\w\s\*]+?\b(\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_built():
    assert os.path.exists(_lib.LIBPATH), "run python -m tridsolver_amd.build"


def test_every_declared_symbol_is_exported_and_bound():
    declared = _declared_symbols()
    assert {"tridDmtsvStridedBatch", "tridSmtsvStridedBatch", "tridDmtsvStridedBatchInc",
            "tridSmtsvStridedBatchInc"} <= declared
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIBPATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = declared - exported
    assert not missing, missing
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    lib = _lib.lib()
    for name in declared:
        assert getattr(lib, name) is not None


def test_struct_and_enum_layout():
    # struct TridParams {void*; int opts[3]; int sync;} is 24 bytes on LP64
    # (reference scalar/include/trid_params.h:31-35).
    assert C.sizeof(_lib.TridParams) == 24
    assert _lib.TridParams.opts.offset == 8 and _lib.TridParams.sync.offset == 20
    hdr = open(os.path.join(INCLUDE, "trid_common.h")).read()
    vals = dict((m.group(1), int(m.group(2)))
                for m in re.finditer(r"(TRID_STATUS_\w+)\s*=\s*(\d+)", hdr))
    assert vals == {v: k for k, v in _lib.STATUS.items()}


def _call_invalid(**kw):
    lib = _lib.lib()
    args = dict(ndim=3, solvedim=0, dims=[4, 4, 4], pads=[4, 4, 4], ptr=C.c_void_p(16))
    args.update(kw)
    p = args["ptr"]
    dims = None if args["dims"] is None else _lib.int3(args["dims"])
    pads = None if args["pads"] is None else _lib.int3(args["pads"])
    return lib.tridDmtsvStridedBatch(None, p, p, p, p, args["ndim"], args["solvedim"], dims,
                                     pads)


@pytest.mark.parametrize("kw", [
    dict(ndim=0), dict(ndim=4), dict(solvedim=3), dict(solvedim=-1),
    dict(dims=[4, -1, 4]), dict(pads=[3, 4, 4]), dict(dims=None), dict(pads=None),
    dict(ptr=None),
    # 2^32 systems: more than one call solves (32-bit grids and system indices)
    dict(dims=[2, 65536, 65536], pads=[2, 65536, 65536]),
])
def test_invalid_arguments_rejected_without_device(kw):
    assert _call_invalid(**kw) == 3  # TRID_STATUS_INVALID_VALUE


def test_inc_requires_u():
    lib = _lib.lib()
    p = C.c_void_p(16)
    st = lib.tridSmtsvStridedBatchInc(None, p, p, p, p, None, 2, 1, _lib.int3([4, 4]),
                                      _lib.int3([4, 4]))
    assert st == 3


def test_adjoint_needs_x_only_for_coefficient_gradients():
    lib = _lib.lib()
    p = C.c_void_p(16)
    dims = _lib.int3([4, 4, 4])
    # a coefficient gradient without the forward solution x
    st = lib.tridDmtsvStridedBatchAdjoint(None, p, p, p, None, p, p, None, p, None, 3, 0,
                                          dims, dims)
    assert st == 3
    # lam without w
    st = lib.tridDmtsvStridedBatchAdjoint(None, p, p, p, None, None, p, None, None, None, 3,
                                          0, dims, dims)
    assert st == 3


def test_workspace_bytes():
    lib = _lib.lib()
    assert lib.tridHipWorkspaceBytes(1, 0, 3, 1, _lib.int3([8, 4, 2])) == 8 * 8 * 4 * 2
    assert lib.tridHipWorkspaceBytes(0, 1, 3, 1, _lib.int3([8, 4, 2])) == 2 * 4 * 8 * 4 * 2
    assert lib.tridHipWorkspaceBytes(1, 0, 0, 0, _lib.int3([8])) == 0


def test_workspace_bytes_long_systems():
    # contiguous FP64 systems of up to 16384 rows fit one workgroup: only the
    # per-lane path's workspace (N * nsys reals)
    lib = _lib.lib()
    assert lib.tridHipWorkspaceBytes(1, 0, 2, 0, _lib.int3([8192, 3])) == 8 * 8192 * 3
    assert lib.tridHipWorkspaceBytes(1, 0, 2, 0, _lib.int3([16384, 3])) == 8 * 16384 * 3
    # 17000 rows: three 8192-row segments of 512 chunks each, 10 reals per
    # chunk (6 records, 4 reduced arrays); the larger of that and the lane
    # path's
    assert lib.tridHipWorkspaceBytes(1, 0, 2, 0, _lib.int3([17000, 3])) == \
        max(8 * 17000 * 3, 8 * 10 * 1536 * 3)
    # 140000 rows: 18 segments -> 9216-row reduced systems, which fit
    assert lib.tridHipWorkspaceBytes(1, 0, 2, 0, _lib.int3([140000, 2])) == \
        max(8 * 140000 * 2, 8 * 10 * 9216 * 2)
    # 270000 rows: 33 segments -> 16896-row reduced systems, themselves too
    # long: a second level of 3 segments (1536 chunks)
    assert lib.tridHipWorkspaceBytes(1, 0, 2, 0, _lib.int3([270000, 2])) == \
        max(8 * 270000 * 2, 8 * 10 * (16896 * 2 + 1536 * 2))
    # strided FP64: up to 2048 rows in one workgroup; 5000 rows: 10 segments
    # of 512 rows, 32 chunks each
    assert lib.tridHipWorkspaceBytes(1, 0, 3, 1, _lib.int3([17, 2048, 2])) == 8 * 17 * 2048 * 2
    assert lib.tridHipWorkspaceBytes(1, 0, 3, 1, _lib.int3([17, 5000, 2])) == \
        max(8 * 17 * 5000 * 2, 8 * 10 * 320 * 34)


def test_force_hybrid_config_validation():
    lib = _lib.lib()
    assert lib.tridHipForceHybridConfig(-1, 1, 1, 1, 0) == 3
    assert lib.tridHipForceHybridConfig(8, 0, 1, 1, 0) == 3
    assert lib.tridHipForceHybridConfig(8, 1, 1, 0, 0) == 3
    assert lib.tridHipForceHybridConfig(8, 1, 1, 1, 5) == 3  # variants 0-4
    assert lib.tridHipForceHybridConfig(8, 1, 1, 1, -1) == 3
    assert lib.tridHipForceHybridConfig(0, 0, 0, 0, 0) == 0


def test_measurement_entry_points_reject_bad_arguments():
    lib = _lib.lib()
    p = C.c_void_p(16)
    dims = _lib.int3([64, 64, 64])
    # copy shape: threads not whole waves / too many, vectors not 1, 2 or 4, size not 16-byte
    assert lib.tridHipStreamCopyShape(p, p, 1 << 20, 100, 1, 1) == 3
    assert lib.tridHipStreamCopyShape(p, p, 1 << 20, 2048, 1, 1) == 3
    assert lib.tridHipStreamCopyShape(p, p, 1 << 20, 256, 3, 1) == 3
    assert lib.tridHipStreamCopyShape(p, p, 1 << 20, 256, 8, 1) == 3
    assert lib.tridHipStreamCopyShape(p, p, 1000, 256, 1, 1) == 3
    # access probe: x with seg 0 and M 4, 8, 16; y, z with 128- or 256-byte rows and M 16 or 32
    assert lib.tridHipAccessProbe(p, p, p, p, 0, dims, 256, 32, 1) == 3
    assert lib.tridHipAccessProbe(p, p, p, p, 0, dims, 0, 32, 1) == 3  # x: M = 4, 8, 16
    assert lib.tridHipAccessProbe(p, p, p, p, 0, dims, 0, 16, 0) == 3
    assert lib.tridHipAccessProbe(p, p, p, p, 3, dims, 0, 16, 1) == 3
    assert lib.tridHipAccessProbe(p, p, p, p, 2, dims, 512, 32, 1) == 3
    assert lib.tridHipAccessProbe(p, p, p, p, 2, dims, 256, 8, 1) == 3
    assert lib.tridHipAccessProbe(p, p, p, p, 2, dims, 256, 32, 0) == 3
    assert lib.tridHipAccessProbe(None, p, p, p, 2, dims, 256, 32, 1) == 3
    # adjoint tuner: the geometry checks of the solve
    assert lib.tridHipAutotuneAdjoint(1, 1, 3, 3, dims, dims, None) == 3
    assert lib.tridHipAutotuneAdjoint(1, 1, 3, 0, _lib.int3([4, -1, 4]), dims, None) == 3
