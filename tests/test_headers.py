"""The drop-in headers compile the reference's kinds of callers (no GPU needed):

- a caller that declares its arrays with FP and literals with F, chosen by
  FPPREC (the reference app's compile-time switch,
  scalar/include/trid_common.h:39-47), builds and links against
  libtridhip.so at FPPREC 0 and 1 and gets the precision it asked for;
- FP stays undefined without FPPREC (the library itself needs none), and a
  bad FPPREC stops the build;
- the alignment macros keep the reference's values
  (scalar/include/trid_common.h:49-55);
- where the reference tree is present: tests/c/drop_in.c compiles against the
  reference's own headers and links against this library (so the symbols and
  types it declares are the ones this library exports), and the reference
  ADI app's compare.c compiles against this library's headers.
"""
import os
import subprocess

import pytest

from tridsolver_amd import _lib
from tridsolver_amd import build as B

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")
CDIR = os.path.join(ROOT, "tests", "c")
REF = "/root/reference"
REF_INCLUDE = os.path.join(REF, "scalar", "include")


def _cc(args, lang="c"):
    cc = ["gcc", "-std=c99"] if lang == "c" else ["g++", "-x", "c++", "-std=c++14"]
    r = subprocess.run(cc + ["-Wall"] + args, capture_output=True, text=True)
    return r.returncode, r.stdout + r.stderr


@pytest.fixture(scope="module")
def libbuilt():
    if not os.path.exists(_lib.LIBPATH):
        pytest.fail("libtridhip.so is not built (python -m tridsolver_amd.build)")


@pytest.mark.parametrize("prec", [0, 1])
def test_fp_caller_builds_and_links(prec, libbuilt, tmp_path):
    out = str(tmp_path / ("fp_caller_%d" % prec))
    B._c_program(os.path.join(CDIR, "fp_caller.c"), out, defines=("FPPREC=%d" % prec,))
    assert os.path.exists(out)
    nm = subprocess.run(["nm", "-u", out], capture_output=True, text=True, check=True).stdout
    want = "tridSmtsvStridedBatch" if prec == 0 else "tridDmtsvStridedBatch"
    other = "tridDmtsvStridedBatch" if prec == 0 else "tridSmtsvStridedBatch"
    assert want in nm and other not in nm


@pytest.mark.parametrize("prec,size", [(0, 4), (1, 8)])
def test_fp_and_f_follow_fpprec(prec, size, tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "tridsolver.h"\n'
                   '#define CAT_(x, s) x##s\n#define CAT(x, s) CAT_(x, s)\n'
                   'typedef char fp_size[sizeof(FP) == %d ? 1 : -1];\n'
                   'typedef char lit_size[sizeof(CAT(1.5, F)) == %d ? 1 : -1];\n'
                   'typedef char align[(ALIGN == 32 && ALIGN_FLOAT == 8 && ALIGN_DOUBLE == 4 &&'
                   ' CUDA_ALIGN_BYTE == 32 && WARP_SIZE == 32 && MAXDIM == 3) ? 1 : -1];\n'
                   'int main(void) { FP x = CAT(0.5, F); return (int)x; }\n' % (size, size))
    rc, log = _cc(["-DFPPREC=%d" % prec, "-I" + INCLUDE, str(src), "-o", str(tmp_path / "t")])
    assert rc == 0, log


def test_fp_undefined_without_fpprec(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "tridsolver.h"\n#ifdef FP\n#error FP defined\n#endif\n'
                   'int main(void) { struct TridParams p; p.sync = 1; return p.sync - 1; }\n')
    rc, log = _cc(["-I" + INCLUDE, str(src), "-o", str(tmp_path / "t")])
    assert rc == 0, log


def test_bad_fpprec_is_an_error(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "trid_common.h"\nint main(void) { return 0; }\n')
    rc, log = _cc(["-DFPPREC=2", "-I" + INCLUDE, str(src), "-o", str(tmp_path / "t")])
    assert rc != 0 and "FPPREC" in log


def test_headers_compile_as_cxx(libbuilt, tmp_path):
    out = str(tmp_path / "drop_in_cxx")
    B._c_program(os.path.join(CDIR, "drop_in.c"), out, lang="c++")
    assert os.path.exists(out)


@pytest.mark.skipif(not os.path.isdir(REF_INCLUDE), reason="reference tree not present")
def test_drop_in_against_reference_headers(libbuilt, tmp_path):
    # The reference header declares TridParams as a C++ struct without a C
    # typedef (scalar/include/trid_params.h:31-35), so its callers are C++;
    # drop_in.c is compiled as C++ here, with the FPPREC the reference
    # header requires.
    out = str(tmp_path / "drop_in_ref")
    B._c_program(os.path.join(CDIR, "drop_in.c"), out, defines=("FPPREC=1",), lang="c++",
                 include=REF_INCLUDE)
    nm = subprocess.run(["nm", "-u", out], capture_output=True, text=True, check=True).stdout
    for sym in ("tridDmtsvStridedBatch", "tridDmtsvStridedBatchInc", "tridSmtsvStridedBatch"):
        assert sym in nm


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree not present")
@pytest.mark.parametrize("prec", [0, 1])
def test_reference_compare_c_against_these_headers(prec, tmp_path):
    src = os.path.join(REF, "apps", "adi", "src", "compare.c")
    rc, log = _cc(["-std=gnu99", "-DFPPREC=%d" % prec, "-I" + INCLUDE, "-c", src,
                   "-o", str(tmp_path / "compare.o")])
    assert rc == 0, log
