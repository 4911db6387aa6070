"""Builds libtridhip.so (the C ABI library) in-tree for gfx950.

Usage: python -m tridsolver_amd.build   (or tridsolver_amd.build.build()).
Every HIP translation unit in csrc/ is compiled by hipcc for gfx950 only, in
parallel, and linked into tridsolver_amd/lib/libtridhip.so.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
LIBNAME = "libtridhip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 in generated code.
CFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall",
    "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang",
    "-packed-fp32-ops", "-munsafe-fp-atomics",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src)[:-4] + ".o")
    cmd = [HIPCC] + CFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, obj, r.returncode, r.stdout + r.stderr


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sources()
    objs = []
    workers = min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(workers) as ex:
        for src, obj, rc, log in ex.map(_compile, srcs):
            if rc != 0:
                raise RuntimeError("hipcc failed for %s:\n%s" % (src, log))
            if verbose and log.strip():
                print(log)
            objs.append(obj)
    out = os.path.join(LIBDIR, LIBNAME)
    tmp = out + ".tmp"
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


def _c_program(src, out, defines=(), lang="c", include=None):
    """Compiles one C test program against the library's headers (or another
    include directory), linked against libtridhip.so and the HIP runtime,
    both found at run time through rpath."""
    root = os.path.dirname(HERE)
    rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
    if lang == "c":
        head = [os.environ.get("CC", "gcc"), "-std=c99"]
    else:
        head = [os.environ.get("CXX", "g++"), "-x", "c++", "-std=c++14"]
    cmd = head + ["-O2", "-Wall"] + ["-D" + d for d in defines] + [
        "-I" + (include or os.path.join(root, "include")), "-I" + os.path.join(rocm, "include"),
        "-D__HIP_PLATFORM_AMD__", src, "-x", "none", "-o", out, "-L" + LIBDIR, "-ltridhip",
        "-L" + os.path.join(rocm, "lib"), "-lamdhip64", "-lm",
        "-Wl,-rpath,$ORIGIN/../../tridsolver_amd/lib", "-Wl,-rpath," + os.path.join(rocm, "lib")]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("C caller build failed (%s):\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    return out


def build_c_caller() -> str:
    """Compiles the plain-C callers under tests/c: drop_in.c (this library's
    headers, C99) and fp_caller.c once per precision (FPPREC 0 and 1, the
    reference app's compile-time switch). Returns drop_in's path."""
    root = os.path.dirname(HERE)
    cdir = os.path.join(root, "tests", "c")
    for prec, name in ((0, "fp_caller_f32"), (1, "fp_caller_f64")):
        _c_program(os.path.join(cdir, "fp_caller.c"), os.path.join(cdir, name),
                   defines=("FPPREC=%d" % prec,))
    return _c_program(os.path.join(cdir, "drop_in.c"), os.path.join(cdir, "drop_in"))


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
    print(build_c_caller())
