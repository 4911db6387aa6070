"""ADI heat diffusion on one MI355X: the reference app's x/y/z sweep pattern
(apps/adi/src/adi_cuda.cu) driven through libtridhip.so.

Library calls:
  tridAdiPreproc{D,S}  right-hand side and coefficients (apps/adi/src/preproc.cu:40-90)
  tridAdiStep{D,S}     preproc -> solve x -> solve y -> u += solve z
  tridAdiRun{D,S}      `iters` steps, optionally one captured hipGraph replayed
  tridAdi{Step,Run}Fused{D,S}  matrix-free step: a, b, c computed in the sweeps

Command line (same options as apps/adi/src/adi_cuda.cu:51-63,137-148):
  python -m tridsolver_amd.adi -nx 256 -ny 256 -nz 256 -iter 20 [-prec 32|64]
         [-optx 0 -opty 0 -optz 0] [-prof 1] [-graph 1] [-init ref|random]
         [-fused 1] [-dump u.dat]
Prints sum(u) and, with -prof 1, the time per grid element of each phase as
the reference does (adi_cuda.cu:377-384). -dump writes u in the reference
apps' raw format (apps/adi/src/print_array.c:74-96; see adi_io.py).
"""
from __future__ import annotations

import argparse
import ctypes as C
import time

import torch

from . import _lib
from .api import make_params, use_torch_stream
from ._lib import check


def _p(t):
    return C.c_void_p(t.data_ptr())


def _fns(dtype):
    sfx = "D" if dtype == torch.float64 else "S"
    L = _lib.lib()
    return (getattr(L, "tridAdiPreproc" + sfx), getattr(L, "tridAdiStep" + sfx),
            getattr(L, "tridAdiRun" + sfx))


class Workspace:
    """du and one shared coefficient set a, b, c for an nx*ny*nz grid."""

    def __init__(self, nx, ny, nz, dtype, device="cuda"):
        n = nx * ny * nz
        self.shape = (nx, ny, nz)
        self.du = torch.empty(n, dtype=dtype, device=device)
        self.a = torch.empty_like(self.du)
        self.b = torch.empty_like(self.du)
        self.c = torch.empty_like(self.du)


def initial_u(nx, ny, nz, dtype, kind="random", seed=0, device="cuda"):
    """'random': U(0,1) inside, 0 on the boundary (Dirichlet 0).
    'ref': the reference app's start, 1 on the boundary, 0 inside
    (apps/adi/src/adi_cuda.cu:191-202)."""
    u = torch.zeros(nz, ny, nx, dtype=dtype, device=device)
    if kind == "ref":
        u[0], u[-1] = 1, 1
        u[:, 0], u[:, -1] = 1, 1
        u[:, :, 0], u[:, :, -1] = 1, 1
    else:
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        inner = [max(0, n - 2) for n in (nz, ny, nx)]
        u[1:-1, 1:-1, 1:-1] = torch.rand(*inner, dtype=dtype, device=device, generator=g)
    return u.reshape(-1).contiguous()


def preproc(lam, u, ws, *, nine=None):
    """Writes du, a, b, c (and, with nine=(ay,by,cy,az,bz,cz), the y/z copies
    the reference app keeps)."""
    fn = _fns(u.dtype)[0]
    use_torch_stream()
    extra = [None] * 6 if nine is None else [_p(t) for t in nine]
    nx, ny, nz = ws.shape
    check(fn(lam, _p(u), _p(ws.du), _p(ws.a), _p(ws.b), _p(ws.c), *extra, nx, ny, nz),
          "tridAdiPreproc")


def step(lam, u, ws, opts=(0, 0, 0), sync=False):
    fn = _fns(u.dtype)[1]
    use_torch_stream()
    p = make_params(opts, 1 if sync else 0)
    nx, ny, nz = ws.shape
    check(fn(C.byref(p), lam, _p(u), _p(ws.du), _p(ws.a), _p(ws.b), _p(ws.c), nx, ny, nz),
          "tridAdiStep")


def run(lam, u, ws, iters, graph=True, opts=(0, 0, 0), sync=False):
    fn = _fns(u.dtype)[2]
    use_torch_stream()
    p = make_params(opts, 1 if sync else 0)
    nx, ny, nz = ws.shape
    check(fn(C.byref(p), lam, _p(u), _p(ws.du), _p(ws.a), _p(ws.b), _p(ws.c), nx, ny, nz,
             iters, 1 if graph else 0), "tridAdiRun")


def run_fused(lam, u, du, iters, shape, graph=True, sync=False):
    """`iters` matrix-free steps (tridAdiRunFused*): du is workspace only."""
    sfx = "D" if u.dtype == torch.float64 else "S"
    fn = getattr(_lib.lib(), "tridAdiRunFused" + sfx)
    use_torch_stream()
    p = make_params((0, 0, 0), 1 if sync else 0)
    nx, ny, nz = shape
    check(fn(C.byref(p), lam, _p(u), _p(du), nx, ny, nz, iters, 1 if graph else 0),
          "tridAdiRunFused")


def _phase_times(lam, u, ws, opts, iters):
    """Per-phase device times (ms per step) measured with events."""
    from . import api
    nx, ny, nz = ws.shape
    dims = [nx, ny, nz]
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(5)]
    acc = [0.0] * 4
    for _ in range(iters):
        ev[0].record()
        preproc(lam, u, ws)
        ev[1].record()
        api.trid_strided_batch(ws.a, ws.b, ws.c, ws.du, solvedim=0, dims=dims, opts=opts)
        ev[2].record()
        api.trid_strided_batch(ws.a, ws.b, ws.c, ws.du, solvedim=1, dims=dims, opts=opts)
        ev[3].record()
        api.trid_strided_batch(ws.a, ws.b, ws.c, ws.du, solvedim=2, dims=dims, u=u, opts=opts)
        ev[4].record()
        torch.cuda.synchronize()
        for k in range(4):
            acc[k] += ev[k].elapsed_time(ev[k + 1])
    return [t / iters for t in acc]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="tridsolver_amd.adi", prefix_chars="-")
    for name, default in (("nx", 256), ("ny", 256), ("nz", 256), ("iter", 10), ("optx", 0),
                          ("opty", 0), ("optz", 0), ("prof", 1), ("prec", 64), ("graph", 1),
                          ("devid", 0)):
        ap.add_argument("-" + name, "--" + name, type=int, default=default)
    ap.add_argument("-init", "--init", default="ref", choices=("ref", "random"))
    ap.add_argument("-fused", "--fused", type=int, default=0)
    ap.add_argument("-dump", "--dump", default=None)
    a = ap.parse_args(argv)
    torch.cuda.set_device(a.devid)
    dtype = torch.float64 if a.prec == 64 else torch.float32
    print("\nGrid dimensions: %d x %d x %d" % (a.nx, a.ny, a.nz))
    u = initial_u(a.nx, a.ny, a.nz, dtype, a.init)
    ws = Workspace(a.nx, a.ny, a.nz, dtype)
    lam = 1.0
    opts = (a.optx, a.opty, a.optz)
    shape = (a.nx, a.ny, a.nz)

    def go(uu, iters):
        if a.fused:
            run_fused(lam, uu, ws.du, iters, shape, graph=bool(a.graph), sync=True)
        else:
            run(lam, uu, ws, iters, graph=bool(a.graph), opts=opts, sync=True)

    go(u.clone(), 1)  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    go(u, a.iter)
    total = time.perf_counter() - t0
    if a.dump:
        from .adi_io import dump_u
        dump_u(a.dump, u, a.nx, a.ny, a.nz)
    print("\nComputing ADI on GPU: %f (s)" % total)
    print("end h_u sum = %.17g" % float(u.double().sum()))
    n = a.nx * a.ny * a.nz
    if a.prof == 1:
        ph = _phase_times(lam, u.clone(), ws, opts, max(1, min(a.iter, 5)))
        print("Time per element averaged on %d iterations: " % a.iter)
        print("[total] \t[prepro] \t[trid_x] \t[trid_y] \t[trid_z]")
        print("%e \t%e \t%e \t%e \t%e" % ((total / a.iter) / n, ph[0] * 1e-3 / n,
                                          ph[1] * 1e-3 / n, ph[2] * 1e-3 / n,
                                          ph[3] * 1e-3 / n))
    else:
        print("Avg(per iter) \n[total]\n%f" % (total / a.iter))
    print("Done.")


if __name__ == "__main__":
    main()
