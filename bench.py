"""Benchmark: grid points solved per second of the batched FP64 tridiagonal
solve on a 512^3 grid along x, y and z, on one MI355X, with the HBM roofline of
the dominant kernel and the reference CPU solver timed on this host's cores.

One step = one solve along x, one along y and one along z of the same
512^3 FP64 arrays (a, b, c read; d overwritten), i.e. 3 * 512^3 grid points.
Inputs are generated on the GPU before the timed region (diagonally dominant:
a, c ~ U(-1,1), b = 2 + |a| + |c| + U(0,1), d ~ U(-1,1); fixed seed).

Algorithmic bytes per grid point (the roofline's `achieved`): read a, b, c, d
and write d = 5 x 8 = 40 B (FP64).

Prints ONE JSON line (the driver's contract) and, with --verbose, a table.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X datasheet HBM3E bandwidth


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512, help="grid edge (n^3 points)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-slab", type=int, default=64,
                    help="z-extent of the CPU baseline sample (n x n x slab)")
    ap.add_argument("--verbose", action="store_true")
    return ap.parse_args()


def gen_gpu(n3, dtype, seed):
    import torch
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    a = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    c = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    b = 2 + a.abs() + c.abs() + torch.rand(n3, dtype=dtype, device="cuda", generator=g)
    d = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    return a, b, c, d


def measure_copy_bw(torch, lib, nbytes=4 << 30, reps=10):
    """HBM bandwidth of the library's stream-copy kernel (read + write bytes)."""
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(2):
        lib.tridHipStreamCopy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), nbytes)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        lib.tridHipStreamCopy(C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), nbytes)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del src, dst
    return 2 * nbytes / (ms * 1e-3) / 1e9


def cpu_baseline(n, slab, reps=3):
    """Reference CPU solver (oracle/_ref, compiled from the reference sources)
    on an n x n x slab FP64 sample, x, y and z solves; falls back to the
    oracle restatement if the reference library is not present."""
    import numpy as np

    import oracle
    rng = np.random.default_rng(7)
    dims = [n, n, slab]
    npts = n * n * slab
    a = rng.uniform(-1, 1, npts)
    c = rng.uniform(-1, 1, npts)
    b = 2 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, npts)
    d = rng.uniform(-1, 1, npts)
    kind = "reference" if oracle.ref_available() else "port"
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()

    def run(sd, x):
        if kind == "reference":
            oracle.ref_batch(a, b, c, x, ndim=3, solvedim=sd, dims=dims)
        else:
            oracle.batch(a, b, c, x, ndim=3, solvedim=sd, dims=dims)

    per_dir = {}
    x = d.copy()
    for sd, name in enumerate("xyz"):
        run(sd, x)  # warm-up
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run(sd, x)
            ts.append(time.perf_counter() - t0)
        per_dir[name] = statistics.median(ts)
    total = sum(per_dir.values())
    pts_s = 3 * npts / total
    desc = ("%s CPU solver, FP64 %dx%dx%d sample (x, y, z solves, median of %d), "
            "%d OpenMP threads, %s" % ("reference" if kind == "reference" else "oracle",
                                      n, n, slab, reps, threads,
                                      oracle.ref_isa() if kind == "reference" else "scalar"))
    return {"value": pts_s, "unit": "grid pts/s", "cores": threads, "kind": kind,
            "sample": desc, "per_dir_s": per_dir}


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU only (the solve has no "
              "multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from tridsolver_amd import _lib, api

    torch.cuda.set_device(0)
    lib = _lib.lib()
    n = args.n
    n3 = n ** 3
    dims = [n, n, n]
    dt = torch.float64
    bytes_per_pt = 40.0

    copy_bw = measure_copy_bw(torch, lib)

    a, b, c, d = gen_gpu(n3, dt, seed=512)
    torch.cuda.synchronize()
    api.use_torch_stream()  # library launches on torch's current stream

    def solve(sd):
        api.trid_strided_batch(a, b, c, d, solvedim=sd, dims=dims, sync=False)

    cfgs = {}
    for _ in range(args.warmup):
        for sd in range(3):
            solve(sd)
            cfgs["xyz"[sd]] = api.last_config()
    torch.cuda.synchronize()

    ev = [[(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
           for _ in range(3)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        for sd in range(3):
            ev[k][sd][0].record()
            solve(sd)
            ev[k][sd][1].record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    per_dir_ms = {}
    for sd in range(3):
        per_dir_ms["xyz"[sd]] = statistics.mean(ev[k][sd][0].elapsed_time(ev[k][sd][1])
                                                for k in range(args.steps))
    ms_per_step = wall * 1e3 / args.steps
    value = 3 * n3 * args.steps / wall

    dirs = {}
    for name, ms in per_dir_ms.items():
        gbs = bytes_per_pt * n3 / (ms * 1e-3) / 1e9
        dirs[name] = {"ms": round(ms, 4), "gpts_s": round(n3 / (ms * 1e-3) / 1e9, 3),
                      "alg_GBs": round(gbs, 1), "frac_peak": round(gbs / PEAK_HBM_GBS, 4),
                      "frac_copy": round(gbs / copy_bw, 4), "config": cfgs[name]}
    dom = max(per_dir_ms, key=per_dir_ms.get)
    ach = bytes_per_pt * n3 / (per_dir_ms[dom] * 1e-3) / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("fp64_512_" + dom)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, args.cpu_slab)

    out = {
        "metric": "grid pts solved/s (FP64 512^3 x/y/z batched tridiagonal solves)",
        "value": value,
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded diagonally dominant a,b,c,d generated on the GPU)",
        "config": {"workload": "512^3 FP64 strided batch, solve along x, y, z (one step = "
                               "3 solves of 134,217,728 unknowns)",
                   "grid": dims, "bytes_per_point": bytes_per_pt},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "kernel": "solvedim %s" % dom},
        "cpu_baseline": None if cpu is None else {k: cpu[k] for k in
                                                  ("value", "unit", "cores", "kind", "sample")},
        "per_direction": dirs,
        "copy_GBs": round(copy_bw, 1),
    }
    if args.verbose:
        for name, r in dirs.items():
            print("%s: %s" % (name, r), file=sys.stderr)
        if cpu:
            print("cpu: %s" % cpu, file=sys.stderr)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
