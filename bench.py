"""Benchmark: grid points solved per second of the batched FP64 tridiagonal
solve on a 512^3 grid along x, y and z, on one MI355X, with the HBM roofline of
the dominant kernel and the reference CPU solver timed on this host's cores.

One step = one solve along x, one along y and one along z of the same 512^3
FP64 arrays (a, b, c read; d overwritten), i.e. 3 * 512^3 grid points.
Inputs are generated on the GPU before the timed region (diagonally dominant:
a, c ~ U(-1,1), b = 2 + |a| + |c| + U(0,1), d ~ U(-1,1); fixed seed). Each
direction's kernel shape is autotuned before the timed region.

Algorithmic bytes per grid point (the roofline's `achieved`): read a, b, c, d
and write d = 5 x 8 = 40 B (FP64).

Extra measurements in the same JSON line ("extras"; skipped by --quick):
FP32 x/y/z; FP32 sweeps + FP64 refinement vs FP64 (time ratio and error),
for FP64-stored operands (40 B per point) and for FP32-stored ones (24 B per
point, FP64 solution); adjoint / forward time on 256^3 FP64; the ragged
40,000-system batch; the hipGraph-replayed ADI step on 256^3; one 8-way
z-slab shard of the partitioned solve.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time

# OpenMP threads of the CPU baseline stay on neighbouring cores (read by
# libgomp when the CPU libraries load, so set before they are)
os.environ.setdefault("OMP_PROC_BIND", "close")

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
    ap.add_argument("--no-tune", action="store_true", help="use default kernel shapes")
    ap.add_argument("--quick", action="store_true", help="headline metric only")
    ap.add_argument("--cpu-slab", type=int, default=64,
                    help="z-extent of the CPU baseline sample (n x n x slab)")
    ap.add_argument("--verbose", action="store_true")
    ap.add_argument("--access-dirs", default="yz", choices=["x", "y", "z", "yz", "xyz"],
                    help="directions of --only access")
    ap.add_argument("--only", choices=["fp32", "mixed", "mixed-f32", "adjoint", "ragged", "adi",
                                       "partitioned", "cpu-extras", "oracle-checks", "access",
                                       "copy", "long"],
                    help="run one secondary workload and print its JSON instead of the bench "
                         "line (for profiling)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def gen_gpu(torch, n3, dtype, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    a = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    c = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    b = 2 + a.abs() + c.abs() + torch.rand(n3, dtype=dtype, device="cuda", generator=g)
    d = torch.rand(n3, dtype=dtype, device="cuda", generator=g) * 2 - 1
    return a, b, c, d


def event_ms(torch, fn, reps=5, warm=1):
    """Median device time of fn() over `reps`, measured with events on the
    current stream (the library is bound to it)."""
    for _ in range(warm):
        fn()
    ts = []
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        fn()
        e1.record()
        e1.synchronize()
        ts.append(e0.elapsed_time(e1))
    return statistics.median(ts)


COPY_SHAPES = [(t, v, nt) for t in (256, 512, 1024) for v in (1, 2, 4) for nt in (1, 0)]


def measure_copy_bw(torch, lib, nbytes=4 << 30, reps=10):
    """HBM bandwidth of the library's stream-copy kernel (read + write bytes):
    the best over a sweep of shapes (threads per workgroup, 16-byte vectors
    per thread, non-temporal or plain accesses). Returns (GB/s, shape, all)."""
    src = torch.empty(nbytes // 8, dtype=torch.float64, device="cuda").fill_(1.0)
    dst = torch.empty_like(src)
    res = {}
    for t, v, nt in COPY_SHAPES:
        ms = event_ms(torch, lambda: lib.tridHipStreamCopyShape(
            C.c_void_p(src.data_ptr()), C.c_void_p(dst.data_ptr()), nbytes, t, v, nt),
            reps=reps, warm=2)
        res["%d/%d/%s" % (t, v, "nt" if nt else "plain")] = round(2 * nbytes / (ms * 1e-3) / 1e9, 1)
    del src, dst
    best = max(res, key=res.get)
    return res[best], best, res


def bench_access(torch, api, tune, dirs="yz", n=512):
    """Bandwidth ceiling of the solve's access pattern (tridHipAccessProbe:
    the solve's loads and stores, an add per value in place of its
    arithmetic) along x, y and z of an n^3 FP64 grid: x at 4, 8 and 16 rows
    per lane; y and z at 128- and 256-byte row segments and 16 and 32 rows
    per lane; XCD groupings 1, 8, 16. 40 B per point as the solve. Beside it
    the solve of the same direction (tuned for the geometry first with tune,
    else the default shape). dirs: the directions to measure (any of "x",
    "y", "z")."""
    lib = api._lib.lib()
    n3 = n ** 3
    dims = [n, n, n]
    a, b, c, d = gen_gpu(torch, n3, torch.float64, seed=512)
    out = {}
    for sd, name in ((0, "x"), (1, "y"), (2, "z")):
        if name not in dirs:
            continue
        rows = {}
        shapes = [(0, M) for M in (4, 8, 16)] if sd == 0 else \
            [(seg, M) for seg in (128, 256) for M in (16, 32)]
        for seg, M in shapes:
            for G in (1, 8, 16):
                st = lib.tridHipAccessProbe(api._ptr(a), api._ptr(b), api._ptr(c), api._ptr(d),
                                            sd, api.int3(dims), seg, M, G)
                if st != 0:
                    continue
                ms = event_ms(torch, lambda: lib.tridHipAccessProbe(
                    api._ptr(a), api._ptr(b), api._ptr(c), api._ptr(d), sd, api.int3(dims), seg,
                    M, G), reps=9)
                key = "M%d_G%d" % (M, G) if sd == 0 else "seg%d_M%d_G%d" % (seg, M, G)
                rows[key] = {"ms": round(ms, 4), "alg_GBs": round(40.0 * n3 / ms / 1e6, 1)}
        d.copy_(a)
        if tune:
            api.autotune(torch.float64, sd, dims)
        ms = event_ms(torch, lambda: api.trid_strided_batch(a, b, c, d, solvedim=sd, dims=dims),
                      reps=9)
        best = max(rows, key=lambda k: rows[k]["alg_GBs"])
        out[name] = {"probe": rows, "probe_best": best,
                     "probe_best_GBs": rows[best]["alg_GBs"],
                     "solve_ms": round(ms, 4), "solve_alg_GBs": round(40.0 * n3 / ms / 1e6, 1),
                     "solve_config": api.last_config(),
                     "solve_over_probe_best": round(40.0 * n3 / ms / 1e6 / rows[best]["alg_GBs"], 3)}
    del a, b, c, d
    torch.cuda.empty_cache()
    return out


def cpu_baseline(n, slab, reps=5):
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

    def timed():
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
        return per_dir

    per_dir = timed()
    total = sum(per_dir.values())
    pts_s = 3 * npts / total
    single = None
    if kind == "reference":  # the same sample on one OpenMP thread
        oracle.ref_set_threads(1)
        one = timed()
        oracle.ref_set_threads(threads)
        single = {"value": 3 * npts / sum(one.values()),
                  "per_dir_gpts_s": {k: round(npts / v / 1e9, 3) for k, v in one.items()}}
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    desc = ("%s CPU solver, FP64 %dx%dx%d sample (x, y, z solves, median of %d), "
            "%d OpenMP threads (OMP_PROC_BIND=%s), %s, %s" % (
                "reference" if kind == "reference" else "oracle", n, n, slab, reps, threads,
                os.environ.get("OMP_PROC_BIND", ""),
                oracle.ref_isa() if kind == "reference" else "scalar",
                model or "cpu model unknown"))
    out = {"value": pts_s, "unit": "grid pts/s", "cores": threads, "kind": kind,
           "sample": desc, "alg_GBs": round(40.0 * pts_s / 1e9, 1),
           "per_dir_gpts_s": {k: round(npts / v / 1e9, 3) for k, v in per_dir.items()},
           "nproc": os.cpu_count()}
    if single is not None:
        out["single_thread"] = {"value": single["value"], "cores": 1,
                                "alg_GBs": round(40.0 * single["value"] / 1e9, 2),
                                "per_dir_gpts_s": single["per_dir_gpts_s"]}
    return out


def cpu_extras(reps=3):
    """CPU timings of the other workloads on bounded samples, OpenMP on all
    threads: ragged (the first 4,000 systems of the 40,000-system set) and the
    FP64 adjoint (all gradients, 256x256x64) by the oracle restatement; one
    FP64 and one FP32 ADI step at 256^3 by the reference CPU solver where it
    is built. Median of `reps` after a warm-up."""
    import numpy as np

    import oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count()

    def med(fn):
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)

    out = {"cores": threads}
    rng = np.random.default_rng(40000)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(8192), 40000))).astype(np.int64)[:4000]
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    npts = int(off[-1])
    a = rng.uniform(-1, 1, npts)
    c = rng.uniform(-1, 1, npts)
    b = 2 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, npts)
    d = rng.uniform(-1, 1, npts)
    x = d.copy()
    t = med(lambda: oracle.ragged(a, b, c, x, off, nthreads=threads))
    out["ragged_fp64"] = {"gpts_s": round(npts / t / 1e9, 3), "kind": "port",
                          "sample": "%d systems, %d unknowns, FP64" % (len(lens), npts)}
    # ADI: the reference app's CPU step (apps/adi/src/adi_cpu.cpp:223-269):
    # preproc, then the reference solver along x, y and z (Inc into u)
    n = 256
    dims3 = [n, n, n]
    use_ref = oracle.ref_available()
    for dt, key in ((np.float64, "adi_fp64_step"), (np.float32, "adi_fp32_step")):
        u = rng.uniform(0, 1, n ** 3).astype(dt)
        bufs = tuple(np.empty_like(u) for _ in range(4)) if use_ref else None

        def step():
            if not use_ref:
                oracle.adi(1.0, u, n, n, n, 1, nthreads=threads)
                return
            du, a3, b3, c3 = oracle.adi_preproc(1.0, u, n, n, n, out=bufs)
            oracle.ref_batch(a3, b3, c3, du, ndim=3, solvedim=0, dims=dims3)
            oracle.ref_batch(a3, b3, c3, du, ndim=3, solvedim=1, dims=dims3)
            oracle.ref_batch(a3, b3, c3, du, u, ndim=3, solvedim=2, dims=dims3, inc=True)

        t = med(step)
        del bufs
        out[key] = {"ms": round(t * 1e3, 2), "gpts_s": round(n ** 3 / t / 1e9, 3),
                    "kind": "reference" if use_ref else "port",
                    "sample": "one ADI step at 256^3: preproc (oracle), x, y, z-Inc solves (%s)"
                              % ("reference CPU library" if use_ref else "oracle")}
    dims = [256, 256, 64]
    m = int(np.prod(dims))
    a = rng.uniform(-1, 1, m)
    c = rng.uniform(-1, 1, m)
    b = 2 + np.abs(a) + np.abs(c) + rng.uniform(0, 1, m)
    xs = rng.uniform(-1, 1, m)
    w = rng.standard_normal(m)
    per = {}
    for sd, name in enumerate("xyz"):
        t = med(lambda: oracle.adjoint(a, b, c, xs, w, ndim=3, solvedim=sd, dims=dims,
                                       nthreads=threads))
        per[name] = round(m / t / 1e9, 3)
    out["adjoint_fp64_all_grads_gpts_s"] = dict(per, kind="port", sample="256x256x64 FP64")
    return out


def oracle_checks(torch, api, n, samp=64):
    """Part of the CPU-baseline leg: the long-double oracle checks the
    mixed-precision results the line reports, on a sample of systems. The
    same seeded inputs as bench_mixed / bench_mixed_f32 / bench_ragged are
    solved on the GPU at one refinement with the shapes tuned there; then, per
    direction, an 8 x samp block of whole systems (all n rows along the solve
    direction) is copied to the host and compared with the oracle's
    long-double solve of those systems (normwise max|x - x_ld| / max|x_ld|).
    The ragged batches are checked on their first 2,000 systems."""
    import numpy as np

    import oracle
    out = {}
    dims = [n, n, n]

    def block(t, sd):
        v = t.view(n, n, n)  # (z, y, x)
        if sd == 0:
            return v[0:8, 0:samp, :].contiguous().cpu().numpy(), [n, samp, 8]
        if sd == 1:
            return v[0:8, :, 0:samp].contiguous().cpu().numpy(), [samp, n, 8]
        return v[:, 0:8, 0:samp].contiguous().cpu().numpy(), [samp, 8, n]

    a, b, c, d0 = gen_gpu(torch, n ** 3, torch.float64, seed=64)
    per = {}
    for sd, name in enumerate("xyz"):
        x = d0.clone()
        api.mixed_strided_batch(a, b, c, x, solvedim=sd, dims=dims, refinements=1, sync=True)
        blk = [block(t, sd) for t in (a, b, c, d0, x)]
        sdims = blk[0][1]
        xl = oracle.batch_ld(*(t[0].ravel() for t in blk[:4]), ndim=3, solvedim=sd, dims=sdims)
        got = blk[4][0].ravel()
        per[name] = float(np.abs(got - xl).max() / np.abs(xl).max())
        del x
    out["fp64_stored_1_refinement"] = dict(per, sample="8 x %d systems of %d rows per direction"
                                           % (samp, n))
    del a, b, c, d0
    a, b, c, d = gen_gpu(torch, n ** 3, torch.float32, seed=32)
    xs = torch.empty(n ** 3, dtype=torch.float64, device="cuda")
    per = {}
    for sd, name in enumerate("xyz"):
        api.mixed_f32_strided_batch(a, b, c, d, xs, solvedim=sd, dims=dims, refinements=1,
                                    sync=True)
        blk = [block(t, sd) for t in (a, b, c, d, xs)]
        xl = oracle.batch_ld(*(t[0].ravel().astype(np.float64) for t in blk[:4]), ndim=3,
                             solvedim=sd, dims=blk[0][1])
        got = blk[4][0].ravel()
        per[name] = float(np.abs(got - xl).max() / np.abs(xl).max())
    out["fp32_stored_1_refinement"] = dict(
        per, sample="8 x %d systems of %d rows per direction; oracle solves the FP32 values "
                    "promoted to FP64" % (samp, n))
    del a, b, c, d, xs
    rng = np.random.default_rng(40000)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(8192), 40000))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(off[-1])
    toff = torch.from_numpy(off).cuda()
    k = 2000
    m = int(off[k])
    a, b, c, d = gen_gpu(torch, total, torch.float64, seed=8192)
    x = d.clone()
    api.ragged_batch_mixed(a, b, c, x, toff, max_len=8192, refinements=1, sync=True)
    host = [t[:m].cpu().numpy() for t in (a, b, c, d, x)]
    xl = oracle.ragged_ld(*host[:4], off[:k + 1])
    out["ragged_fp64_stored_1_refinement"] = float(np.abs(host[4] - xl).max() / np.abs(xl).max())
    del a, b, c, d, x
    a, b, c, d = gen_gpu(torch, total, torch.float32, seed=8192)
    x = api.ragged_batch_mixed_f32(a, b, c, d, toff, max_len=8192, refinements=1, sync=True)
    host = [t[:m].cpu().numpy().astype(np.float64) for t in (a, b, c, d, x)]
    xl = oracle.ragged_ld(*host[:4], off[:k + 1])
    out["ragged_fp32_stored_1_refinement"] = float(np.abs(host[4] - xl).max() / np.abs(xl).max())
    out["ragged_sample"] = "first %d of the 40,000 systems (%d unknowns)" % (k, m)
    del a, b, c, d, x
    torch.cuda.empty_cache()
    return out


def bench_fp32_dirs(torch, api, n, tune):
    n3 = n ** 3
    a, b, c, d = gen_gpu(torch, n3, torch.float32, seed=32)
    out = {}
    for sd, name in enumerate("xyz"):
        if tune:
            api.autotune(torch.float32, sd, [n, n, n])
        ms = event_ms(torch, lambda: api.trid_strided_batch(a, b, c, d, solvedim=sd,
                                                            dims=[n, n, n]))
        gbs = 20.0 * n3 / (ms * 1e-3) / 1e9
        out[name] = {"ms": round(ms, 4), "gpts_s": round(n3 / ms / 1e6, 2),
                     "alg_GBs": round(gbs, 1), "frac_peak": round(gbs / PEAK_HBM_GBS, 4),
                     "config": api.last_config()}
    del a, b, c, d
    torch.cuda.empty_cache()
    return out


def bench_mixed(torch, api, n, fp64_ms, tune):
    n3 = n ** 3
    a, b, c, d0 = gen_gpu(torch, n3, torch.float64, seed=64)
    out = {}
    for sd, name in enumerate("xyz"):
        if tune:
            api.autotune_mixed(sd, [n, n, n])
        x64 = d0.clone()
        api.trid_strided_batch(a, b, c, x64, solvedim=sd, dims=[n, n, n])
        runs = {}
        for ref in (1, 2):  # FP64 refinement passes after the FP32 solve
            xm = d0.clone()
            api.mixed_strided_batch(a, b, c, xm, solvedim=sd, dims=[n, n, n], refinements=ref)
            err = float((xm - x64).abs().max() / x64.abs().max())
            del xm
            d = d0.clone()
            ms = event_ms(torch, lambda: api.mixed_strided_batch(a, b, c, d, solvedim=sd,
                                                                 dims=[n, n, n],
                                                                 refinements=ref))
            del d
            runs[ref] = {"ms": round(ms, 4), "speedup_vs_fp64": round(fp64_ms[name] / ms, 3),
                         "max_rel_err_vs_fp64": err}
        del x64
        # the fewest refinements that meet 1e-12 against the FP64 solve
        ref = 1 if runs[1]["max_rel_err_vs_fp64"] <= 1e-12 else 2
        best = runs[ref]
        out[name] = {"ms": best["ms"], "gpts_s": round(n3 / best["ms"] / 1e6, 2),
                     "refinements": ref, "speedup_vs_fp64": best["speedup_vs_fp64"],
                     "max_rel_err_vs_fp64": best["max_rel_err_vs_fp64"],
                     "config": api.last_config(),
                     "by_refinements": {str(k): v for k, v in runs.items()}}
    del a, b, c, d0
    torch.cuda.empty_cache()
    return out


def bench_mixed_f32(torch, api, n, fp64_ms, tune=False):
    """FP64 solution of the n^3 system stored in FP32 (tridDmtsvStridedBatchMixedF32):
    16 B read + 8 B written per point. Speed is compared with the FP64 solve
    of the same size (fp64_ms, 40 B per point); accuracy with the FP64 solve
    of the FP32 values promoted to FP64 (the same system)."""
    n3 = n ** 3
    dims = [n, n, n]
    a, b, c, d = gen_gpu(torch, n3, torch.float32, seed=32)
    a64, b64, c64, d64 = (t.double() for t in (a, b, c, d))
    x = torch.empty(n3, dtype=torch.float64, device="cuda")
    out = {}
    for sd, name in enumerate("xyz"):
        if tune:
            api.autotune_mixed_f32(sd, dims)
        ref = d64.clone()
        api.trid_strided_batch(a64, b64, c64, ref, solvedim=sd, dims=dims)
        runs = {}
        for nref in (1, 2):
            api.mixed_f32_strided_batch(a, b, c, d, x, solvedim=sd, dims=dims, refinements=nref)
            err = float((x - ref).abs().max() / ref.abs().max())
            ms = event_ms(torch, lambda: api.mixed_f32_strided_batch(
                a, b, c, d, x, solvedim=sd, dims=dims, refinements=nref), reps=10, warm=3)
            runs[nref] = {"ms": round(ms, 4), "speedup_vs_fp64": round(fp64_ms[name] / ms, 3),
                          "max_rel_err_vs_fp64": err, "config": api.last_config()}
        del ref
        nref = 1 if runs[1]["max_rel_err_vs_fp64"] <= 1e-12 else 2
        best = runs[nref]
        out[name] = {"ms": best["ms"], "gpts_s": round(n3 / best["ms"] / 1e6, 2),
                     "alg_GBs": round(24.0 * n3 / best["ms"] / 1e6, 1), "refinements": nref,
                     "speedup_vs_fp64": best["speedup_vs_fp64"],
                     "max_rel_err_vs_fp64": best["max_rel_err_vs_fp64"],
                     "config": best["config"],
                     "by_refinements": {str(k): v for k, v in runs.items()}}
    del a, b, c, d, a64, b64, c64, d64, x
    torch.cuda.empty_cache()
    return out


def bench_adjoint(torch, api, tune, n=256):
    """Adjoint / forward time at n^3 FP64 per direction. With tune, the
    forward solve and both adjoint variants are autotuned for the geometry
    first, so the ratio compares tuned shapes."""
    n3 = n ** 3
    dims = [n, n, n]
    a, b, c, d = gen_gpu(torch, n3, torch.float64, seed=256)
    w = torch.randn(n3, dtype=torch.float64, device="cuda")
    out = {}
    for sd, name in enumerate("xyz"):
        if tune:
            api.autotune(torch.float64, sd, dims)
            api.autotune_adjoint(torch.float64, sd, dims, grads=True)
            api.autotune_adjoint(torch.float64, sd, dims, grads=False)
        x = d.clone()
        fwd = event_ms(torch, lambda: api.trid_strided_batch(a, b, c, x, solvedim=sd, dims=dims))
        x = d.clone()
        api.trid_strided_batch(a, b, c, x, solvedim=sd, dims=dims)
        lam, ga, gb, gc = (torch.empty_like(x) for _ in range(4))
        adj_all = event_ms(torch, lambda: api.adjoint_strided_batch(
            a, b, c, x, w, solvedim=sd, dims=dims, lam=lam, grad_a=ga, grad_b=gb, grad_c=gc))
        adj_d = event_ms(torch, lambda: api.adjoint_strided_batch(
            a, b, c, x, w, solvedim=sd, dims=dims, lam=lam))
        out[name] = {"fwd_ms": round(fwd, 4), "adj_all_ms": round(adj_all, 4),
                     "adj_dL_dd_ms": round(adj_d, 4), "tuned": bool(tune),
                     "adj_all_alg_GBs": round(72.0 * n3 / adj_all / 1e6, 1),
                     "fwd_alg_GBs": round(40.0 * n3 / fwd / 1e6, 1),
                     "adj_over_fwd_all_grads": round(adj_all / fwd, 3),
                     "adj_over_fwd_dL_dd_only": round(adj_d / fwd, 3),
                     "byte_floor_all_grads": 72 / 40, "byte_floor_dL_dd_only": 40 / 40}
    del a, b, c, d, w
    torch.cuda.empty_cache()
    return out


def ragged_graph_ms(torch, api, a, b, c, d, toff, reps=20):
    """Mean time per replay of the ragged solve (plan + persistent launch)
    captured once in a hipGraph on a side stream whose workspace is reserved
    beforehand; events on that stream around `reps` back-to-back replays."""
    s = torch.cuda.Stream()
    try:
        with torch.cuda.stream(s):
            api.reserve_workspace(api.ragged_workspace_bytes(toff.numel() - 1), stream=s)
            api.ragged_batch(a, b, c, d, toff, max_len=8192, stream=s)  # warm-up
        s.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=s):
            api.ragged_batch(a, b, c, d, toff, max_len=8192, stream=s)
        with torch.cuda.stream(s):
            g.replay()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            for _ in range(reps):
                g.replay()
            e1.record(s)
        e1.synchronize()
        ms = e0.elapsed_time(e1) / reps
        del g
        return ms
    finally:
        api.use_torch_stream(torch.cuda.current_stream())


def bench_ragged(torch, api, tune):
    import numpy as np
    rng = np.random.default_rng(40000)
    lens = np.round(np.exp(rng.uniform(np.log(8), np.log(8192), 40000))).astype(np.int64)
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    total = int(off[-1])
    toff = torch.from_numpy(off).cuda()
    out = {"systems": 40000, "unknowns": total}
    for dt, key, bpp in ((torch.float64, "fp64", 40.0), (torch.float32, "fp32", 20.0)):
        a, b, c, d = gen_gpu(torch, total, dt, seed=8192)
        if tune:
            api.autotune_ragged(dt, toff, 8192)
        ms = event_ms(torch, lambda: api.ragged_batch(a, b, c, d, toff, max_len=8192))
        gbs = (bpp * total + 8 * (len(off))) / (ms * 1e-3) / 1e9
        out[key] = {"ms": round(ms, 4), "gpts_s": round(total / ms / 1e6, 2),
                    "alg_GBs": round(gbs, 1), "frac_peak": round(gbs / PEAK_HBM_GBS, 4),
                    "config": api.last_config()}
        gms = ragged_graph_ms(torch, api, a, b, c, d, toff)
        gbs = (bpp * total + 8 * (len(off))) / (gms * 1e-3) / 1e9
        out[key]["graph"] = {"ms_per_replay": round(gms, 4),
                             "gpts_s": round(total / gms / 1e6, 2), "alg_GBs": round(gbs, 1),
                             "what": "planning kernels + persistent solve captured once in a "
                                     "hipGraph, replayed back to back"}
        del a, b, c, d
    # FP32 sweeps + FP64 refinement on the same 40,000 FP64 systems, at 1 and
    # 2 refinements; the line reports the fewest that meet 1e-12 against the
    # FP64 solve
    a, b, c, d = gen_gpu(torch, total, torch.float64, seed=8192)
    ref = d.clone()
    api.ragged_batch(a, b, c, ref, toff, max_len=8192, sync=True)
    ms64 = out["fp64"]["ms"]
    runs = {}
    for nref in (1, 2):
        x = d.clone()
        api.ragged_batch_mixed(a, b, c, x, toff, max_len=8192, refinements=nref, sync=True)
        err = float((x - ref).abs().max() / ref.abs().max())
        ms = event_ms(torch, lambda: api.ragged_batch_mixed(a, b, c, x, toff, max_len=8192,
                                                            refinements=nref))
        runs[nref] = {"ms": round(ms, 4), "speedup_vs_fp64": round(ms64 / ms, 3),
                      "max_rel_err_vs_fp64": err}
    nref = 1 if runs[1]["max_rel_err_vs_fp64"] <= 1e-12 else 2
    out["mixed"] = dict(runs[nref], refinements=nref, config=api.last_config(),
                        by_refinements={str(k): v for k, v in runs.items()})
    del a, b, c, d, ref, x
    # the same batch stored in FP32, FP64 solution (one refinement): 24 B per
    # point; error against the FP64 solve of the FP32 values promoted
    a, b, c, d = gen_gpu(torch, total, torch.float32, seed=8192)
    ref = d.double()
    api.ragged_batch(a.double(), b.double(), c.double(), ref, toff, max_len=8192, sync=True)
    x = api.ragged_batch_mixed_f32(a, b, c, d, toff, max_len=8192, refinements=1, sync=True)
    err = float((x - ref).abs().max() / ref.abs().max())
    ms = event_ms(torch, lambda: api.ragged_batch_mixed_f32(a, b, c, d, toff, x, max_len=8192,
                                                            refinements=1))
    out["mixed_f32_stored"] = {"ms": round(ms, 4), "speedup_vs_fp64": round(ms64 / ms, 3),
                               "alg_GBs": round((24.0 * total + 8 * len(off)) / ms / 1e6, 1),
                               "max_rel_err_vs_fp64": err, "refinements": 1,
                               "config": api.last_config()}
    del a, b, c, d, ref, x
    torch.cuda.empty_cache()
    return out


def bench_long(torch, api, tune):
    """Long FP64 systems, 40 B per point algorithmic. x: 8192 contiguous
    systems of 16384 rows, the longest one workgroup holds (512 lanes of 32
    rows, 40 B per point moved once). y: 8192 strided systems of 16384 rows,
    longer than one workgroup holds (2048 strided rows): the long-system path
    (hybrid segments reduced to one equation per chunk, the reduced systems
    solved by a second batched solve, the segments read again for the
    substitution; about 85 B per point moved). x8192: 8192 contiguous systems
    of 8192 rows (256 lanes of 32 rows, two workgroups per CU)."""
    out = {}
    for name, sd, dims in (("x", 0, [16384, 8192, 1]), ("y", 1, [8192, 16384, 1]),
                           ("x8192", 0, [8192, 8192, 1])):
        n = dims[0] * dims[1]
        a, b, c, d = gen_gpu(torch, n, torch.float64, seed=16384)
        if tune:
            api.autotune(torch.float64, sd, dims)
        ms = event_ms(torch, lambda: api.trid_strided_batch(a, b, c, d, solvedim=sd, dims=dims))
        out[name] = {"dims": dims, "ms": round(ms, 4), "gpts_s": round(n / ms / 1e6, 2),
                     "alg_GBs": round(40.0 * n / ms / 1e6, 1), "config": api.last_config()}
        del a, b, c, d
    torch.cuda.empty_cache()
    return out


def bench_partitioned(torch, api, n=512, shards=8):
    """One GPU's share when each z system of an n^3 FP64 grid is split into
    `shards` z-slabs (n x n x n/shards): the boundary step, the finishing
    solve (given every shard's boundary block), and a plain local solve of
    the same slab for comparison. The allgather between the two steps is not
    timed (it needs the other GPUs); its size per GPU is reported."""
    nz = n // shards
    dims = [n, n, nz]
    npts = n * n * nz
    nsys = n * n
    a, b, c, d = gen_gpu(torch, npts, torch.float64, seed=64)
    bnd = torch.empty(6 * nsys, dtype=torch.float64, device="cuda")
    k = shards // 2  # an interior shard: couplings on both sides
    t_bnd = event_ms(torch, lambda: api.part_boundary(a, b, c, d, bnd, solvedim=2, shard=k,
                                                      nshards=shards, dims=dims))
    bnd_all = bnd.repeat(shards)  # stands in for the gathered blocks (timing only)
    x = d.clone()
    t_fin = event_ms(torch, lambda: api.part_solve(a, b, c, x, bnd_all, solvedim=2, shard=k,
                                                   nshards=shards, dims=dims))
    x = d.clone()
    t_loc = event_ms(torch, lambda: api.trid_strided_batch(a, b, c, x, solvedim=2, dims=dims))
    del a, b, c, d, x, bnd, bnd_all
    torch.cuda.empty_cache()
    return {"shard": dims, "shards": shards, "boundary_ms": round(t_bnd, 4),
            # algorithmic bytes: a, b, c, d, read once by the one-pass sweep
            "boundary_alg_GBs": round(32.0 * npts / t_bnd / 1e6, 1),
            "finish_ms": round(t_fin, 4), "local_solve_ms": round(t_loc, 4),
            "partitioned_over_local": round((t_bnd + t_fin) / t_loc, 3),
            "allgather_bytes_per_gpu": 6 * 8 * nsys * shards}


def bench_adi(torch, n=256, iters=20):
    """ADI n^3 steps run as hipGraph replays (tridAdiRun*, one captured step
    replayed `iters` times per call). step_ms is the replay time per step:
    the difference between calls of iters and 3 * iters steps (best of 3
    each), which removes the capture, instantiation and final wait that one
    call adds; step_ms_incl_capture is one call of iters steps divided by
    iters."""
    from tridsolver_amd import adi

    def timed(call):
        ts = {}
        for k in (iters, 3 * iters):
            best = float("inf")
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                call(k)
                best = min(best, time.perf_counter() - t0)
            ts[k] = best
        return (ts[3 * iters] - ts[iters]) * 1e3 / (2 * iters), ts[iters] * 1e3 / iters

    out = {}
    for dt, key in ((torch.float64, "fp64"), (torch.float32, "fp32")):
        u = adi.initial_u(n, n, n, dt, "random", seed=1)
        ws = adi.Workspace(n, n, n, dt)
        adi.run(1.0, u, ws, 2, graph=True, sync=True)
        step_ms, incl = timed(lambda k: adi.run(1.0, u, ws, k, graph=True, sync=True))
        out[key] = {"step_ms": round(step_ms, 4), "gpts_s": round(n ** 3 / step_ms / 1e6, 2),
                    "step_ms_incl_capture": round(incl, 4), "steps": iters,
                    "bytes_per_point_step": 168 if key == "fp64" else 84}
        # matrix-free step: coefficients and right-hand side computed in the
        # sweeps, 56 (28) B/pt
        du = ws.du
        del ws
        adi.run_fused(1.0, u, du, 2, (n, n, n), graph=True, sync=True)
        step_ms, incl = timed(lambda k: adi.run_fused(1.0, u, du, k, (n, n, n), graph=True,
                                                      sync=True))
        byts = 56 if key == "fp64" else 28
        out[key + "_fused"] = {"step_ms": round(step_ms, 4),
                               "gpts_s": round(n ** 3 / step_ms / 1e6, 2),
                               "step_ms_incl_capture": round(incl, 4), "steps": iters,
                               "bytes_per_point_step": byts,
                               "gb_s": round(byts * n ** 3 / step_ms / 1e6, 1)}
    del u, du
    torch.cuda.empty_cache()
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the library has no multi-GPU path); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    from tridsolver_amd import _lib, api

    torch.cuda.set_device(0)
    lib = _lib.lib()
    n = args.n
    if args.only:
        api.use_torch_stream()
        tune = not args.no_tune
        if args.only in ("mixed", "mixed-f32"):
            a, b, c, d0 = gen_gpu(torch, n ** 3, torch.float64, seed=64)
            ref = {}
            for sd, name in enumerate("xyz"):
                if tune:
                    api.autotune(torch.float64, sd, [n] * 3)
                d = d0.clone()
                ref[name] = event_ms(torch, lambda: api.trid_strided_batch(
                    a, b, c, d, solvedim=sd, dims=[n] * 3))
            del a, b, c, d0, d
            torch.cuda.empty_cache()
            res = bench_mixed(torch, api, n, ref, tune) if args.only == "mixed" else \
                bench_mixed_f32(torch, api, n, ref, tune)
        else:
            run = {"fp32": lambda: bench_fp32_dirs(torch, api, n, tune),
                   "adjoint": lambda: bench_adjoint(torch, api, tune),
                   "ragged": lambda: bench_ragged(torch, api, tune),
                   "adi": lambda: bench_adi(torch),
                   "partitioned": lambda: bench_partitioned(torch, api),
                   "long": lambda: bench_long(torch, api, tune),
                   "cpu-extras": cpu_extras,
                   "oracle-checks": lambda: oracle_checks(torch, api, n),
                   "access": lambda: bench_access(torch, api, tune, args.access_dirs, n),
                   "copy": lambda: measure_copy_bw(torch, lib)}[args.only]
            res = run()
        print(json.dumps({"only": args.only, "result": res}))
        return
    n3 = n ** 3
    dims = [n, n, n]
    bytes_per_pt = 40.0
    tune = not args.no_tune

    api.use_torch_stream()  # library launches on torch's current stream
    copy_bw, copy_shape, copy_sweep = measure_copy_bw(torch, lib)
    log("copy bandwidth %.1f GB/s (threads/vectors/access %s)" % (copy_bw, copy_shape))

    if tune:
        for sd in range(3):
            api.autotune(torch.float64, sd, dims)
    a, b, c, d = gen_gpu(torch, n3, torch.float64, seed=512)
    torch.cuda.synchronize()

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
    del a, b, c, d
    torch.cuda.empty_cache()

    dirs = {}
    for name, ms in per_dir_ms.items():
        gbs = bytes_per_pt * n3 / (ms * 1e-3) / 1e9
        dirs[name] = {"ms": round(ms, 4), "gpts_s": round(n3 / (ms * 1e-3) / 1e9, 3),
                      "alg_GBs": round(gbs, 1), "frac_peak": round(gbs / PEAK_HBM_GBS, 4),
                      "frac_copy": round(gbs / copy_bw, 4), "config": cfgs[name]}
        log("fp64 %s: %s" % (name, dirs[name]))
    dom = max(per_dir_ms, key=per_dir_ms.get)
    ach = bytes_per_pt * n3 / (per_dir_ms[dom] * 1e-3) / 1e9
    # HBM bytes per launch from the committed counter profile, reported only
    # when this run's tuned shape is the one the counters were taken with
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        key = "fp64_%d_%s" % (n, dom)
        shapes = tj.get("configs", {}).get(key, [])
        if shapes and not isinstance(shapes[0], list):
            shapes = [shapes]  # one profiled shape
        if any(list(sh) == list(cfgs[dom]) for sh in shapes):
            traffic = tj.get(key)
        else:
            log("traffic: the profiled shape of %s differs from this run's %s; not reported"
                % (key, cfgs[dom]))

    extras = {}
    if not args.quick:
        extras["fp32_per_direction"] = bench_fp32_dirs(torch, api, n, tune)
        log("fp32: %s" % extras["fp32_per_direction"])
        extras["fp32_refined_vs_fp64"] = bench_mixed(torch, api, n, per_dir_ms, tune)
        log("mixed: %s" % extras["fp32_refined_vs_fp64"])
        extras["fp32_stored_refined_vs_fp64"] = bench_mixed_f32(torch, api, n, per_dir_ms, tune)
        log("mixed, FP32-stored: %s" % extras["fp32_stored_refined_vs_fp64"])
        extras["adjoint_256"] = bench_adjoint(torch, api, not args.no_tune)
        log("adjoint: %s" % extras["adjoint_256"])
        extras["ragged_40k"] = bench_ragged(torch, api, tune)
        log("ragged: %s" % extras["ragged_40k"])
        extras["adi_256_graph"] = bench_adi(torch)
        log("adi: %s" % extras["adi_256_graph"])
        extras["long_systems_16384_fp64"] = bench_long(torch, api, tune)
        log("long systems: %s" % extras["long_systems_16384_fp64"])
        extras["partitioned_z_512_8_shards"] = bench_partitioned(torch, api)
        log("partitioned: %s" % extras["partitioned_z_512_8_shards"])

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, args.cpu_slab)
        log("cpu: %s" % cpu)
        if not args.quick:
            extras["cpu_other_workloads"] = cpu_extras()
            log("cpu extras: %s" % extras["cpu_other_workloads"])
            chk = oracle_checks(torch, api, n)
            log("oracle checks: %s" % chk)
            extras["mixed_max_rel_err_vs_long_double_oracle"] = chk
            for key, src in (("fp32_refined_vs_fp64", "fp64_stored_1_refinement"),
                             ("fp32_stored_refined_vs_fp64", "fp32_stored_1_refinement")):
                for name in "xyz":
                    extras[key][name]["max_rel_err_vs_oracle_1_refinement_sample"] = \
                        chk[src][name]
            extras["ragged_40k"]["mixed"]["max_rel_err_vs_oracle_1_refinement_sample"] = \
                chk["ragged_fp64_stored_1_refinement"]
            extras["ragged_40k"]["mixed_f32_stored"][
                "max_rel_err_vs_oracle_1_refinement_sample"] = chk["ragged_fp32_stored_1_refinement"]

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
        "config": {"workload": "%d^3 FP64 strided batch, solve along x, y, z (one step = "
                               "3 solves of %d unknowns)" % (n, n3),
                   "grid": dims, "bytes_per_point": bytes_per_pt, "autotuned": tune},
        "roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(ach / PEAK_HBM_GBS, 4),
                     "traffic": traffic, "kernel": "solvedim %s (slowest direction)" % dom},
        "cpu_baseline": None if cpu is None else {
            k: cpu[k] for k in ("value", "unit", "cores", "kind", "sample", "alg_GBs",
                                "per_dir_gpts_s", "nproc", "single_thread") if k in cpu},
        "per_direction": dirs,
        "copy_GBs": round(copy_bw, 1),
        "copy_shape": copy_shape,
        "copy_sweep_GBs": copy_sweep,
        "extras": extras,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
