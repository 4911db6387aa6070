"""The reference ADI apps' solution dump, and a comparison of two dumps.

Format (apps/adi/src/print_array.c:74-96): the nx*ny*nz values of u, raw
native-endian FP32 or FP64 (the app's precision), x fastest, then y, then z;
no header. The reference writes it to `<executable>.dat`; a padded leading
dimension is dropped on write.

The reference's compare tool (apps/adi/src/compare.c:71-154) reports the
signed sum of the differences, which can cancel. `compare` reports that sum
for continuity, plus the max absolute and max relative difference and the
normwise relative error, which cannot.

Command line:
  python -m tridsolver_amd.adi_io file1.dat file2.dat -nx 256 -ny 256 -nz 256 [-prec 64]
"""
from __future__ import annotations

import argparse
import sys

import numpy as np


def _dtype(prec):
    if prec not in (32, 64):
        raise ValueError("prec must be 32 or 64")
    return np.float64 if prec == 64 else np.float32


def dump_u(path, u, nx, ny, nz, ldim=None):
    """Writes u (host array or tensor; leading dimension `ldim` >= nx,
    default nx) in the reference's dump format."""
    if hasattr(u, "detach"):
        u = u.detach().cpu().numpy()
    u = np.asarray(u)
    ldim = nx if ldim is None else ldim
    if ldim < nx or u.size < ldim * ny * nz:
        raise ValueError("u is smaller than ldim*ny*nz")
    grid = u.reshape(-1)[:ldim * ny * nz].reshape(nz, ny, ldim)[:, :, :nx]
    np.ascontiguousarray(grid).tofile(path)


def load_u(path, nx, ny, nz, prec=64):
    """Reads a dump of an nx*ny*nz grid; returns a flat array, x fastest."""
    dt = _dtype(prec)
    n = nx * ny * nz
    data = np.fromfile(path, dtype=dt)
    if data.size != n:
        raise ValueError("%s holds %d values of %d bits, expected %d (%dx%dx%d)"
                         % (path, data.size, prec, n, nx, ny, nz))
    return data


def compare(path1, path2, nx, ny, nz, prec=64):
    """Differences between two dumps (file2 - file1, as the reference)."""
    u1 = load_u(path1, nx, ny, nz, prec).astype(np.float64)
    u2 = load_u(path2, nx, ny, nz, prec).astype(np.float64)
    diff = u2 - u1
    nan = np.isnan(diff)
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = diff / u1
    rel[~np.isfinite(rel)] = 0.0  # the reference zeroes NaN quotients
    scale = np.abs(u1).max() if u1.size else 0.0
    return {
        "points": int(u1.size),
        "nan_points": int(nan.sum()),
        "signed_sum": float(np.nansum(diff)),
        "max_abs": float(np.nanmax(np.abs(diff))) if u1.size else 0.0,
        "max_rel": float(np.abs(rel).max()) if u1.size else 0.0,
        "normwise_rel": float(np.nanmax(np.abs(diff)) / scale) if scale > 0 else 0.0,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(prog="tridsolver_amd.adi_io", prefix_chars="-")
    ap.add_argument("file1")
    ap.add_argument("file2")
    for name in ("nx", "ny", "nz"):
        ap.add_argument("-" + name, "--" + name, type=int, default=256)
    ap.add_argument("-prec", "--prec", type=int, default=64)
    a = ap.parse_args(argv)
    print("\nGrid dimensions: %d x %d x %d" % (a.nx, a.ny, a.nz))
    r = compare(a.file1, a.file2, a.nx, a.ny, a.nz, a.prec)
    print("signed sum of differences  %.17g" % r["signed_sum"])
    print("max |difference|           %.6e" % r["max_abs"])
    print("max |relative difference|  %.6e" % r["max_rel"])
    print("normwise relative error    %.6e" % r["normwise_rel"])
    print("NaN points                 %d" % r["nan_points"])
    return 0 if r["nan_points"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
