"""Autotuner of the batched LU's work-group shape, run on the GPU.

For each Chebyshev order N and each precision of the LU (FP32 for the mixed
path, FP64 for the plain one) it times the LU kernel launch of a batch of
temporal points of that N with 256-thread work-groups (two per CU, up to three
panel rows per thread) and with 512-thread work-groups (one per CU), and writes
the faster one per padded order to the table the library reads when a context
is created (stab_amd/data/lu_tuning.txt: lines "f32|f64 npad threads") and the
measurements to a report.

    python -m stab_amd.autotune [--points 1024] [--reps 3] [--table PATH] [--report PATH]

The timing is the device time of the LU launches (stab_stats.ms_lu_kernel) of
stab_solve_leading with one inverse-iteration step, the work-group size forced
through STAB_LU_NT for each trial.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORDERS = (32, 48, 64, 96, 128, 160, 192, 256)


def _points(tab, ny, npts, seed):
    from . import solver as sv
    rng = np.random.default_rng(seed)
    pts = []
    for _ in range(npts):
        a, b, re = rng.uniform(0.05, 0.4), rng.uniform(0.0, 0.2), rng.uniform(300.0, 3000.0)
        inp = sv.Inputs(mattyp=1, T0=288.15, Te=288.15, Ma=0.5, Re=float(re), Pr=0.72, ny=int(ny), yi=2.0 / 46.0,
                        ymax=50.0, itype=1, alpha=complex(a, 0.0), beta=complex(b, 0.0))
        pts.append(sv.make_point(inp, complex(0.33 * a, 0.0), 4, 0))
    return pts


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=1024)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--table", default=os.path.join(HERE, "data", "lu_tuning.txt"))
    ap.add_argument("--report", default=None)
    args = ap.parse_args(argv)
    from . import _lib
    from . import meanflow as mf
    from .solver import Solver
    os.environ["STAB_GRAPH"] = "0"          # every trial issues its kernels directly
    tab = mf.profile(mf.BlasiusSpec(Ma=0.5, Pr=0.72, law="sutherland", Te=288.15, eta_max=20.0, nrows=1200))
    S = Solver(0)
    S.set_profile(0, tab[:, 0], tab[:, 1:6])
    rows, table = [], []
    for prec, mode in (("f32", _lib.MODE_MIXED), ("f64", _lib.MODE_PLAIN_F64)):
        for ny in ORDERS:
            npad = -(-5 * ny // 32) * 32
            pts = _points(tab, ny, args.points, 7 + ny)
            best = None
            for nt in (256, 512):
                if npad > 3 * nt:
                    continue
                os.environ["STAB_LU_NT"] = str(nt)
                S.solve_leading(pts[:16], mode, _lib.LEAD_INVIT, maxit=1)         # warm the variant
                ms = []
                for _ in range(args.reps):
                    r = S.solve_leading(pts, mode, _lib.LEAD_INVIT, maxit=1)
                    ms.append(r["stats"]["ms_lu_kernel"])
                t = float(np.median(ms))
                flops = args.points * 8.0 / 3.0 * (5 * ny) ** 3
                rows.append((prec, ny, npad, nt, t, flops / (t * 1e-3) / 1e12))
                print(f"{prec} N={ny:3d} npad={npad:4d} threads={nt}: {t:8.3f} ms  "
                      f"{rows[-1][5]:6.2f} TFLOP/s", flush=True)
                if best is None or t < best[1]:
                    best = (nt, t)
            table.append((prec, npad, best[0]))
    os.environ.pop("STAB_LU_NT", None)
    S.close()
    stamp = time.strftime("%Y-%m-%d %H:%M:%S")
    with open(args.table, "w") as f:
        f.write(f"# LU work-group size per padded order, written by stab_amd/autotune.py ({stamp}, "
                f"{args.points} points per order)\n# precision npad threads\n")
        for prec, npad, nt in table:
            f.write(f"{prec} {npad} {nt}\n")
    print(f"wrote {args.table}")
    if args.report:
        with open(args.report, "w") as f:
            f.write("# LU work-group shape: autotuner measurements\n\n")
            f.write(f"Command: `python -m stab_amd.autotune --points {args.points} --reps {args.reps}`"
                    f" ({stamp}, one MI355X).\n\n")
            f.write("Batch: temporal points of the compressible Blasius family (M = 0.5, Sutherland, Pr = 0.72), "
                    "one N per batch; time = device time of the LU launch (median of the repetitions); "
                    "TFLOP/s from 8/3 n^3 per point.\n\n")
            f.write("| precision | N | npad | threads | LU ms | TFLOP/s | chosen |\n|---|---|---|---|---|---|---|\n")
            chosen = {(p, n): t for p, n, t in table}
            for prec, ny, npad, nt, t, tf in rows:
                f.write(f"| {prec} | {ny} | {npad} | {nt} | {t:.3f} | {tf:.2f} | "
                        f"{'yes' if chosen[(prec, npad)] == nt else ''} |\n")
        print(f"wrote {args.report}")


if __name__ == "__main__":
    main()
