"""Throughput benchmark of the batched Chebyshev stability solve on one MI355X.

Workload (one step = one pass of the hot path over the batch): a seeded sweep
of temporal stability points of a compressible Blasius boundary layer at
M = 0.5 (Pr = 0.72, gamma = 1.4, Sutherland viscosity with Te = 288.15 K,
adiabatic wall): alpha ~ U[0.05, 0.4], beta ~ U[0, 0.2], Re ~ U[300, 3000],
N ~ U{64, 96, 128} (pencils 320^2 .. 640^2 complex), Streett mapping
ymax = 50 with half-point y = 2.  Each point carries an input-file shift
sigma = 0.33 alpha (a phase-speed guess, as a getevec -v guess) and asks for
the 8 eigenpairs nearest it.  A mode counts as converged when its FP64
relative residual ||(A - w B) x|| / ((||A|| + |w| ||B||) ||x||) < 1e-10.

value = converged modes per second of the chosen path (FP32 LU with FP64
refinement); the plain FP64 path is timed on the same batch and reported as
``plain``; the CPU baseline times the oracle (ZGESV + ZGEEV, the reference's
method) on a bounded sample and checks the GPU eigenvalues against ZGGEV on
the same points.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# MI355X published peaks
PEAK_HBM_GBS = 8000.0
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_FP64_MFMA_TFLOPS = 78.6


def make_workload(npts: int, seed: int):
    from stab_amd import meanflow as mf
    tab = mf.profile(mf.BlasiusSpec(Ma=0.5, Pr=0.72, law="sutherland", Te=288.15, eta_max=20.0, nrows=2000))
    rng = np.random.default_rng(seed)
    al = rng.uniform(0.05, 0.4, npts)
    be = rng.uniform(0.0, 0.2, npts)
    re = rng.uniform(300.0, 3000.0, npts)
    ny = rng.choice([64, 96, 128], npts)
    return tab, al, be, re, ny


def make_points(al, be, re, ny, nev=8):
    from stab_amd import solver as sv
    pts = []
    for a, b, r, n in zip(al, be, re, ny):
        inp = sv.Inputs(mattyp=1, T0=288.15, Te=288.15, Ma=0.5, Re=float(r), Pr=0.72, ny=int(n),
                        yi=2.0 / 46.0, ymax=50.0, itype=1, alpha=complex(a, 0.0), beta=complex(b, 0.0))
        pts.append(sv.make_point(inp, complex(0.33 * a, 0.0), nev, 0))
    return pts


def cpu_worker(args):
    """One point of the CPU baseline: the oracle's reference method
    (ZGESV + ZGEEV) timed, plus ZGGEV on the same pencil for the check."""
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    tab, a, b, r, n = args
    from oracle import stab_oracle as so
    txt = f"1\n288.15\n0.5 {r!r} 0.72\n{int(n)} {2.0 / 46.0!r} 50\n1\n1\n0 0 0 0\n1\n{a!r} 0\n{b!r} 0\n0\n"
    p = so.parse_input(txt)
    t0 = time.perf_counter()
    pb = so.setup(p, tab[:, 0], tab[:, 1:6].copy(), spatial=False)
    A0, B0 = so.temporal_pencil(pb)
    w, v = so.temporal_solve_reference(A0, B0)
    dt = time.perf_counter() - t0
    wq, _ = so.qz(A0, B0, right=False)
    wq = wq[np.isfinite(wq)]
    return dt, w, wq


def cpu_baseline(tab, al, be, re, ny, gpu_evals, nev, sample, workers):
    import multiprocessing as mp
    idx = list(range(sample))
    jobs = [(tab, al[i], be[i], re[i], ny[i]) for i in idx]
    t0 = time.perf_counter()
    ctx = mp.get_context("spawn")
    with ctx.Pool(workers) as pool:
        out = pool.map(cpu_worker, jobs)
    wall = time.perf_counter() - t0
    # QZ check of the GPU eigenvalues on the sampled points
    match = total = 0
    for (dt, w, wq), i in zip(out, idx):
        for k in range(nev):
            ev = gpu_evals[i, k]
            d = np.min(np.abs(wq - ev))
            total += 1
            match += int(d <= 1e-8 * abs(ev))
    per_point = sum(o[0] for o in out) / len(out)
    return dict(value=nev * len(out) / wall, per_point_s=per_point, wall_s=wall, match=match, total=total)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--nev", type=int, default=8)
    ap.add_argument("--cpu-sample", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from stab_amd import _lib
    from stab_amd import solver as sv

    tab, al, be, re, ny = make_workload(args.points, args.seed)
    pts = make_points(al, be, re, ny, args.nev)
    S = sv.Solver(0)
    S.set_profile(0, tab[:, 0], tab[:, 1:6])

    def run(mode, steps, warm):
        for _ in range(warm):
            S.solve(pts, mode, want_vectors=False)
        times, stats, res = [], [], None
        for _ in range(steps):
            t0 = time.perf_counter()
            res = S.solve(pts, mode, want_vectors=False)
            times.append(time.perf_counter() - t0)
            stats.append(res.stats)
        return res, times, stats

    print(f"# bench: {args.points} points, mode mixed", file=sys.stderr, flush=True)
    res, times, stats = run(_lib.MODE_MIXED, args.steps, args.warmup)
    ok = (res.resid[:, :args.nev] < 1e-10) & (res.info[:, None] == 0)
    nconv = int(ok.sum())
    t_step = float(np.mean(times))
    value = nconv / t_step
    st = stats[-1]
    ms_lu = float(np.mean([s["ms_lu"] for s in stats]))
    ms_eig = float(np.mean([s["ms_eig"] for s in stats]))
    lu_tflops = st["lu_flops"] / (ms_lu * 1e-3) / 1e12

    plain = None
    if not args.no_plain:
        print("# bench: plain FP64 path", file=sys.stderr, flush=True)
        pres, ptimes, pstats = run(_lib.MODE_PLAIN_F64, 1, 0)
        pok = (pres.resid[:, :args.nev] < 1e-10) & (pres.info[:, None] == 0)
        pt = float(np.mean(ptimes))
        plain = dict(value=int(pok.sum()) / pt, ms_per_step=pt * 1e3, converged=int(pok.sum()),
                     ms_lu=pstats[-1]["ms_lu"], ms_eig=pstats[-1]["ms_eig"],
                     lu_tflops=pstats[-1]["lu_flops"] / (pstats[-1]["ms_lu"] * 1e-3) / 1e12)

    cpu = None
    if not args.no_cpu:
        print("# bench: CPU baseline sample", file=sys.stderr, flush=True)
        workers = max(1, min(16, os.cpu_count() or 1, args.cpu_sample))
        cb = cpu_baseline(tab, al, be, re, ny, res.evals, args.nev, args.cpu_sample, workers)
        cpu = {"value": cb["value"], "unit": "leading eigenmodes/s", "cores": workers, "kind": "port",
               "sample": f"first {args.cpu_sample} points of the same seeded sweep, reference method "
                         f"(ZGESV+ZGEEV full spectrum, 8 modes counted per point), one point per process, "
                         f"OPENBLAS_NUM_THREADS=1; {cb['per_point_s']:.2f} s per point on one core",
               "qz_match": f"{cb['match']}/{cb['total']}"}

    out = {
        "metric": "leading eigenmodes/sec at |dw|<=1e-8 (FP32 LU + FP64 refinement path)",
        "value": value,
        "unit": "leading eigenmodes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64 LU / f64 refinement",
        "data": "synthetic (seeded parameter sweep over a computed compressible Blasius profile)",
        "config": {"workload": "temporal sweep, compressible Blasius M=0.5, 4096 points, N in {64,96,128}, "
                               "8 modes nearest sigma=0.33 alpha",
                   "points": args.points, "nev": args.nev, "seed": args.seed,
                   "converged_modes": nconv, "total_modes": int(args.points * args.nev),
                   "ms_lu": ms_lu, "ms_eig": ms_eig, "ms_prep": st["ms_prep"], "ms_assemble": st["ms_assemble"],
                   "ks_cycles_max": st["restarts_max"]},
        "plain": plain,
        "vs_plain": (value / plain["value"]) if plain and plain["value"] > 0 else None,
        "roofline": {"bound": "mfma", "kernel": "k_lu (batched complex LU, FP32)", "achieved": lu_tflops,
                     "peak": PEAK_FP32_MFMA_TFLOPS, "unit": "TFLOP/s", "frac": lu_tflops / PEAK_FP32_MFMA_TFLOPS,
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
