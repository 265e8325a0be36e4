"""Write tests/golden/crossflow_qz.npz: QZ (ZGGEV) eigenvalues of the first 64
points of the seeded crossflow batch used by tests/test_gpu_crossflow.py.

Run once in the build container (a few minutes on 8 cores):

    python tests/golden/make_crossflow_qz.py [--points 64] [--workers 8]

The batch: temporal points on the swept (Falkner-Skan-Cooke) profile
tests/golden/FSCtest/profile.0 (edge velocity u = 0.582, w = 0.813), constant
viscosity, M = 0.3, Re = 400, Pr = 1, N = 256 (pencils of order 1280),
Streett mapping ymax = 50 with half-point y = 2, alpha and beta ~ U[0.2, 1.0]
and the shift omega ~ U[0, 0.01] (seed 5).  For each point the oracle's FP64
pencil with the solver's Dirichlet rows (unit row in A, zero row in B) goes to
ZGGEV; the 16 finite eigenvalues nearest the shift are stored, by distance.
"""
from __future__ import annotations

import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

SEED = 5
KEEP = 16
NY = 256


def batch(npts: int, seed: int = SEED):
    rng = np.random.default_rng(seed)
    al = rng.uniform(0.2, 1.0, npts)
    be = rng.uniform(0.2, 1.0, npts)
    om = rng.uniform(0.0, 0.01, npts)
    return al, be, om


def inp_text(a, b, ny=NY):
    return (f"0\n0.3 400.0 1.0\n{ny} {2.0 / 46.0!r} 50\n1\n1\n0 0 0 0\n1\n"
            f"{float(a)!r} 0\n{float(b)!r} 0\n0\n")


def one(args):
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    import scipy.linalg as sla
    from oracle import stab_oracle as so
    a, b, s = args
    ym, vt = so.read_profile(os.path.join(HERE, "FSCtest", "profile.0"))
    pb = so.setup(so.parse_input(inp_text(a, b)), ym, vt, spatial=False)
    A, B = so.temporal_pencil(pb)
    d = np.where(~np.any(A != 0, axis=1))[0]
    A[d, :] = 0.0
    A[d, d] = 1.0
    B[d, :] = 0.0
    w = sla.eigvals(A, B, check_finite=False)
    w = w[np.isfinite(w)]
    return w[np.argsort(np.abs(w - s))[:KEEP]]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=64)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "crossflow_qz.npz"))
    args = ap.parse_args()
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    import multiprocessing as mp
    al, be, om = batch(1024)
    t0 = time.time()
    with mp.get_context("spawn").Pool(args.workers) as pool:
        W = np.array(pool.map(one, [(al[i], be[i], om[i]) for i in range(args.points)], chunksize=1))
    np.savez_compressed(args.out, alpha=al[:args.points], beta=be[:args.points], shift=om[:args.points], evals=W,
                        seed=SEED, ny=NY)
    print(f"wrote {args.out} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
