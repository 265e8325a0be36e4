"""Write tests/golden/sweep4096_qz.npz: the QZ (ZGGEV) eigenvalues nearest the
input-file shift for every point of the 4096-point temporal benchmark sweep.

Run once in the build container (about half an hour on 8 cores):

    python tests/golden/make_sweep_qz.py [--points 4096] [--workers 8]

For each point of the seeded sweep (compressible Blasius, M = 0.5, Pr = 0.72,
Sutherland with Te = 288.15 K, adiabatic wall; alpha ~ U[0.05, 0.4],
beta ~ U[0, 0.2], Re ~ U[300, 3000], N ~ U{64, 96, 128}, seed 2024) the
oracle's FP64 pencil (A0, B0) of temporal.f90:604-752 is built, its Dirichlet
rows are replaced as the solver does (unit row in A, zero row in B: the
reference's spurious omega = 0 eigenvalues of those rows move to infinity,
every other eigenpair is unchanged), and scipy's ZGGEV returns all
eigenvalues with left and right eigenvectors.  Stored per point: the KEEP
finite eigenvalues nearest sigma = 0.33 alpha, sorted by distance, and each
one's relative condition number
    kappa = ||y|| ||x|| (||A|| + |w| ||B||) / (|w| |y^H B x|)
(Frobenius norms), so that kappa * 1e-16 bounds how well any FP64 method can
fix that eigenvalue.
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

KEEP = 16
SEED = 2024
TE = 288.15


def sweep(npts: int, seed: int = SEED):
    from stab_amd import meanflow as mf
    tab = mf.profile(mf.BlasiusSpec(Ma=0.5, Pr=0.72, law="sutherland", Te=TE, eta_max=20.0, nrows=2000))
    rng = np.random.default_rng(seed)
    al = rng.uniform(0.05, 0.4, npts)
    be = rng.uniform(0.0, 0.2, npts)
    re = rng.uniform(300.0, 3000.0, npts)
    ny = rng.choice([64, 96, 128], npts)
    return tab, al, be, re, ny


def solver_pencil(tab, a, b, r, n):
    from oracle import stab_oracle as so
    txt = (f"1\n{TE}\n0.5 {float(r)!r} 0.72\n{int(n)} {2.0 / 46.0!r} 50\n1\n1\n0 0 0 0\n1\n"
           f"{float(a)!r} 0\n{float(b)!r} 0\n0\n")
    p = so.parse_input(txt)
    pb = so.setup(p, tab[:, 0], tab[:, 1:6].copy(), spatial=False)
    A, B = so.temporal_pencil(pb)
    dirichlet = np.where(~np.any(A != 0, axis=1))[0]
    A[dirichlet, :] = 0.0
    A[dirichlet, dirichlet] = 1.0
    B[dirichlet, :] = 0.0
    return A, B


def one(args):
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    import scipy.linalg as sla
    tab, a, b, r, n = args
    A, B = solver_pencil(tab, a, b, r, n)
    w, yl, xr = sla.eig(A, B, left=True, right=True, check_finite=False)
    fin = np.isfinite(w)
    w, yl, xr = w[fin], yl[:, fin], xr[:, fin]
    sig = 0.33 * a
    idx = np.argsort(np.abs(w - sig))[:KEEP]
    nA, nB = np.linalg.norm(A), np.linalg.norm(B)
    ws, ys, xs = w[idx], yl[:, idx], xr[:, idx]
    yBx = np.abs(np.einsum("ij,ij->j", ys.conj(), B @ xs))
    kap = np.linalg.norm(ys, axis=0) * np.linalg.norm(xs, axis=0) * (nA + np.abs(ws) * nB) / (np.abs(ws) * yBx)
    return ws, kap


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--points", type=int, default=4096)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=os.path.join(HERE, "sweep4096_qz.npz"))
    args = ap.parse_args()
    os.environ["OPENBLAS_NUM_THREADS"] = "1"
    import multiprocessing as mp
    tab, al, be, re, ny = sweep(args.points)
    jobs = [(tab, al[i], be[i], re[i], ny[i]) for i in range(args.points)]
    t0 = time.time()
    W = np.zeros((args.points, KEEP), dtype=np.complex128)
    K = np.zeros((args.points, KEEP))
    with mp.get_context("spawn").Pool(args.workers) as pool:
        for i, (ws, kap) in enumerate(pool.imap(one, jobs, chunksize=4)):
            W[i, :len(ws)] = ws
            K[i, :len(kap)] = kap
            if (i + 1) % 256 == 0:
                print(f"{i + 1}/{args.points} points, {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(args.out, alpha=al, beta=be, Re=re, ny=ny, sigma=0.33 * al, evals=W, kappa=K,
                        seed=SEED, keep=KEEP)
    print(f"wrote {args.out} in {time.time() - t0:.0f} s")


if __name__ == "__main__":
    main()
