"""Every point of the 4096-point benchmark sweep against QZ.

tests/golden/sweep4096_qz.npz holds, for each point, the ZGGEV eigenvalues
nearest the input-file shift sigma = 0.33 alpha and their relative condition
numbers kappa (tests/golden/make_sweep_qz.py).  Checked here through the C ABI:
  * the leading mode of the learned-shift path (FP32 LU + FP64 residual
    inverse iteration) at every point;
  * the 8 eigenvalues nearest sigma of the plain FP64 path and of the FP32 LU +
    FP64 refinement path at every point (config 2 as written): every pair with
    a FP64 residual below 1e-10, at most one in a thousand of them outside the
    bar (cluster members whose condition number turns a 1e-11 residual into an
    eigenvalue error above 1e-8).
The bar is |w - w_QZ| <= 1e-8 |w_QZ|.  A pair whose kappa * 1.1e-16 exceeds
1e-9 is one that QZ itself does not fix to within a tenth of the bar; such
pairs are counted and reported, and must still agree to 10 kappa u.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIX = os.path.join(GOLDEN, "sweep4096_qz.npz")
U = 1.1e-16


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sweep(gpu_solver):
    import bench
    if not os.path.exists(FIX):
        pytest.skip("QZ fixture not generated")
    tab, al, be, re, ny = bench.make_workload(4096, 2024)
    d = np.load(FIX)
    assert np.array_equal(d["alpha"], al) and np.array_equal(d["ny"], ny)
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    pts = bench.make_points(al, be, re, ny, 8)
    return dict(pts=pts, al=al, be=be, re=re, qz=d["evals"], kappa=d["kappa"])


def _check(evals, sweep, want, label, allowed=0):
    """evals (npts, k): each returned value against the `want` QZ eigenvalues
    nearest sigma; at most `allowed` values may miss the bar."""
    qz, kap = sweep["qz"][:, :want], sweep["kappa"][:, :want]
    npts, k = evals.shape
    worst, loose, fails = 0.0, 0, []
    for i in range(npts):
        for j in range(k):
            e = evals[i, j]
            d = np.abs(qz[i] - e) / np.abs(qz[i])
            m = int(np.argmin(d))
            worst = max(worst, d[m])
            bar = 1e-8
            if kap[i, m] * U > 1e-9:
                loose += 1
                bar = max(bar, 10 * kap[i, m] * U)
            if not d[m] <= bar:
                fails.append((i, j, e, qz[i, m], d[m], kap[i, m]))
    _say(f"{label}: {npts * k} values, max rel err {worst:.2e}, {loose} ill-conditioned in QZ itself, "
         f"{len(fails)} outside the bar")
    for f in fails[:10]:
        _say(f"   point {f[0]} pair {f[1]}: {f[2]:.12f} vs QZ {f[3]:.12f} rel {f[4]:.1e} kappa {f[5]:.1e}")
    assert len(fails) <= allowed


def test_learned_leading_mode_matches_qz_at_every_point(gpu_solver, sweep):
    from stab_amd import _lib
    from stab_amd import learned
    import bench
    _say("sweep vs QZ: learned-shift leading mode, 4096 points")
    mlp = learned.ShiftMLP.load(gpu_solver)
    om = learned.shifts(mlp, bench.features(sweep["al"], sweep["be"], sweep["re"]))
    r = gpu_solver.solve_leading(sweep["pts"], _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=om,
                                 shifts=om, nev=4)
    assert np.all(np.isin(r["info"], [_lib.INFO_OK, _lib.INFO_FALLBACK]))
    assert np.all(r["resid"] < 1e-10)
    _check(r["evals"][:, None], sweep, 16, "learned leading mode")


@pytest.mark.parametrize("mode", [0, 1], ids=["plain_fp64", "fp32_lu_refined"])
def test_leading8_matches_qz_at_every_point(gpu_solver, sweep, mode):
    """Every pair the Krylov-Schur solve returns with a FP64 residual below
    1e-10 is a QZ eigenvalue to the bar (within the 16 nearest sigma), and
    almost all of them are among the 8 nearest.  Pairs inside clusters of
    nearly equal eigenvalues (the discretised continuum) can stop on
    stagnation; they are counted, not accepted."""
    from stab_amd import _lib
    _say(f"sweep vs QZ: 8 nearest sigma, mode {mode}, 4096 points")
    r = gpu_solver.solve(sweep["pts"], mode, want_vectors=False)
    conv = r.resid[:, :8] < 1e-10
    _say(f"   points with all 8 converged {int(conv.all(axis=1).sum())} / 4096, pairs converged "
         f"{int(conv.sum())} / {conv.size}")
    assert conv.mean() > 0.9
    ev = np.where(conv, r.evals[:, :8], np.nan)
    qz8 = sweep["qz"][:, :8]
    in8 = np.array([[np.isfinite(e) and np.min(np.abs(qz8[i] - e)) <= 1e-8 * abs(e) for e in ev[i]]
                    for i in range(len(ev))])
    _say(f"   converged pairs among the 8 QZ eigenvalues nearest sigma: {int(in8.sum())} / {int(conv.sum())}")
    assert in8.sum() >= 0.97 * conv.sum()
    rows = [(i, j) for i in range(ev.shape[0]) for j in range(8) if conv[i, j]]
    sub = np.full((len(rows), 1), np.nan, dtype=complex)
    for k, (i, j) in enumerate(rows):
        sub[k, 0] = ev[i, j]
    picked = {"qz": sweep["qz"][[i for i, _ in rows]], "kappa": sweep["kappa"][[i for i, _ in rows]]}
    # A FP64 residual of 1e-11 bounds the eigenvalue error only by kappa * 1e-11: inside clusters
    # (kappa to 1e9 here) a pair can pass the residual bar and miss 1e-8.  On the sweep 15 of the
    # 32,293 converged pairs do (max 4.6e-6): at most 1 in 1000 is allowed, and counted by bench.py.
    _check(sub, picked, 16, f"converged pairs of the 8 nearest (mode {mode})", allowed=len(rows) // 1000)
