"""Learned alpha shifts for the spatial problem at its full batch: 2048 seeded
(omega, Re) points on the Mach 2 compressible Blasius layer, N = 128
(linearisation of order 1280), omega ~ U[0.02, 0.15], beta = 0,
Re ~ U[500, 2500] -- the held-out set of the shipped network
(stab_amd/data/alpha_mlp.npz, trained by `python -m stab_amd.learned spatial`
on 24,576 points with other seeds).

Each point's leading mode (least damped discrete downstream mode) from the
network's alpha shift (FP32 LU + FP64 residual inverse iteration, plain FP64
Krylov-Schur fallback from the input-file shift omega / 0.6 - 0.002 i) must
equal the plain FP64 path's answer to 1e-8 with residual below 1e-10; against
the same method from the input-file shift the network must need fewer solves,
and at most 5% of the points may fall back.
"""
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_learned_alpha_shift_config4(gpu_solver):
    from stab_amd import _lib
    from stab_amd import learned
    from stab_amd import meanflow as mf
    if not os.path.exists(learned.WEIGHTS_ALPHA):
        pytest.skip("alpha-shift network not trained")
    tab = mf.profile(mf.BlasiusSpec(Ma=2.0, Pr=0.72, law="sutherland", Te=288.15, eta_max=20.0, nrows=2000))
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    mlp = learned.ShiftMLP.load(gpu_solver, learned.WEIGHTS_ALPHA)
    seed = int(mlp.meta["seed"]) + 1
    pts, feats = learned.sample_spatial(2048, seed, 128)
    _say("config 4: plain FP64 Krylov-Schur labels, 2048 spatial points, N = 128")
    ref = learned.labels(gpu_solver, pts)
    have = np.isfinite(ref)
    _say(f"   {int(have.sum())} / 2048 points with a leading discrete mode")
    ev = learned.evaluate(gpu_solver, mlp, pts, feats, ref, shift_fn=learned.shifts_spatial)
    r = ev["result"]
    fb = [complex(p.shift.re, p.shift.im) for p in pts]
    base = gpu_solver.solve_leading(pts, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=fb)
    mb = int((have & (np.abs(base["evals"] - ref) <= 1e-8 * np.abs(ref))).sum())
    nl, nb = ev["info"][_lib.INFO_FALLBACK], int((base["info"] == _lib.INFO_FALLBACK).sum())
    _say(f"   learned: info {ev['info']}, match {ev['match']} / {ev['comparable']}, solves "
         f"{ev['stats']['op_applies']}, {ev['stats']['ms_total']:.1f} ms; shift rel err median "
         f"{ev['shift_rel_err_median']:.2e} p95 {ev['shift_rel_err_p95']:.2e}")
    _say(f"   input shift: fallbacks {nb}, match {mb}, solves {base['stats']['op_applies']}, "
         f"{base['stats']['ms_total']:.1f} ms")
    assert ev["match"] == ev["comparable"]
    ok = have & np.isin(r["info"], [_lib.INFO_OK, _lib.INFO_FALLBACK])
    assert np.all(r["resid"][ok] < 1e-10)
    # the input-file shift already lies near the leading spatial mode here (phase speed 0.6), so the
    # network's gain is in solves, not in fallbacks
    assert nl <= 0.05 * len(pts)
    assert ev["stats"]["op_applies"] < 0.9 * base["stats"]["op_applies"]
    mlp.close()
