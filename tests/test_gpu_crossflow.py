"""Crossflow batch at the largest temporal order (N = 256, pencils of order 1280).

The swept-wing (Falkner-Skan-Cooke) profile of the reference's FSCtest
(tests/golden/FSCtest/profile.0: edge u = 0.582, w = 0.813) exercises the
crossflow terms (u3m != 0) with beta != 0.  1024 seeded points, alpha and
beta ~ U[0.2, 1.0], shift omega ~ U[0, 0.01], M = 0.3, Re = 400, Pr = 1,
constant viscosity.  The plain FP64 path and the FP32 LU + FP64 refinement path
each return the 4 eigenpairs nearest the shift; every converged pair must have
a FP64 residual below 1e-10, the two paths must agree to 1e-8, and the first 64
points must match ZGGEV on the same pencils (tests/golden/crossflow_qz.npz,
tests/golden/make_crossflow_qz.py) to 1e-8.
"""
import os
import sys

import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

FIX = os.path.join(GOLDEN, "crossflow_qz.npz")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def batch(gpu_solver):
    sys.path.insert(0, GOLDEN)
    import make_crossflow_qz as mc
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    ym, vt = so.read_profile(os.path.join(GOLDEN, "FSCtest", "profile.0"))
    gpu_solver.set_profile(0, ym, vt)
    al, be, om = mc.batch(1024)
    pts = [sv.make_point(sv.read_input(mc.inp_text(a, b)), complex(s, 0.0), 4) for a, b, s in zip(al, be, om)]
    return dict(pts=pts, al=al, be=be, om=om)


def test_crossflow_n256_batch(gpu_solver, batch):
    from stab_amd import _lib
    pts = batch["pts"]
    _say("crossflow N=256: plain FP64, 1024 points")
    rp = gpu_solver.solve(pts, _lib.MODE_PLAIN_F64, want_vectors=False)
    _say(f"   plain: {rp.stats['ms_total']:.1f} ms, converged {int((rp.info == 0).sum())}")
    _say("crossflow N=256: FP32 LU + FP64 refinement, 1024 points")
    rm = gpu_solver.solve(pts, _lib.MODE_MIXED, want_vectors=False)
    _say(f"   mixed: {rm.stats['ms_total']:.1f} ms, converged {int((rm.info == 0).sum())}")
    ok = (rp.info == _lib.INFO_OK) & (rm.info == _lib.INFO_OK)
    assert ok.mean() > 0.9
    assert np.all(rp.resid[ok, :4] < 1e-10) and np.all(rm.resid[ok, :4] < 1e-10)
    d = np.abs(rp.evals[ok, :4] - rm.evals[ok, :4]) / np.abs(rp.evals[ok, :4])
    _say(f"   paths agree to {d.max():.1e} on {ok.sum()} points")
    assert d.max() <= 1e-8
    if not os.path.exists(FIX):
        pytest.skip("crossflow QZ fixture not generated")
    f = np.load(FIX)
    n = len(f["alpha"])
    assert np.array_equal(f["alpha"], batch["al"][:n])
    worst, bad = 0.0, []
    for i in range(n):
        for r in (rp, rm):
            if r.info[i] != _lib.INFO_OK:
                continue
            for e in r.evals[i, :4]:
                rel = np.min(np.abs(f["evals"][i] - e)) / abs(e)
                worst = max(worst, rel)
                if rel > 1e-8:
                    bad.append((i, e, rel))
    _say(f"   {n} points vs ZGGEV: max rel err {worst:.1e}, {len(bad)} outside 1e-8")
    for b in bad[:8]:
        _say(f"     point {b[0]}: {b[1]:.12f} rel {b[2]:.1e}")
    assert not bad
