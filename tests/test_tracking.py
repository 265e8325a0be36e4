"""Mode tracking: the reference's selection rules (getab.f90, getalpha.f90)
on constructed spectra (CPU), and the batched GPU tracker against those rules
applied to QZ spectra of the oracle pencils."""
import numpy as np
import pytest

from conftest import golden


def _crossing_spectra(ps):
    """Two straight branches crossing at p = 0.5 plus a far third one."""
    return [np.array([0.2 + 0.6 * p + 0.01j, 0.5 + 0.0 * p + 0.011j, 3.0 + 1.0j]) for p in ps]


def test_getab_follows_a_branch_through_a_crossing():
    ps = np.linspace(0.0, 1.0, 11)
    sp = _crossing_spectra(ps)
    idx = __import__("stab_amd.tracking", fromlist=["x"]).select_getab(sp, ps, 0.2 + 0.01j, 0.26 + 0.01j)
    # the extrapolation keeps the rising branch (index 0) on both sides of the crossing
    assert idx == [0] * 11


def test_getalpha_jumps_branches_at_a_crossing():
    from stab_amd import tracking as tr
    ps = np.linspace(0.0, 1.0, 11)
    sp = _crossing_spectra(ps)
    idx = tr.select_getalpha(sp, 0.2 + 0.01j)
    # nearest-to-previous: at p = 0.5 the rising branch sits on the flat one and the pick stays flat
    assert idx[:5] == [0] * 5 and idx[-1] == 1


def test_extrapolation_and_nearest_rules():
    from stab_amd import tracking as tr
    assert tr.extrapolate(1.0 + 1j, 2.0, 0.5 + 0.5j, 1.0, 3.0) == 1.5 + 1.5j
    assert tr.nearest([1.0, 2.0, 1.0], 1.0) == 0          # first minimiser wins ties


@pytest.mark.gpu
def test_gpu_tracking_matches_getab_on_qz_spectra(gpu_solver, ts_profile):
    """Six curves (beta, Re) of the thesis temporal case tracked over alpha in
    [0.26, 0.33] in eight steps at N = 48: every step equals getab's pick from
    the QZ spectrum of the same pencil (1e-8)."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    from stab_amd import tracking as tr
    base = open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", "48 1.0 0.0")
    gpu_solver.set_profile(0, *ts_profile)
    alphas = np.linspace(0.26, 0.33, 8)
    cases = [(0.0, 1000.0), (0.02, 1000.0), (0.05, 1200.0), (0.0, 800.0), (0.03, 1500.0), (0.01, 900.0)]
    curves, params = [], []
    for be, re in cases:
        row = []
        for a in alphas:
            inp = sv.read_input(base)
            inp.alpha, inp.beta, inp.Re = complex(a, 0.0), complex(be, 0.0), re
            row.append(sv.make_point(inp, 0j, 4))
        curves.append(row)
        params.append(alphas)
    # the thesis TS mode (alpha 0.3086, Re 1000) has phase speed ~0.372: start the curves there
    g0 = np.array([0.372 * alphas[0] + 0.002j] * len(cases))
    g1 = np.array([0.372 * alphas[1] + 0.002j] * len(cases))
    ev, info = tr.track(gpu_solver, curves, np.array(params), g0, g1, rule="getab")
    print(f"tracked info {np.bincount(info.ravel(), minlength=4)}", flush=True)
    for c, (be, re) in enumerate(cases):
        spectra = []
        for a in alphas:
            p = so.parse_input(base)
            p.alpha, p.beta, p.Re = complex(a, 0.0), complex(be, 0.0), re
            pb = so.setup(p, *ts_profile, spatial=False)
            A0, B0 = so.temporal_pencil(pb)
            w, _ = so.qz(A0, B0, right=False)
            spectra.append(np.where(np.isfinite(w) & (np.abs(w) > 1e-9), w, 1e300))
        idx = tr.select_getab(spectra, alphas, g0[c], g1[c])
        ref = np.array([s[j] for s, j in zip(spectra, idx)])
        err = np.abs(ev[c] - ref) / np.abs(ref)
        print(f"  curve beta={be} Re={re}: max rel err {err.max():.1e}; {ev[c][-1]:.6f}", flush=True)
        assert np.all(err <= 1e-8)
