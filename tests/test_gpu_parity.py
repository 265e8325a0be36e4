"""GPU parity: libstab_hip (through its C ABI) against the CPU oracle.

Tolerances: pencil entries 1e-13 relative to the matrix max; eigenvalues
|w - w_QZ| <= 1e-8 |w_QZ| against ZGGEV on the same FP64 pencil; normalised
eigenvectors 1e-6; reference eigenfunction tables |d| <= 1e-8 (the ndiff
-abserr 1e-8 check of thesis/TStest/run.sh:35-39).
"""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _inp(path):
    from stab_amd import solver as sv
    return sv.read_input(open(path).read())


def _oracle_params(path):
    from oracle import stab_oracle as so
    return so.parse_input(open(path).read())


def test_grid_and_operators(gpu_solver, ts_profile):
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    _say("grid: thesis temporal N=96")
    inp = _inp(golden("thesis_TStest", "temporal.inp"))
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, 0.1 + 0j, 4)
    g = gpu_solver.grid(pt)
    p = _oracle_params(golden("thesis_TStest", "temporal.inp"))
    y, eta, deta, d2eta = so.sgengrid(p)
    vm = so.getmean(y, *ts_profile)
    d1, d2, _, _ = so.derivative_operators(p)
    fin = y < 1e98
    assert np.allclose(g["y"][fin], y[fin], rtol=1e-14, atol=0)
    assert np.allclose(g["eta"], eta, rtol=0, atol=1e-15)
    assert np.allclose(g["deta"], deta, rtol=1e-14, atol=1e-300)
    assert np.allclose(g["d2eta"], d2eta, rtol=1e-14, atol=1e-300)
    assert np.max(np.abs(g["vm"] - vm)) < 1e-13
    assert np.max(np.abs(g["d1"] - d1)) <= 1e-13 * np.max(np.abs(d1))
    assert np.max(np.abs(g["d2"] - d2)) <= 1e-12 * np.max(np.abs(d2))


@pytest.mark.parametrize("case", ["temporal", "spatial"])
def test_pencil_matches_oracle(gpu_solver, ts_profile, case):
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"pencil: thesis {case}")
    path = golden("thesis_TStest", f"{case}.inp")
    inp = _inp(path)
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, 0.1 + 0j, 4)
    p = _oracle_params(path)
    if case == "temporal":
        pb = so.setup(p, *ts_profile, spatial=False)
        mats = dict(zip((_lib.MAT_A0, _lib.MAT_B0), so.temporal_pencil(pb)))
    else:
        pb = so.setup(p, *ts_profile, spatial=True)
        mats = dict(zip((_lib.MAT_C0, _lib.MAT_C1, _lib.MAT_C2), so.spatial_pencil(pb)))
    for which, ref in mats.items():
        got = gpu_solver.pencil(pt, which)
        scale = max(np.max(np.abs(ref)), 1e-300)
        err = np.max(np.abs(got - ref)) / scale
        _say(f"  matrix {which}: max rel err {err:.2e}")
        assert err < 1e-13


@pytest.mark.parametrize("mode", [0, 1])
def test_thesis_temporal_golden(gpu_solver, ts_profile, mode):
    """thesis/TStest/time.ref through the GPU path (plain FP64 and mixed)."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    _say(f"thesis temporal golden, mode {mode}")
    inp = _inp(golden("thesis_TStest", "temporal.inp"))
    hdr, tab = so.parse_table(open(golden("thesis_TStest", "time.ref")).read())
    guess = hdr["Omega"]
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, guess, 4)
    res = gpu_solver.solve([pt], mode)
    _say(f"  info {res.info[0]} stats {res.stats}")
    ev = res.evals[0, :4]
    _say(f"  evals {ev}")
    _say(f"  resid {res.resid[0, :4]}")
    j = sv.select_nearest(ev, guess)
    assert abs(ev[j] - guess) <= 1e-8 * abs(guess)
    assert res.resid[0, j] < 1e-10
    vec = sv.normalize(res.evecs[0][:, j], 5 * inp.ny)
    g = gpu_solver.grid(pt)
    mine = np.array([[g["y"][i]] + sum(([vec[i * 5 + k].real, vec[i * 5 + k].imag] for k in range(5)), [])
                     for i in range(inp.ny - 1, -1, -1)])
    d = np.abs(mine - tab)
    d[-1, 0] = 0.0
    _say(f"  table max abs diff {d.max():.2e}")
    assert d.max() <= 1e-8


@pytest.mark.parametrize("mode,ir,tol", [(0, 0, 1e-10), (1, 0, 1e-4), (1, 1, 1e-8), (1, 2, 1e-10), (1, -12, 1e-13)])
def test_shifted_solve(gpu_solver, ts_profile, mode, ir, tol):
    """LU + triangular solves of the shifted pencil against numpy.linalg.solve,
    with the shift away from the spectrum (forward error)."""
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"shifted solve mode {mode} ir {ir}")
    inp = _inp(golden("thesis_TStest", "temporal.inp"))
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, 0.2 + 0.05j, 4)
    M = gpu_solver.pencil(pt, _lib.MAT_SHIFTED)
    rng = np.random.default_rng(3)
    b = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    x, info = gpu_solver.shifted_solve(pt, b, mode, ir)
    xr = np.linalg.solve(M, b)
    err = np.linalg.norm(x - xr) / np.linalg.norm(xr)
    _say(f"  info {info} rel err {err:.2e}")
    assert info == 0
    assert err < tol


@pytest.mark.parametrize("mode,ir,tol", [(0, 0, 1e-15), (1, 0, 1e-6), (1, 2, 1e-11), (1, -12, 1e-15)])
def test_shifted_solve_near_eigenvalue(gpu_solver, ts_profile, mode, ir, tol):
    """Shift within 1e-4 of an eigenvalue: backward error ||b - M x|| / (||M|| ||x||)."""
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"near-singular shifted solve mode {mode} ir {ir}")
    inp = _inp(golden("thesis_TStest", "temporal.inp"))
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, 0.1146 + 0.0023j, 4)
    M = gpu_solver.pencil(pt, _lib.MAT_SHIFTED)
    rng = np.random.default_rng(4)
    b = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    x, info = gpu_solver.shifted_solve(pt, b, mode, ir)
    berr = np.linalg.norm(b - M @ x) / (np.linalg.norm(M, 2) * np.linalg.norm(x))
    _say(f"  info {info} backward err {berr:.2e}")
    assert info == 0
    assert berr < tol
