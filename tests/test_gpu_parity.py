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


@pytest.mark.parametrize("mode,ir,tol", [(0, 0, 1e-10), (1, 0, 1e-4), (1, 1, 1e-8), (1, 2, 1e-10), (1, -12, 1e-10)])
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


@pytest.mark.parametrize("mode,ir,tol", [(0, 0, 1e-13), (1, -12, 1e-13)])
@pytest.mark.parametrize("ny", [8, 70, 160])
def test_shifted_solve_orders(gpu_solver, ts_profile, ny, mode, ir, tol):
    """Solves whose LU leaves each block's interchanges out of the columns
    before it (the forward solve applies them block by block) at orders 64
    (one block), 352 (last block 32 wide) and 800 (the 512-thread LU): backward
    error ||b - M x|| / (||M|| ||x||), which a misplaced interchange would
    raise to O(1)."""
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"shifted solve ny {ny} mode {mode} ir {ir}")
    text = open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", f"{ny} 1.0 0.0")
    inp = sv.read_input(text)
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, 0.2 + 0.05j, 4)
    M = gpu_solver.pencil(pt, _lib.MAT_SHIFTED)
    rng = np.random.default_rng(ny)
    b = rng.standard_normal(M.shape[0]) + 1j * rng.standard_normal(M.shape[0])
    x, info = gpu_solver.shifted_solve(pt, b, mode, ir)
    berr = np.linalg.norm(b - M @ x) / (np.linalg.norm(M, 2) * np.linalg.norm(x))
    _say(f"  info {info} backward err {berr:.2e}")
    assert info == 0
    assert berr < tol


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


@pytest.mark.parametrize("mode,ny,tol", [(0, 8, 1e-14), (0, 70, 1e-13), (0, 96, 1e-13), (0, 128, 1e-13),
                                         (0, 160, 1e-13), (1, 8, 1e-5), (1, 70, 1e-5), (1, 96, 1e-5),
                                         (1, 128, 1e-5), (1, 160, 1e-5)])
def test_lu_factors(gpu_solver, ts_profile, mode, ny, tol):
    """P (Dr M Dc) = L U for the batched LU kernel (Dr = Dc = I in FP64).
    Orders 64 (one 64-column block), 352 (last block 32 wide), 480, 640 and
    800 (the 640-thread variant)."""
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"lu factors mode {mode} ny {ny}")
    text = open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", f"{ny} 1.0 0.0")
    inp = sv.read_input(text)
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, 0.2 + 0.05j, 4)
    M = gpu_solver.pencil(pt, _lib.MAT_SHIFTED)
    lu, perm, dr, dc, info = gpu_solver.lu_factors(pt, mode)
    n, npad = M.shape[0], lu.shape[0]
    Mp = np.eye(npad, dtype=complex)
    Mp[:n, :n] = M
    Ms = (dr[:, None] * Mp) * dc[None, :]
    L = np.tril(lu, -1) + np.eye(npad)
    U = np.triu(lu)
    err = np.linalg.norm(Ms[perm] - L @ U) / np.linalg.norm(Ms)
    _say(f"  info {info} ||PA - LU||/||A|| = {err:.2e}  perm ok {sorted(perm) == list(range(npad))}")
    assert sorted(perm) == list(range(npad))
    assert err < tol


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case,ny,ref_file", [("thesis", 96, "thesis_TStest/space.ref"), ("test", 32, "test/space.1")])
def test_spatial_golden(gpu_solver, ts_profile, mode, case, ny, ref_file):
    """Spatial alpha and eigenfunction through the GPU path: thesis/TStest/space.ref
    (N=96) and test/space.1 (N=32), ndiff -abserr 1e-8 semantics."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    _say(f"spatial golden {case} N={ny} mode {mode}")
    inp_path = golden("thesis_TStest", "spatial.inp") if case == "thesis" else golden("test", "input.dat")
    inp = _inp(inp_path)
    hdr, tab = so.parse_table(open(golden(*ref_file.split("/"))).read())
    guess = hdr["Alpha"]
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, guess, 4)
    res = gpu_solver.solve([pt], mode)
    ev = res.evals[0, :4]
    _say(f"  info {res.info[0]} evals {ev} resid {res.resid[0, :4]}")
    j = sv.select_nearest(ev, guess)
    assert abs(ev[j] - guess) <= 1e-8 * abs(guess)
    assert res.resid[0, j] < 1e-10
    vec = sv.normalize(res.evecs[0][:, j], 5 * inp.ny)
    g = gpu_solver.grid(pt)
    mine = np.array([[g["y"][i]] + sum(([vec[i * 5 + k].real, vec[i * 5 + k].imag] for k in range(5)), [])
                     for i in range(inp.ny - 1, -1, -1)])
    d = np.abs(mine - tab)
    d[np.isclose(tab, 1e99)] = 0.0
    _say(f"  table max abs diff {d.max():.2e}")
    assert d.max() <= 1e-8
    # second half of the companion vector is alpha * phi (spatial.f90:1007-1016)
    z = res.evecs[0][:, j]
    n = 5 * inp.ny
    assert np.linalg.norm(z[n:] - ev[j] * z[:n]) <= 1e-10 * np.linalg.norm(z[n:])


@pytest.mark.parametrize("mode", [0, 1])
def test_batch_matches_qz(gpu_solver, mode):
    """A small seeded batch of the benchmark workload (compressible Blasius,
    M = 0.5, Sutherland): all six pairs of every point converge, each
    matches the nearest ZGGEV eigenvalue of the oracle pencil to 1e-8,
    eigenvectors to 1e-6 after normalisation, and the nearest-sigma ordering
    agrees with QZ."""
    import bench
    from oracle import stab_oracle as so
    _say(f"batch vs QZ mode {mode}")
    tab, al, be, re, ny = bench.make_workload(6, 77)
    pts = bench.make_points(al, be, re, ny, 6)
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    res = gpu_solver.solve(pts, mode)
    nconv = 0
    for i, p in enumerate(pts):
        txt = (f"1\n288.15\n0.5 {float(re[i])!r} 0.72\n{int(ny[i])} {2.0 / 46.0!r} 50\n1\n1\n0 0 0 0\n1\n"
               f"{float(al[i])!r} 0\n{float(be[i])!r} 0\n0\n")
        pp = so.parse_input(txt)
        pb = so.setup(pp, tab[:, 0], tab[:, 1:6].copy(), spatial=False)
        A0, B0 = so.temporal_pencil(pb)
        wq, vq = so.qz(A0, B0)
        fin = np.isfinite(wq) & (np.abs(wq) > 1e-8)
        wq, vq = wq[fin], vq[:, fin]
        for k in range(6):
            if res.resid[i, k] >= 1e-10:
                continue
            ev = res.evals[i, k]
            q = int(np.argmin(np.abs(wq - ev)))
            assert abs(wq[q] - ev) <= 1e-8 * abs(wq[q]), (i, k, ev, wq[q])
            a = so.normalize_evec(vq[:, q], 5 * p.ny)
            b = so.normalize_evec(res.evecs[i][:, k], 5 * p.ny)
            assert np.max(np.abs(a - b)) <= 1e-6, (i, k)
            nconv += 1
    _say(f"  converged pairs checked: {nconv} of {6 * len(pts)}")
    assert nconv == 6 * len(pts)     # every pair of this seeded batch converges
