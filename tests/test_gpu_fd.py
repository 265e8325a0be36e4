"""GPU parity of the finite-difference paths (fd_temporal.f90, itype 3;
fd_spatial.f90, itype 4) through the C ABI, against the oracle on thesis/TStest's
fd_temporal.inp and fd_spatial.inp (yi = 1, ymax = 50) at N = 10 to 128.

Tolerances: grid 1e-14 relative, derivative rows bitwise, pencil entries
1e-13 relative to the matrix max, eigenvalues 1e-9 relative to ZGEEV on the
oracle's pencil, normalised eigenvectors 1e-6, FP64 residuals 1e-10.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

TS_OMEGA = 1.1467880189410E-001 + 2.3844535276592E-003j      # thesis/TStest/README.md:25


def _say(msg):
    print(msg, flush=True)


def _text(ny):
    return open(golden("thesis_TStest", "fd_temporal.inp")).read().replace("128 1.0 50.0", f"{ny} 1.0 50.0")


def _point(ny, shift, nev=4):
    from stab_amd import solver as sv
    inp = sv.read_input(_text(ny))
    assert inp.itype == 3
    return inp, sv.make_point(inp, shift, nev)


@pytest.fixture(scope="module")
def oracle128(ts_profile):
    from oracle import stab_oracle as so
    p = so.parse_input(_text(128))
    return (p,) + so.run_fd_temporal(p, *ts_profile)


@pytest.mark.parametrize("ny", [10, 64, 128])
def test_fd_grid_and_operators(gpu_solver, ts_profile, ny):
    """ogengrid, the interpolated mean flow and the pencil's D1/D2 rows (the
    latter bitwise: both sides form coefficient times 1/deta or 1/deta^2)."""
    from oracle import stab_oracle as so
    _say(f"fd grid ny={ny}")
    gpu_solver.set_profile(0, *ts_profile)
    _, pt = _point(ny, 0.1 + 0j)
    g = gpu_solver.grid(pt)
    p = so.parse_input(_text(ny))
    y, eta, deta, d2eta, dy = so.ogengrid(p)
    assert np.allclose(g["y"], y, rtol=1e-14, atol=0)
    assert np.array_equal(g["eta"], eta)
    assert np.allclose(g["deta"], deta, rtol=1e-14, atol=0)
    assert np.allclose(g["d2eta"], d2eta, rtol=1e-14, atol=0)
    assert np.max(np.abs(g["vm"] - so.getmean(y, *ts_profile))) < 1e-13
    d1, d2 = so.fd_operators(ny, dy)
    assert np.array_equal(g["d1"], d1)
    assert np.array_equal(g["d2"], d2)


@pytest.mark.parametrize("ny", [16, 128])
def test_fd_pencil_matches_oracle(gpu_solver, ts_profile, ny):
    """A0, B0 (fd_temporal.f90:582-890) assembled on the device."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    _say(f"fd pencil ny={ny}")
    gpu_solver.set_profile(0, *ts_profile)
    _, pt = _point(ny, 0.1 + 0j)
    pb = so.fd_setup(so.parse_input(_text(ny)), *ts_profile)
    for which, ref in zip((_lib.MAT_A0, _lib.MAT_B0), so.fd_temporal_pencil(pb)):
        got = gpu_solver.pencil(pt, which)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        _say(f"  matrix {which}: max rel err {err:.2e}")
        assert err < 1e-13


@pytest.mark.parametrize("mode", [0, 1])
def test_fd_solve_matches_zgeev(gpu_solver, ts_profile, oracle128, mode):
    """The nev pairs nearest the shift (plain FP64 and mixed FP32-LU paths)
    equal the oracle's ZGESV + ZGEEV eigenvalues nearest it; the TS wave is
    among them."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    p, pb, A0, B0, w, v = oracle128
    _say(f"fd solve N=128 mode {mode}")
    gpu_solver.set_profile(0, *ts_profile)
    shift = 0.35 * p.alpha
    _, pt = _point(128, shift, 6)
    res = gpu_solver.solve([pt], mode)
    ev = res.evals[0, :6]
    _say(f"  info {res.info[0]} evals {ev}")
    assert res.info[0] == 0
    fin = np.where(np.isfinite(w) & (np.abs(w) > 1e-9))[0]
    near = fin[np.argsort(np.abs(w[fin] - shift))[:6]]
    for k in range(6):
        j = near[np.argmin(np.abs(w[near] - ev[k]))]
        assert abs(ev[k] - w[j]) <= 1e-9 * abs(w[j])
        assert res.resid[0, k] < 1e-10
        x = so.normalize_evec(v[:, j], 5 * p.ny)
        assert np.max(np.abs(sv.normalize(res.evecs[0][:, k], 5 * p.ny) - x)) <= 1e-6
    assert np.min(np.abs(ev - TS_OMEGA)) < 1e-4


@pytest.mark.parametrize("method", ["krylov", "invit"])
def test_fd_leading_mode(gpu_solver, ts_profile, oracle128, method):
    """stab_solve_leading on a finite-difference point: the least stable
    discrete mode among the 8 eigenvalues nearest the fallback shift, with the
    farfield taken over the outer half of the domain (the last nodes)."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    p, pb, A0, B0, w, v = oracle128
    _say(f"fd leading {method}")
    gpu_solver.set_profile(0, *ts_profile)
    fb = 0.35 * p.alpha
    fin = np.where(np.isfinite(w) & (np.abs(w) > 1e-9))[0]
    near = list(fin[np.argsort(np.abs(w[fin] - fb))[:8]])
    nodes = so.fd_farfield_nodes(pb.y, p.ymax)
    jl = so.leading_mode(1, w, v, p.ny, p.alpha.real, candidates=near, nodes=nodes)
    assert abs(w[jl] - TS_OMEGA) < 1e-4
    _, pt = _point(128, fb, 8)
    if method == "krylov":
        out = gpu_solver.solve_leading([pt], _lib.MODE_PLAIN_F64, _lib.LEAD_KRYLOV, want_vectors=True)
    else:
        out = gpu_solver.solve_leading([pt], _lib.MODE_MIXED, _lib.LEAD_INVIT, want_vectors=True,
                                       shifts=[w[jl] * (1.0 + 1e-3)], fallback_shifts=[fb])
    _say(f"  info {out['info']} eval {out['evals'][0]} ff {out['farfield'][0]:.3e} iters {out['iters']}")
    assert out["info"][0] == _lib.INFO_OK
    assert abs(out["evals"][0] - w[jl]) <= 1e-9 * abs(w[jl])
    assert out["resid"][0] < 1e-10
    ff = so.farfield_ratio(v[:, jl], p.ny, nodes)
    assert out["farfield"][0] == pytest.approx(ff, rel=1e-5)
    assert np.max(np.abs(out["evecs"][0] - so.normalize_evec(v[:, jl], 5 * p.ny))) <= 1e-6


def test_fd_and_chebyshev_points_share_a_batch(gpu_solver, ts_profile):
    """A Chebyshev and a finite-difference point of the same ny in one batch
    keep their own operator sets: each gives what it gives alone, bitwise."""
    from stab_amd import solver as sv
    _say("fd + chebyshev batch")
    gpu_solver.set_profile(0, *ts_profile)
    cheb = sv.read_input(open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", "64 1.0 0.0"))
    pc = sv.make_point(cheb, TS_OMEGA, 4)
    _, pf = _point(64, TS_OMEGA, 4)
    both = gpu_solver.solve([pc, pf], 0)
    alone = [gpu_solver.solve([q], 0) for q in (pc, pf)]
    for i in range(2):
        assert np.array_equal(both.evals[i, :4], alone[i].evals[0, :4])
        assert np.array_equal(both.evecs[i], alone[i].evecs[0])
    assert abs(both.evals[0, :4] - TS_OMEGA).min() < 1e-6
    assert abs(both.evals[1, :4] - TS_OMEGA).min() < 1e-4


# --- finite-difference spatial (fd_spatial.f90, itype 4) ----------------------------

TS_ALPHA = 2.2804739411185E-001 - 6.5163146952733E-003j       # thesis/TStest/README.md:14


def _stext(ny, top=0, wallt=0):
    txt = open(golden("thesis_TStest", "fd_spatial.inp")).read().replace("128 1.0 50.0", f"{ny} 1.0 50.0")
    return txt.replace("0 0 0 0\t", f"{top} 0 {wallt} 0\t")


def _spoint(ny, shift, nev=4, top=0, wallt=0):
    from stab_amd import solver as sv
    inp = sv.read_input(_stext(ny, top, wallt))
    assert inp.itype == 4 and inp.top == top and inp.wallt == wallt
    return inp, sv.make_point(inp, shift, nev)


@pytest.mark.parametrize("ny", [10, 64])
def test_fds_operators(gpu_solver, ts_profile, ny):
    """The spatial rows (central to node ny-3, grad rows at the top), bitwise."""
    from oracle import stab_oracle as so
    _say(f"fd spatial operators ny={ny}")
    gpu_solver.set_profile(0, *ts_profile)
    _, pt = _spoint(ny, 0.2 + 0j)
    g = gpu_solver.grid(pt)
    d1, d2 = so.fd_spatial_operators(ny, 1.0 / float(ny - 1))
    assert np.array_equal(g["d1"], d1)
    assert np.array_equal(g["d2"], d2)


@pytest.mark.parametrize("top,wallt", [(0, 0), (1, 2), (0, 2)])
def test_fds_pencil_matches_oracle(gpu_solver, ts_profile, top, wallt):
    """C0, C1, C2 (fd_spatial.f90:650-1190) for the wall energy row absent
    (wallt 0) and present (wallt 2, its time term with G of node ny-1 or ny-2),
    and the top continuity row absent and present."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    _say(f"fd spatial pencil top={top} wallt={wallt}")
    gpu_solver.set_profile(0, *ts_profile)
    _, pt = _spoint(24, 0.2 + 0j, top=top, wallt=wallt)
    pb = so.fd_setup(so.parse_input(_stext(24, top, wallt)), *ts_profile, spatial=True)
    for which, ref in zip((_lib.MAT_C0, _lib.MAT_C1, _lib.MAT_C2), so.fd_spatial_pencil(pb)):
        got = gpu_solver.pencil(pt, which)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        _say(f"  matrix {which}: max rel err {err:.2e}")
        assert err < 1e-13


@pytest.mark.parametrize("mode", [0, 1])
def test_fds_solve_matches_zgeev(gpu_solver, ts_profile, mode):
    """The alpha nearest omega / 0.35 (plain FP64 and mixed FP32-LU), the TS
    wave, equals the oracle's companion ZGEEV eigenvalue.  One pair: the next
    ones lie 0.144 from the shift in a cluster of entropy/vorticity pairs that
    Pr = 1 makes nearly degenerate (distances equal to 1e-5), which no
    nearest-k window separates."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    _say(f"fd spatial solve N=96 mode {mode}")
    p = so.parse_input(_stext(96))
    pb, _, alp, v = so.run_fd_spatial(p, *ts_profile)
    gpu_solver.set_profile(0, *ts_profile)
    shift = p.omega / 0.35
    _, pt = _spoint(96, shift, 1)
    res = gpu_solver.solve([pt], mode)
    ev = res.evals[0, :1]
    _say(f"  info {res.info[0]} evals {ev} resid {res.resid[0, :1]}")
    assert res.info[0] == 0
    fin = np.where(np.isfinite(alp) & (np.abs(alp) > 1e-9))[0]
    near = fin[np.argsort(np.abs(alp[fin] - shift))[:1]]
    for k in range(1):
        j = near[np.argmin(np.abs(alp[near] - ev[k]))]
        assert abs(ev[k] - alp[j]) <= 1e-9 * abs(alp[j])
        assert res.resid[0, k] < 1e-10
        x = so.normalize_evec(v[:, j], 5 * p.ny)
        assert np.max(np.abs(sv.normalize(res.evecs[0][:, k], 5 * p.ny) - x)) <= 1e-6
    assert np.min(np.abs(ev - TS_ALPHA)) < 1e-6


def test_fds_leading_mode(gpu_solver, ts_profile):
    """stab_solve_leading (Krylov-Schur, two pairs) on a finite-difference
    spatial point returns the TS wave, the oracle's leading discrete mode
    among the two alphas nearest the shift, farfield over the outer half of
    the domain; inverse iteration from a perturbed shift reaches it too."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    _say("fd spatial leading")
    p = so.parse_input(_stext(96))
    pb, _, alp, v = so.run_fd_spatial(p, *ts_profile)
    shift = p.omega / 0.35
    fin = np.where(np.isfinite(alp) & (np.abs(alp) > 1e-9))[0]
    near = list(fin[np.argsort(np.abs(alp[fin] - shift))[:2]])
    nodes = so.fd_farfield_nodes(pb.y, p.ymax)
    jl = so.leading_mode(2, alp, v[:5 * p.ny], p.ny, omega_r=p.omega.real, candidates=near, nodes=nodes)
    gpu_solver.set_profile(0, *ts_profile)
    _, pt = _spoint(96, shift, 2)
    runs = [gpu_solver.solve_leading([pt], _lib.MODE_PLAIN_F64, _lib.LEAD_KRYLOV, want_vectors=True),
            gpu_solver.solve_leading([pt], _lib.MODE_MIXED, _lib.LEAD_INVIT, want_vectors=True,
                                     shifts=[alp[jl] * (1.0 + 1e-3)], fallback_shifts=[shift])]
    for out in runs:
        _say(f"  info {out['info']} eval {out['evals'][0]} ff {out['farfield'][0]:.3e}")
        assert out["info"][0] == _lib.INFO_OK
        assert abs(out["evals"][0] - alp[jl]) <= 1e-9 * abs(alp[jl])
        assert abs(out["evals"][0] - TS_ALPHA) < 1e-6
        assert out["farfield"][0] == pytest.approx(so.farfield_ratio(v[:, jl], p.ny, nodes), rel=1e-5)


@pytest.mark.parametrize("case", ["ny9", "ymax", "itype5"])
def test_fd_arguments_are_checked(gpu_solver, ts_profile, case):
    """The finite-difference closures need ny >= 10 and ogengrid needs
    ymax > 2 yi (ogengrid.f90:21 divides by ymax - 2 yi); itypes past 4 are
    not solves of this library.  Each is refused with STAB_E_ARG before any
    kernel runs, and the context stays usable."""
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"fd argument check: {case}")
    gpu_solver.set_profile(0, *ts_profile)
    _, pt = _point(64, TS_OMEGA, 2)
    if case == "ny9":
        pt.ny, msg = 9, "below 10"
    elif case == "ymax":
        pt.ymax, msg = 2.0 * pt.yi, "ymax > 2 yi"
    else:
        pt.itype, msg = 5, "not 1 to 4"
    with pytest.raises(sv.StabError, match=msg):
        gpu_solver.solve([pt], _lib.MODE_PLAIN_F64)
    _, ok = _point(64, TS_OMEGA, 2)
    res = gpu_solver.solve([ok], _lib.MODE_PLAIN_F64)
    assert res.info[0] == _lib.INFO_OK


@pytest.mark.parametrize("itype", [3, 4])
def test_fd_pencils_sutherland_oblique(gpu_solver, itype):
    """Every tensor term with the finite-difference rows: Sutherland mu(T),
    Pr = 0.72, M = 0.5, beta != 0 on a computed compressible Blasius profile
    (temporal A0/B0, spatial C0/C1/C2 with wallt 2 and top 1), device against
    the oracle at 1e-13 of each matrix's largest entry."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    from stab_amd import meanflow as mf
    from stab_amd import solver as sv
    _say(f"fd pencils, Sutherland, oblique, itype {itype}")
    te = 288.15
    tab = mf.profile(mf.BlasiusSpec(Ma=0.5, Pr=0.72, law="sutherland", Te=te, eta_max=20.0, nrows=1200))
    ym, vt = tab[:, 0], tab[:, 1:6].copy()
    gpu_solver.set_profile(0, ym, vt)
    head = f"1\n{te}\n0.5 1200 0.72\n24 {2.0 / 46.0!r} 50\n1\n1\n"
    if itype == 3:
        txt = head + "0 0 0 0\n3\n0.21 0.003\n0.12 0\n0\n"
    else:
        txt = head + "1 0 2 0\n4\n0.06 0.001\n0.12 0\n0\n0\n"
    pt = sv.make_point(sv.read_input(txt), 0.1 + 0j, 2)
    p = so.parse_input(txt)
    if itype == 3:
        refs = zip((_lib.MAT_A0, _lib.MAT_B0), so.fd_temporal_pencil(so.fd_setup(p, ym, vt)))
    else:
        refs = zip((_lib.MAT_C0, _lib.MAT_C1, _lib.MAT_C2), so.fd_spatial_pencil(so.fd_setup(p, ym, vt, spatial=True)))
    for which, ref in refs:
        got = gpu_solver.pencil(pt, which)
        err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
        _say(f"  matrix {which}: max rel err {err:.2e}")
        assert err < 1e-13
