"""CPU tests that pin the oracle (and the host-side product code) without a GPU.

The oracle is pinned by the reference's own golden files (thesis/TStest
time.ref/space.ref, TStest/README.md, test/space.1, all in tests/golden) and
by what the mathematics fixes: closed forms, polynomial exactness, finite
differences.
"""
import math
import os
import re

import numpy as np
import pytest

from conftest import golden, ROOT
from oracle import stab_oracle as so


def _table_from(pb, vec):
    p = pb.p
    rows = []
    for i in range(p.ny - 1, -1, -1):
        r = [pb.y[i]]
        for k in range(5):
            r += [vec[i * 5 + k].real, vec[i * 5 + k].imag]
        rows.append(r)
    return np.array(rows)


def _ndiff_ok(mine, ref, tol=1e-8):
    d = np.abs(mine - ref)
    # the top node is at y = 1e99 in both files
    d[np.isclose(ref, 1e99)] = 0.0
    return d.max() <= tol, d.max()


# --- end-to-end goldens ------------------------------------------------------

def test_thesis_temporal_time_ref(ts_profile):
    """thesis/TStest/run.sh:12-17,38-39: omega and the eigenfunction table."""
    p = so.parse_input(open(golden("thesis_TStest", "temporal.inp")).read())
    pb, A0, B0, w, v = so.run_temporal(p, *ts_profile)
    hdr, tab = so.parse_table(open(golden("thesis_TStest", "time.ref")).read())
    j = so.select_nearest(w, hdr["Omega"])
    assert abs(w[j] - hdr["Omega"]) <= 1e-8 * abs(hdr["Omega"])
    ok, dmax = _ndiff_ok(_table_from(pb, so.normalize_evec(v[:, j], 5 * p.ny)), tab)
    assert ok, dmax


def test_thesis_spatial_space_ref(ts_profile):
    """thesis/TStest/run.sh:21-26,35-36: alpha and the eigenfunction table."""
    p = so.parse_input(open(golden("thesis_TStest", "spatial.inp")).read())
    pb, Cs, alp, v = so.run_spatial(p, *ts_profile)
    hdr, tab = so.parse_table(open(golden("thesis_TStest", "space.ref")).read())
    j = so.select_nearest(alp, hdr["Alpha"])
    assert abs(alp[j] - hdr["Alpha"]) <= 1e-8 * abs(hdr["Alpha"])
    ok, dmax = _ndiff_ok(_table_from(pb, so.normalize_evec(v[:, j], 5 * p.ny)), tab)
    assert ok, dmax


def test_tstest_readme_spatial_n64(ts_profile):
    """TStest/README.md:10 (N=64, in-repo profile.0)."""
    txt = open(golden("TStest", "input.dat")).read().replace("96 1.0 0.0", "64 1.0 0.0")
    p = so.parse_input(txt)
    _, _, alp, _ = so.run_spatial(p, *ts_profile)
    ref = 2.2804739410500E-01 - 6.5163146761218E-03j
    j = so.select_nearest(alp, ref)
    assert abs(alp[j] - ref) <= 1e-8 * abs(ref)


def test_test_space1_n32(ts_profile):
    """test/space.1 (N=32, test/input.dat): header alpha and the table."""
    p = so.parse_input(open(golden("test", "input.dat")).read())
    pb, _, alp, v = so.run_spatial(p, *ts_profile)
    hdr, tab = so.parse_table(open(golden("test", "space.1")).read())
    j = so.select_nearest(alp, hdr["Alpha"])
    assert abs(alp[j] - hdr["Alpha"]) <= 1e-8 * abs(hdr["Alpha"])
    ok, dmax = _ndiff_ok(_table_from(pb, so.normalize_evec(v[:, j], 5 * p.ny)), tab)
    assert ok, dmax


def test_qz_matches_reference_method(ts_profile):
    """ZGGEV on (A0, B0) and ZGESV+ZGEEV give the same finite, nonzero spectrum."""
    txt = open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", "32 1.0 0.0")
    p = so.parse_input(txt)
    pb, A0, B0, w, _ = so.run_temporal(p, *ts_profile)
    wq, _ = so.qz(A0, B0, right=False)
    wq = wq[np.isfinite(wq)]
    for x in w[np.abs(w) > 1e-6]:
        assert np.min(np.abs(wq - x)) <= 1e-7 * max(abs(x), 1.0)


# --- closed forms --------------------------------------------------------------

def test_sgengrid_matches_time_ref_y(ts_profile):
    """Infinite algebraic mapping (sgengrid.f90:28-38) against column 1 of time.ref."""
    p = so.parse_input(open(golden("thesis_TStest", "temporal.inp")).read())
    y, eta, deta, d2eta = so.sgengrid(p)
    _, tab = so.parse_table(open(golden("thesis_TStest", "time.ref")).read())
    yref = tab[::-1, 0]
    fin = yref < 1e98
    assert np.allclose(y[fin], yref[fin], rtol=1e-12, atol=1e-16)
    assert y[0] == 1e99


@pytest.mark.parametrize("ymax", [0.0, 50.0])
def test_sgengrid_metrics_are_derivatives(ymax):
    """deta = d eta/dy and d2eta = d2 eta/dy2 of the mapping (finite differences)."""
    p = so.Params(ny=33, yi=2.0 / 46.0 if ymax else 1.0, ymax=ymax)
    y, eta, deta, d2eta = so.sgengrid(p)
    if ymax:
        f = lambda e: ymax * p.yi * (1 + e) / (1 + 2 * p.yi - e)
    else:
        f = lambda e: p.yi * (1 + e) / (1 - e)
    for i in range(2, 30):
        # complex step for y'(eta); centred difference scaled to the pole distance for y''
        dy = (f(eta[i] + 1e-30j)).imag / 1e-30
        h = 1e-4 * (1.0 - eta[i])
        d2y = (f(eta[i] + h) - 2 * f(eta[i]) + f(eta[i] - h)) / h ** 2
        assert deta[i] == pytest.approx(1.0 / dy, rel=1e-12)
        # d2eta/dy2 = -y''(eta) / y'(eta)^3
        assert d2eta[i] == pytest.approx(-d2y / dy ** 3, rel=1e-5)


@pytest.mark.parametrize("n", [8, 31, 64])
def test_chebyd_exact_on_polynomials(n):
    """D differentiates polynomials of degree <= n exactly; rows sum to zero;
    D_00 = (2 n^2 + 1)/6 (Chebyshev-Gauss-Lobatto corner)."""
    d = so.chebyd(n)
    x = np.cos(np.pi * np.arange(n + 1) / n)
    for k in (1, 2, min(5, n), n):
        f = x ** k
        df = k * x ** (k - 1)
        assert np.max(np.abs(d @ f - df)) <= 1e-10 * n ** 2 * max(1, k)
    assert np.max(np.abs(d.sum(axis=1))) <= 1e-10 * n ** 2
    assert d[0, 0] == pytest.approx((2 * n * n + 1) / 6.0, rel=1e-12)
    assert d[n, n] == pytest.approx(-(2 * n * n + 1) / 6.0, rel=1e-12)


def test_spline_reproduces_quadratics():
    """With the cantilever end condition (FDP(1) = FDP(2)) the spline of a
    quadratic is the quadratic itself."""
    x = np.sort(np.random.default_rng(1).uniform(0, 3, 25))
    y = 0.3 - 1.2 * x + 0.7 * x ** 2
    f = so.spline(x, y)
    assert np.allclose(f, 1.4, rtol=1e-10)
    for xx in np.linspace(x[0], x[-1], 57):
        assert so.speval(x, y, f, xx) == pytest.approx(0.3 - 1.2 * xx + 0.7 * xx ** 2, rel=1e-12, abs=1e-12)


def test_spline_matches_dense_tridiagonal_solve():
    """Brute force: the same tridiagonal system solved densely."""
    rng = np.random.default_rng(2)
    x = np.sort(rng.uniform(0, 1, 9))
    y = rng.standard_normal(9)
    f = so.spline(x, y)
    n = 9
    h = np.diff(x)
    A = np.zeros((n - 2, n - 2))
    r = np.zeros(n - 2)
    for i in range(1, n - 1):
        k = i - 1
        A[k, k] = 2 * (h[i - 1] + h[i])
        if k > 0:
            A[k, k - 1] = h[i - 1]
        if k < n - 3:
            A[k, k + 1] = h[i]
        r[k] = 6 * ((y[i + 1] - y[i]) / h[i] - (y[i] - y[i - 1]) / h[i - 1])
    A[0, 0] += h[0]          # FDP(1) = FDP(2)
    A[-1, -1] += h[-1]       # FDP(N) = FDP(N-1)
    fi = np.linalg.solve(A, r)
    assert np.allclose(f[1:-1], fi, rtol=1e-12, atol=1e-12)
    assert f[0] == f[1] and f[-1] == f[-2]


def test_sutherland_derivatives():
    p = so.Params(mattyp=1, Pr=0.72, datmat=(1.715336725523065e-05, 273.0, 110.4))
    t = np.array([150.0, 288.15, 600.0])
    h = 1e-3
    mu, _, _, dmu, d2mu, *_ = so.getmat(p, t)
    mup = so.getmat(p, t + h)[0]
    mum = so.getmat(p, t - h)[0]
    assert np.allclose(dmu, (mup - mum) / (2 * h), rtol=1e-8)
    assert np.allclose(d2mu, (mup - 2 * mu + mum) / h ** 2, rtol=1e-5)
    # mu(273) = mu_ref
    assert so.getmat(p, np.array([273.0]))[0][0] == pytest.approx(1.715336725523065e-05, rel=1e-15)


def test_input_quirk_te_equals_t0():
    """input.f90:24-25 computes Te before Ma is read: Te = T0 as entered."""
    txt = "1\n300.0\n2.0 1000 0.72\n64 1 0\n1\n1\n0 0 0 0\n1\n0.1 0\n0 0\n0\n"
    p = so.parse_input(txt)
    assert p.Te == 300.0
    mu = 1.715336725523065e-05 * 300 / 273 * math.sqrt(300 / 273) * (273 + 110.4) / (300 + 110.4)
    assert p.rmue == pytest.approx(mu, rel=1e-15)
    assert p.cone == pytest.approx(mu * 1003.1 / 0.72, rel=1e-15)


def test_pencil_residual_of_reference_eigenfunction(ts_profile):
    """The eigenfunction printed in time.ref (14 digits) nearly satisfies our
    pencil: ||(A0 - w B0) x|| <= 1e-9 ||A0|| ||x||."""
    p = so.parse_input(open(golden("thesis_TStest", "temporal.inp")).read())
    pb = so.setup(p, *ts_profile, spatial=False)
    A0, B0 = so.temporal_pencil(pb)
    hdr, tab = so.parse_table(open(golden("thesis_TStest", "time.ref")).read())
    x = np.zeros(5 * p.ny, complex)
    for row, i in zip(tab, range(p.ny - 1, -1, -1)):
        for k in range(5):
            x[i * 5 + k] = row[1 + 2 * k] + 1j * row[2 + 2 * k]
    r = (A0 - hdr["Omega"] * B0) @ x
    assert np.linalg.norm(r) <= 1e-9 * np.linalg.norm(A0) * np.linalg.norm(x)


# --- host-side product code (no GPU) ------------------------------------------

@pytest.mark.parametrize("path", ["thesis_TStest/temporal.inp", "thesis_TStest/spatial.inp", "TStest/input.dat",
                                  "test/input.dat", "thesis_CFtest/temporal.inp", "thesis_CFtest/spatial.inp"])
def test_product_input_reader_matches_oracle(path):
    from stab_amd import solver as sv
    text = open(golden(*path.split("/"))).read()
    a, b = sv.read_input(text), so.parse_input(text)
    for f in ("mattyp", "Ma", "Re", "Pr", "ny", "yi", "ymax", "ievec", "top", "wall", "wallt", "curve", "itype",
              "alpha", "beta", "omega", "ind", "x", "Te"):
        assert getattr(a, f) == getattr(b, f), f


def test_blasius_solver_matches_profile0():
    """Compressible Blasius (M=0.3, Pr=1, C=1, adiabatic) against the reference's
    TStest/profile.0 at the same y: u to 1e-10, T and rho to 1e-11."""
    from scipy.optimize import brentq
    from stab_amd import meanflow as mf
    raw = np.loadtxt(golden("TStest", "profile.0"))
    spec = mf.BlasiusSpec(Ma=0.3, Pr=1.0, law="chapman")
    sol, x = mf.solve(spec)
    assert x[0] == pytest.approx(0.4696, abs=1e-4)       # thesis/TStest/blasius.inp:5
    assert x[1] == pytest.approx(1.018, rel=1e-9)        # T_aw/T_e = 1 + (gamma-1)/2 M^2 for Pr = 1
    dstar = sol.y[6, -1]
    for k in range(0, 2000, 97):
        yt = raw[k, 0] * dstar
        e = brentq(lambda et: sol.sol(et)[5] - yt, 0, 19.999, xtol=1e-15)
        s = sol.sol(e)
        assert abs(s[1] - raw[k, 2]) < 1e-10
        assert abs(s[3] - raw[k, 5]) < 1e-11
        assert abs(1 / s[3] - raw[k, 1]) < 1e-11


def test_evec_dat_round_trip(tmp_path):
    from stab_amd import solver as sv
    inp = sv.read_input(open(golden("thesis_TStest", "temporal.inp")).read())
    ny = inp.ny
    g = dict(y=np.linspace(0, 1, ny), eta=np.linspace(1, -1, ny), deta=np.ones(ny), d2eta=np.zeros(ny))
    ev = np.array([0.1 + 0.01j, 0.2 - 0.03j])
    vec = (np.arange(5 * ny * 2) + 1j).reshape(5 * ny, 2, order="F")
    path = str(tmp_path / "evec.dat")
    sv.write_evec_dat(path, inp, g, ev, vec)
    d = so.read_evec_dat(path)
    assert d["ny"] == ny and d["itype"] == 1 and d["ievec"] == 1
    assert np.array_equal(d["evals"], ev)
    assert np.array_equal(d["y"], g["y"])
    assert np.array_equal(d["evec"], vec)
    d2 = sv.read_evec_dat(path)
    assert np.array_equal(d2["evec"], vec)


def test_eigenfunction_text_format(ts_profile):
    """time.N layout: 4 header lines and 11 columns in 1pe21.13E3 (getevec.f90:196-227)."""
    from stab_amd import solver as sv
    p = so.parse_input(open(golden("thesis_TStest", "temporal.inp")).read())
    pb, A0, B0, w, v = so.run_temporal(p, *ts_profile)
    hdr, _ = so.parse_table(open(golden("thesis_TStest", "time.ref")).read())
    j = so.select_nearest(w, hdr["Omega"])
    d = dict(ny=p.ny, Re=p.Re, Ma=p.Ma, Pr=p.Pr, itype=1, evals=w, alpha=p.alpha, beta=p.beta, omega=p.omega,
             y=pb.y, evec=v)
    text = sv.eigenfunction_text(d, j)
    ref = open(golden("thesis_TStest", "time.ref")).read().splitlines()
    mine = text.splitlines()
    assert mine[0] == ref[0]
    assert mine[2] == ref[2] and mine[3] == ref[3]
    assert len(mine) == len(ref)
    assert all(len(a.split()) == 11 for a in mine[4:])
    assert mine[4][:22] == ref[4][:22]


def test_library_exports_every_declared_symbol():
    """The built C-ABI library loads and exports each function of include/stab_hip.h."""
    import ctypes
    from stab_amd import _lib
    header = open(os.path.join(ROOT, "include", "stab_hip.h")).read()
    names = set(re.findall(r"^\s*(?:int|size_t|const char \*)\s*(stab_\w+)\s*\(", header, re.M))
    assert names >= set(_lib.EXPORTS)
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n


def test_leading_mode_rule_picks_thesis_modes(ts_profile):
    """The leading-mode rule (discrete: decays in the freestream, off the
    c = 1 continuum; least stable) applied to the QZ spectrum (|eigenvalue|
    < 1; beyond lie under-resolved modes of |alpha| ~ 1e6) picks the modes
    the reference's thesis cases select with getevec -v
    (thesis/TStest/time.ref, space.ref)."""
    ym, vt = ts_profile
    # temporal: A0 x = w B0 x
    p = so.parse_input(open(golden("thesis_TStest", "temporal.inp")).read())
    pb = so.setup(p, ym, vt, spatial=False)
    A0, B0 = so.temporal_pencil(pb)
    w, V = so.qz(A0, B0)
    hdr, _ = so.parse_table(open(golden("thesis_TStest", "time.ref")).read())
    phys = [k for k in range(len(w)) if np.isfinite(w[k]) and 1e-9 < abs(w[k]) < 1.0]   # under-resolved modes excluded
    j = so.leading_mode(1, w, V, p.ny, alpha_r=p.alpha.real, candidates=phys)
    assert j >= 0 and abs(w[j] - hdr["Omega"]) <= 1e-8 * abs(hdr["Omega"])
    # the continuum modes nearest Im = 0 are rejected, and they exist above the TS mode
    near = [k for k in phys if w[k].imag > w[j].imag]
    assert all(not so.is_discrete(1, w[k], so.farfield_ratio(V[:, k], p.ny), alpha_r=p.alpha.real) for k in near)
    # spatial: the companion pencil, alpha = eigenvalue
    p = so.parse_input(open(golden("thesis_TStest", "spatial.inp")).read())
    pb = so.setup(p, ym, vt, spatial=True)
    C0, C1, C2 = so.spatial_pencil(pb)
    L, M = so.spatial_linearization(C0, C1, C2)
    w, V = so.qz(L, M)
    hdr, _ = so.parse_table(open(golden("thesis_TStest", "space.ref")).read())
    phys = [k for k in range(len(w)) if np.isfinite(w[k]) and 1e-9 < abs(w[k]) < 1.0]
    j = so.leading_mode(2, w, V[:5 * p.ny], p.ny, omega_r=p.omega.real, candidates=phys)
    assert j >= 0 and abs(w[j] - hdr["Alpha"]) <= 1e-8 * abs(hdr["Alpha"])


def test_discrete_rule_rejects_continuum_and_growing_farfield():
    """Closed-form cases of the rule: a vector that is largest in the outer
    eighth is not discrete; one that decays is, unless its phase speed is on
    the continuum."""
    ny = 16
    decay = np.exp(-np.linspace(0, 8, ny))[::-1].repeat(5)     # node 1 = freestream: small there
    grow = decay[::-1].copy()
    assert so.farfield_ratio(decay, ny) < 0.01 and so.farfield_ratio(grow, ny) == 1.0
    assert so.is_discrete(1, 0.03 - 0.001j, so.farfield_ratio(decay, ny), alpha_r=0.1)
    assert not so.is_discrete(1, 0.03 - 0.001j, so.farfield_ratio(grow, ny), alpha_r=0.1)
    assert not so.is_discrete(1, 0.1 - 0.001j, so.farfield_ratio(decay, ny), alpha_r=0.1)   # c = 1
    assert not so.is_discrete(2, 0.1 - 0.001j, so.farfield_ratio(decay, ny), omega_r=0.1)   # c = 1
    assert not so.is_discrete(2, -0.01 - 0.9j, so.farfield_ratio(decay, ny), omega_r=0.1)   # upstream


def test_ndiff_equivalent_on_goldens(ts_profile, tmp_path):
    """stab_amd.ndiff: the oracle's thesis eigenfunction text agrees with
    time.ref under -abserr 1e-8 and a perturbation of 2e-8 in one field is
    caught at the right place."""
    from stab_amd import ndiff
    ref = open(golden("thesis_TStest", "time.ref")).read()
    ok, worst, bad = ndiff.compare(ref, ref, abserr=1e-8)
    assert ok and worst == 0.0
    rows = ref.splitlines()
    k = next(i for i, l in enumerate(rows) if not l.startswith("#"))
    f = rows[k + 3].split()
    f[4] = "%.13E" % (float(f[4]) + 2e-8)
    rows[k + 3] = " ".join(f)
    ok, worst, bad = ndiff.compare(ref, "\n".join(rows), abserr=1e-8)
    assert not ok and bad == (3, 4) and abs(worst - 2e-8) < 1e-12
    p1, p2 = tmp_path / "a", tmp_path / "b"
    p1.write_text(ref)
    p2.write_text(ref)
    assert ndiff.main(["-abserr", "1e-8", str(p1), str(p2)]) == 0
