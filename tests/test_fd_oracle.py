"""The oracle's finite-difference temporal path (fd_temporal.f90, itype 3).

The reference ships no eigenvalue for this path, so it is pinned in pieces
and end to end against the Chebyshev golden:
  * every derivative row against polynomials: the interior rows to their
    fourth order, the closures to their third (a mistyped digit of a
    Carpenter coefficient breaks a moment condition), the two
    second-derivative rows whose signs the reference flips against the
    central row they are made from, and the mirror symmetry of the closures;
  * the grid against its closed form (y(0) = 0, y(1) = ymax, y(1/2) = yi) and
    its metrics against derivatives of the map;
  * the pencil's rows against the Navier-Stokes equations linearised
    independently (tests/ns_residual.py) on the same derivative rows;
  * the eigenvalue: with the wall energy row replaced by an isothermal wall
    the discretisation converges at fourth order to the Chebyshev value of
    thesis/TStest/README.md (the same flow and profile), and with the
    reference's wall energy row it converges to a Chebyshev pencil given that
    same row.
"""
import numpy as np
import pytest

import ns_residual as ns
from conftest import golden
from oracle import stab_oracle as so

TS_OMEGA = 1.1467880189410E-001 + 2.3844535276592E-003j      # thesis/TStest/README.md:25 (Chebyshev, N = 96)


def _fd_params(ny, ymax=50.0):
    txt = open(golden("thesis_TStest", "fd_temporal.inp")).read()
    return so.parse_input(txt.replace("128 1.0 50.0", f"{ny} 1.0 {ymax}"))


# --- derivative rows ------------------------------------------------------------

def _moments_ok(D, rows, h, deg, order):
    """D applied to eta^k equals the exact derivative of order `order` on the
    given rows for k <= deg (eta in node units scaled by h)."""
    ny = D.shape[0]
    x = np.arange(ny) * h
    worst = 0.0
    for k in range(deg + 1):
        f = x ** k
        if order == 1:
            exact = k * x ** (k - 1) if k >= 1 else np.zeros(ny)
        else:
            exact = k * (k - 1) * x ** (k - 2) if k >= 2 else np.zeros(ny)
        err = np.abs(D[rows] @ f - exact[rows]) / np.maximum(1.0, np.abs(D[rows]) @ np.abs(f))
        worst = max(worst, float(err.max()))
    return worst


def test_fd_first_derivative_rows_exact_on_polynomials():
    """fd_temporal.f90:633-890 rows of D1: central rows exact to degree 4,
    Carpenter closures (stencils.f90:56-90) and their mirror images to degree 3,
    and not to degree 4 (third order)."""
    ny, dy = 24, 1.0 / 23.0
    D1, _ = so.fd_operators(ny, dy)
    assert _moments_ok(D1, list(range(4, ny - 4)), dy, 4, 1) < 1e-13
    clos = [0, 1, 2, 3, ny - 4, ny - 3, ny - 2]
    assert _moments_ok(D1, clos, dy, 3, 1) < 1e-13
    for r in clos:
        assert _moments_ok(D1, [r], dy, 4, 1) > 1e-6
    # the far-end closures are the near-end ones mirrored and negated
    for m in (1, 2, 3):
        assert np.array_equal(D1[ny - 1 - m, ::-1], -D1[m])
    assert not D1[ny - 1].any()


def test_fd_second_derivative_rows():
    """D2: central rows exact to degree 5, the biased (db) and one-sided (dd)
    rows to degree 4; rows ny-4 and ny-3 are the central row with the signs of
    the entries right of the diagonal flipped, except the last one of row ny-3
    (fd_temporal.f90:826-878), so they are not even exact on constants."""
    ny, dy = 24, 1.0 / 23.0
    _, D2 = so.fd_operators(ny, dy)
    assert _moments_ok(D2, list(range(2, ny - 4)), dy, 5, 2) < 1e-12
    assert _moments_ok(D2, [0, 1, ny - 2], dy, 4, 2) < 1e-12
    central = np.array(so.FD_DA) / dy ** 2
    for r, sgn in ((ny - 4, [1, 1, 1, -1, -1]), (ny - 3, [1, 1, 1, -1, 1])):
        assert np.allclose(D2[r, r - 2:r + 3] * np.array(sgn, float), central, rtol=1e-14, atol=0)
        assert abs(D2[r].sum()) > 1.0 / dy ** 2
    assert not D2[ny - 1].any()


def test_fd_grad_matrix_exact_to_fourth_order():
    """grad.f90:47-85: every row exact on polynomials of degree <= 4."""
    ny, dy = 17, 1.0 / 16.0
    G = so.fd_grad_matrix(ny, dy)
    assert _moments_ok(G, list(range(ny)), dy, 4, 1) < 1e-13
    assert _moments_ok(G, [0], dy, 5, 1) > 1e-6


@pytest.mark.parametrize("yi,ymax", [(1.0, 50.0), (2.0, 30.0)])
def test_ogengrid_closed_form(yi, ymax):
    """ogengrid.f90:17-48: y(0) = 0, y(1) = ymax, y(1/2) = yi; deta = d eta/dy
    and d2eta = d^2 eta/dy^2 of the map (complex-step derivatives of y(eta))."""
    ny = 33
    p = _fd_params(ny)
    p.yi, p.ymax = yi, ymax
    y, eta, deta, d2eta, dy = so.ogengrid(p)
    assert dy == 1.0 / 32.0 and eta[0] == 0.0 and eta[-1] == 1.0
    assert y[0] == 0.0
    assert y[-1] == pytest.approx(ymax, rel=1e-14)
    assert y[16] == pytest.approx(yi, rel=1e-14)
    aa = ymax * yi / (ymax - 2.0 * yi)
    bb = 1.0 + aa / ymax
    yfun = lambda e: aa * e / (bb - e)
    h = 1e-30
    dyde = np.imag(yfun(eta + 1j * h)) / h
    assert np.allclose(deta, 1.0 / dyde, rtol=1e-13)
    # d/dy (deta) = deta d/deta (deta)
    detafun = lambda e: (bb - e) ** 2 / (aa * bb)
    assert np.allclose(d2eta, deta * np.imag(detafun(eta + 1j * h)) / h, rtol=1e-13)


# --- pencil rows ------------------------------------------------------------------

@pytest.fixture(scope="module")
def fd_problem(ts_profile):
    ny = 20
    al, be = 0.23 + 0.011j, 0.17 - 0.004j
    p = _fd_params(ny)
    p.alpha, p.beta = al, be
    pb = so.fd_setup(p, *ts_profile)
    jets = ns.mean_jets(pb.vm, pb.dg, pb.deta, pb.d2eta)
    Ti = ns.linear_tensors(jets, p.Ma, p.Re, p.Pr, p.mattyp, p.Te, "temporal")
    return p, pb, Ti


def test_fd_mean_flow_tensors_match_linearised_equations(fd_problem):
    """fd_temporal.f90:290-578 (grad-derived mean gradients, the tensors of
    temporal.f90) equal the complex-step Frechet derivative of the nonlinear
    equations on the grad-derived jets."""
    from test_oracle_pinning import _tensor_mismatch
    _, pb, Ti = fd_problem
    assert _tensor_mismatch(Ti, pb.T) < 1e-9


def test_fd_interior_rows_match_linearised_operator(fd_problem):
    """Rows of nodes 1..ny-2 (fd_temporal.f90:703-890) equal the operator
    assembled from the independent linearisation on the same D1/D2 rows."""
    p, pb, Ti = fd_problem
    om = 0.061 + 0.0023j
    A0, B0 = so.fd_temporal_pencil(pb)
    Lo = (A0 - om * B0)[5:5 * (p.ny - 1)]
    Lind = ns.interior_operator(Ti, pb.d1, pb.deta, pb.d2eta, p.alpha, p.beta, om, d2=pb.d2)
    scale = np.abs(Lo).max(axis=1, keepdims=True)
    assert np.max(np.abs(Lo - Lind) / scale) < 1e-12


def test_fd_wall_and_top_rows(fd_problem):
    """Wall continuity row: the full continuity equation (it has no second
    derivative); wall energy row: the full energy equation in the columns of
    rho, u, v, w and, in the T column, its second-derivative term only
    (fd_temporal.f90:633-700); wall momentum rows and every top row Dirichlet
    (A0 zero, B0 = i), the wall energy row of B0 i G (:606-629)."""
    p, pb, Ti = fd_problem
    ny = p.ny
    om = 0.061 + 0.0023j
    A0, B0 = so.fd_temporal_pencil(pb)
    Lo = A0 - om * B0
    Lall = ns.interior_operator(Ti, pb.d1, pb.deta, pb.d2eta, p.alpha, p.beta, om, d2=pb.d2, all_rows=True)
    scale = np.abs(Lo[0]).max()
    assert np.max(np.abs(Lo[0] - Lall[0])) / scale < 1e-12
    # T column: without the first-derivative term (its d/dy coefficient
    # times eta' D1, and the eta'' D1 part of the second derivative)
    al, be = p.alpha, p.beta
    bhT = ((Ti["B"][0, 4, 4] - 1j * al * Ti["Vxy"][0, 4, 4] - 1j * be * Ti["Vyz"][0, 4, 4]) * pb.deta[0]
           - Ti["Vyy"][0, 4, 4] * pb.d2eta[0])
    row = Lall[4].copy()
    row[4::5] -= bhT * pb.d1[0]
    scale = np.abs(Lo[4]).max()
    assert np.max(np.abs(Lo[4] - row)) / scale < 1e-12
    for r in (1, 2, 3) + tuple(range(5 * (ny - 1), 5 * ny)):
        assert not A0[r].any()
        e = np.zeros(5 * ny, complex)
        e[r] = 1j
        assert np.array_equal(B0[r], e)
    assert np.allclose(B0[4, :5], 1j * pb.T["G"][0, 4, :], rtol=0, atol=0)


# --- eigenvalues ------------------------------------------------------------------

def _fd_ts_eigenvalue(ts_profile, ny, isothermal):
    pb = so.fd_setup(_fd_params(ny), *ts_profile)
    A0, B0 = so.fd_temporal_pencil(pb)
    if isothermal:
        A0[4, :] = 0.0
        B0[4, :] = 0.0
        B0[4, 4] = 1j
    w, _ = so.temporal_solve_reference(A0, B0)
    return w[np.argmin(np.abs(w - TS_OMEGA))]


def test_fd_isothermal_variant_converges_to_chebyshev_golden(ts_profile):
    """With T = 0 at the wall (the wall condition of thesis/TStest/temporal.inp)
    the finite-difference eigenvalue tends to the Chebyshev golden at fourth
    order (the domain ends at ymax = 50 here, which moves it by < 1e-9)."""
    e64 = abs(_fd_ts_eigenvalue(ts_profile, 64, True) - TS_OMEGA)
    e128 = abs(_fd_ts_eigenvalue(ts_profile, 128, True) - TS_OMEGA)
    assert e128 < 5e-8
    assert e64 / e128 > 12.0


def test_fd_converges_to_chebyshev_pencil_with_its_wall_row(ts_profile):
    """The reference's finite-difference problem converges to the Chebyshev
    problem given the same wall energy row (full energy equation in columns
    0-3, the second-derivative term only in column 4); the limit differs from
    the isothermal golden by about 4e-5."""
    p = so.parse_input(open(golden("thesis_TStest", "temporal.inp")).read())
    pb = so.setup(p, *ts_profile, spatial=False)
    A0, B0 = so.temporal_pencil(pb)
    T = pb.T
    al, be = p.alpha, p.beta
    Dh = T["D"] + 1j * al * T["A"] + 1j * be * T["C"] + al ** 2 * T["Vxx"] + al * be * T["Vxz"] + be ** 2 * T["Vzz"]
    Bh = (T["B"] - 1j * al * T["Vxy"] - 1j * be * T["Vyz"]) * pb.deta[:, None, None] - T["Vyy"] * pb.d2eta[:, None, None]
    Vyy = T["Vyy"] * (pb.deta ** 2)[:, None, None]
    i = p.ny - 1
    row = np.zeros((p.ny, 5), complex)
    for q in range(4):
        row[:, q] = Bh[i][4, q] * pb.d1[i] - Vyy[i][4, q] * pb.d2[i]
    row[:, 4] = -Vyy[i][4, 4] * pb.d2[i]
    row[i, :] += Dh[i][4]
    A0[5 * i + 4] = row.reshape(-1)
    B0[5 * i + 4] = 0.0
    B0[5 * i + 4, 5 * i:5 * i + 5] = 1j * T["G"][i, 4, :]
    w, _ = so.temporal_solve_reference(A0, B0)
    cheb = w[np.argmin(np.abs(w - TS_OMEGA))]
    assert 2e-5 < abs(cheb - TS_OMEGA) < 1e-4
    e64 = abs(_fd_ts_eigenvalue(ts_profile, 64, False) - cheb)
    e128 = abs(_fd_ts_eigenvalue(ts_profile, 128, False) - cheb)
    assert e128 < 1e-7
    assert e64 / e128 > 8.0


@pytest.mark.parametrize("ny", [64, 128, 256])
def test_fd_leading_mode_is_the_ts_wave(ts_profile, ny):
    """With the wall at node 0 the farfield is the outer half of the domain
    (the last nodes): the TS wave has decayed there at every resolution and is
    the least stable discrete mode among the eight eigenvalues nearest the
    default shift 0.35 alpha (at ny = 64 the whole spectrum also holds an
    unstable spurious mode near omega = -0.96)."""
    p = _fd_params(ny)
    pb, _, _, w, v = so.run_fd_temporal(p, *ts_profile)
    nodes = so.fd_farfield_nodes(pb.y, p.ymax)
    assert nodes[-1] == ny - 1 and len(nodes) >= 2
    near = list(np.argsort(np.abs(w - 0.35 * p.alpha))[:8])
    j = so.leading_mode(1, w, v, p.ny, p.alpha.real, candidates=near, nodes=nodes)
    assert abs(w[j] - TS_OMEGA) < 1e-4
    assert so.farfield_ratio(v[:, j], p.ny, nodes) < 1e-3


# --- the device's stencil rows, compiled for the host -----------------------------

@pytest.mark.parametrize("ny", [10, 11, 37, 128, 256])
def test_device_stencil_rows_equal_oracle(tmp_path, ny):
    """stab_amd/csrc/fd_stencils.h (the rows k_fd_ops writes, for both
    solves), compiled with g++ and run on the host, gives the oracle's D1, D2
    and grad matrices bit for bit, with the same scalings 1/dy, 1/dy^2,
    1/(12 dy)."""
    import os
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    src = tmp_path / "fd.cpp"
    src.write_text(r'''
#define __host__
#define __device__
#include "fd_stencils.h"
#include <cstdio>
#include <cstdlib>
int main(int argc, char **argv) {
  const int ny = std::atoi(argv[1]);
  const double dy = 1.0 / (double)(ny - 1), h1 = 1.0 / dy, h2 = 1.0 / (dy * dy), w = 12.0 * dy;
  for (int i = 0; i < ny; ++i)
    for (int j = 0; j < ny; ++j) {
      const double s1 = (i >= ny - 2) ? fd::g_coef(ny, i, j) / w : fd::d1_coef(ny, i, j, true) * h1;
      std::printf("%a %a %a %a %a\n", fd::d1_coef(ny, i, j) * h1, fd::d2_coef(ny, i, j) * h2,
                  fd::g_coef(ny, i, j) / w, s1, fd::d2_coef(ny, i, j, true) * h2);
    }
  return 0;
}
''')
    exe = tmp_path / "fd"
    r = subprocess.run([cxx, "-O1", "-std=c++17", "-I", os.path.join(root, "stab_amd", "csrc"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe), str(ny)], capture_output=True, text=True)
    assert out.returncode == 0
    vals = np.array([[float.fromhex(t) for t in ln.split()] for ln in out.stdout.split("\n") if ln.strip()])
    vals = vals.reshape(ny, ny, 5)
    dy = 1.0 / float(ny - 1)
    d1, d2 = so.fd_operators(ny, dy)
    assert np.array_equal(vals[:, :, 0], d1)
    assert np.array_equal(vals[:, :, 1], d2)
    assert np.array_equal(vals[:, :, 2], so.fd_grad_matrix(ny, dy))
    s1, s2 = so.fd_spatial_operators(ny, dy)
    assert np.array_equal(vals[:, :, 3], s1)
    assert np.array_equal(vals[:, :, 4], s2)


# --- finite-difference spatial (fd_spatial.f90, itype 4) ----------------------------

TS_ALPHA = 2.2804739411185E-001 - 6.5163146952733E-003j       # thesis/TStest/README.md:14 (Chebyshev, N = 96)


def _fds_params(ny, top=0, wallt=0):
    txt = open(golden("thesis_TStest", "fd_spatial.inp")).read().replace("128 1.0 50.0", f"{ny} 1.0 50.0")
    p = so.parse_input(txt)
    p.top, p.wallt = top, wallt
    return p


def test_fd_spatial_rows_exact_on_polynomials():
    """fd_spatial.f90:669-903: central rows on nodes 4 .. ny-3 (degree 4 for
    D1, 5 for D2), the grad rows at ny-2 and ny-1 (degree 4), the Carpenter
    closures (degree 3), the biased and one-sided D2 rows (degree 4)."""
    ny, dy = 24, 1.0 / 23.0
    D1, D2 = so.fd_spatial_operators(ny, dy)
    assert _moments_ok(D1, list(range(4, ny)), dy, 4, 1) < 1e-13
    assert _moments_ok(D1, [0, 1, 2, 3], dy, 3, 1) < 1e-13
    assert _moments_ok(D2, list(range(2, ny - 2)), dy, 5, 2) < 1e-12
    assert _moments_ok(D2, [0, 1, ny - 2], dy, 4, 2) < 1e-12
    assert not D2[ny - 1].any()


@pytest.fixture(scope="module")
def fds_problem(ts_profile):
    ny = 20
    p = _fds_params(ny, top=1, wallt=2)
    p.beta = 0.17 - 0.004j
    p.omega = 0.061 + 0.0023j
    pb = so.fd_setup(p, *ts_profile, spatial=True)
    jets = ns.mean_jets(pb.vm, pb.dg, pb.deta, pb.d2eta)
    Ti = ns.linear_tensors(jets, p.Ma, p.Re, p.Pr, p.mattyp, p.Te, "spatial")
    return p, pb, Ti


def test_fd_spatial_rows_match_linearised_operator(fds_problem):
    """Interior rows of L(alpha) = C0 + alpha C1 + alpha^2 C2 equal the
    independently linearised operator; the wall continuity and top
    continuity (top = 1) rows too; the wall energy row (wallt = 2) too except
    that its T column has no first-derivative term and its time term takes G
    of node ny-1 (fd_spatial.f90:899-907, the loop index left at ny); the
    other wall and top rows are Dirichlet (C0 = -1, C1 = C2 = 0)."""
    p, pb, Ti = fds_problem
    ny = p.ny
    al, be, om = 0.23 + 0.011j, p.beta, p.omega
    C0, C1, C2 = so.fd_spatial_pencil(pb)
    Lo = C0 + al * C1 + al * al * C2
    Lall = ns.interior_operator(Ti, pb.d1, pb.deta, pb.d2eta, al, be, om, d2=pb.d2, all_rows=True)
    rows = list(range(5, 5 * (ny - 1))) + [0, 5 * (ny - 1)]
    scale = np.abs(Lo[rows]).max(axis=1, keepdims=True)
    assert np.max(np.abs(Lo[rows] - Lall[rows]) / scale) < 1e-12
    bhT = ((Ti["B"][0, 4, 4] - 1j * al * Ti["Vxy"][0, 4, 4] - 1j * be * Ti["Vyz"][0, 4, 4]) * pb.deta[0]
           - Ti["Vyy"][0, 4, 4] * pb.d2eta[0])
    row = Lall[4].copy()
    row[4::5] -= bhT * pb.d1[0]
    row[0:5] += 1j * om * (Ti["G"][0, 4, :] - Ti["G"][ny - 1, 4, :])
    assert np.max(np.abs(Lo[4] - row)) / np.abs(Lo[4]).max() < 1e-12
    for r in (1, 2, 3) + tuple(range(5 * (ny - 1) + 1, 5 * ny)):
        e = np.zeros(5 * ny, complex)
        e[r] = -1.0
        assert np.array_equal(C0[r], e) and not C1[r].any() and not C2[r].any()


def _fds_ts_alpha(ts_profile, ny, top, wallt, near):
    pb = so.fd_setup(_fds_params(ny, top, wallt), *ts_profile, spatial=True)
    alp, _ = so.spatial_solve_reference(*so.fd_spatial_pencil(pb), 0)
    return alp[np.argmin(np.abs(alp - near))]


def test_fd_spatial_converges_to_chebyshev_golden(ts_profile):
    """thesis/TStest/fd_spatial.inp (T = 0 at the wall, top = 0): alpha tends
    to the Chebyshev golden of README.md at fourth order."""
    e64 = abs(_fds_ts_alpha(ts_profile, 64, 0, 0, TS_ALPHA) - TS_ALPHA)
    e128 = abs(_fds_ts_alpha(ts_profile, 128, 0, 0, TS_ALPHA) - TS_ALPHA)
    assert e128 < 1e-7
    assert e64 / e128 > 12.0


def test_fd_spatial_energy_wall_row_converges_to_chebyshev_with_that_row(ts_profile):
    """wallt = 2: the finite-difference alpha tends to the Chebyshev spatial
    pencil given the same wall energy row, time term with the top node's G
    included; with the wall's own G in that term the Chebyshev limit is
    missed by more than 3e-6."""
    p = so.parse_input(open(golden("thesis_TStest", "spatial.inp")).read())
    pb = so.setup(p, *ts_profile, spatial=True)
    C0, C1, C2 = so.spatial_pencil(pb)
    T = pb.T
    be, om = p.beta, p.omega
    Dh = T["D"] + 1j * be * T["C"] + be ** 2 * T["Vzz"]
    Bh = (T["B"] - 1j * be * T["Vyz"]) * pb.deta[:, None, None] - T["Vyy"] * pb.d2eta[:, None, None]
    Vyy = T["Vyy"] * (pb.deta ** 2)[:, None, None]
    Dh1 = 1j * T["A"] + be * T["Vxz"]
    Bh1 = -1j * T["Vxy"] * pb.deta[:, None, None]
    i = p.ny - 1
    w = 5 * i
    r0 = np.zeros((p.ny, 5), complex)
    r1 = np.zeros((p.ny, 5), complex)
    for q in range(4):
        r0[:, q] = Bh[i][4, q] * pb.d1[i] - Vyy[i][4, q] * pb.d2[i]
        r1[:, q] = Bh1[i][4, q] * pb.d1[i]
    r0[:, 4] = -Vyy[i][4, 4] * pb.d2[i]
    r0[i, :] += Dh[i][4] - 1j * om * T["G"][0, 4, :]
    r1[i, :] += Dh1[i][4]
    C0[w + 4], C1[w + 4] = r0.reshape(-1), r1.reshape(-1)
    C2[w + 4] = 0.0
    C2[w + 4, w:w + 5] = T["Vxx"][i, 4, :]
    alp, _ = so.spatial_solve_reference(C0, C1, C2, 0)
    cheb = alp[np.argmin(np.abs(alp - TS_ALPHA))]
    e64 = abs(_fds_ts_alpha(ts_profile, 64, 0, 2, cheb) - cheb)
    e128 = abs(_fds_ts_alpha(ts_profile, 128, 0, 2, cheb) - cheb)
    assert e128 < 5e-7
    assert e64 / e128 > 4.0
    fp = _fds_params(128, 0, 2)
    fpb = so.fd_setup(fp, *ts_profile, spatial=True)
    D0, D1_, D2_ = so.fd_spatial_pencil(fpb)
    D0[4, 0:5] += 1j * fp.omega * (fpb.T["G"][fp.ny - 2, 4, :] - fpb.T["G"][0, 4, :])
    alt, _ = so.spatial_solve_reference(D0, D1_, D2_, 0)
    assert abs(alt[np.argmin(np.abs(alt - cheb))] - cheb) > 3e-6


def test_fd_spatial_leading_mode_is_the_ts_wave(ts_profile):
    """The TS wave is the least stable discrete spatial mode among the eight
    alphas nearest omega / 0.35, farfield over the outer half of the domain."""
    p = _fds_params(96)
    pb, _, alp, v = so.run_fd_spatial(p, *ts_profile)
    nodes = so.fd_farfield_nodes(pb.y, p.ymax)
    fin = np.where(np.isfinite(alp) & (np.abs(alp) > 1e-9))[0]
    near = list(fin[np.argsort(np.abs(alp[fin] - p.omega / 0.35))[:8]])
    j = so.leading_mode(2, alp, v[:5 * p.ny], p.ny, omega_r=p.omega.real, candidates=near, nodes=nodes)
    assert abs(alp[j] - TS_ALPHA) < 1e-6
