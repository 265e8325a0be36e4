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
    """stab_amd/csrc/fd_stencils.h (the rows k_fd_ops writes), compiled with
    g++ and run on the host, gives the oracle's D1, D2 and grad matrices
    bit for bit, with the same scalings 1/dy, 1/dy^2, 1/(12 dy)."""
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
    for (int j = 0; j < ny; ++j)
      std::printf("%a %a %a\n", fd::d1_coef(ny, i, j) * h1, fd::d2_coef(ny, i, j) * h2, fd::g_coef(ny, i, j) / w);
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
    vals = vals.reshape(ny, ny, 3)
    dy = 1.0 / float(ny - 1)
    d1, d2 = so.fd_operators(ny, dy)
    assert np.array_equal(vals[:, :, 0], d1)
    assert np.array_equal(vals[:, :, 1], d2)
    assert np.array_equal(vals[:, :, 2], so.fd_grad_matrix(ny, dy))
