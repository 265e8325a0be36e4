"""Independent linearisation of the compressible Navier-Stokes equations.

TEST INFRASTRUCTURE ONLY: a second, independently written derivation of the
linear stability operator, used to pin the oracle's coefficient tensors
(oracle/stab_oracle.py::temporal_coefficients / spatial_coefficients, which
restate temporal.f90:309-598 and spatial.f90:354-673) and its pencils.

Nothing here is read off the reference's coefficient statements.  The
nonlinear equations are written from the textbook (nondimensional by the edge
state, p = rho T / (gamma M^2), Stokes' hypothesis lambda = -2/3 mu,
Fourier heat flux with k proportional to mu at constant Pr):

  continuity  rho_t + u . grad rho + rho div u
  momentum    rho (u_i,t + u_j u_i,j) + p_,i - (1/Re) d_j tau_ij
              tau_ij = mu (u_i,j + u_j,i) + lambda div u delta_ij
  energy (E)  rho (T_t + u . grad T) + (gamma-1) rho T div u
              - gamma/(Pr Re) div(k grad T) - gamma (gamma-1) M^2 Phi / Re
              Phi = tau_ij u_i,j

The spatial formulation uses E as its energy row; the temporal one uses
(E - (gamma-1) T * continuity) / gamma, the same equation with rho div u
eliminated through continuity (pointwise row operations, so the spectrum of a
collocation pencil does not change).  The viscosity law is Sutherland's in
edge units, mu(tau) = tau^(3/2) (Te + S) / (tau Te + S), tau = T / Te, or
mu = 1 for the constant law.

The residual is a function of the local "jet" of the fields -- values and
first and second derivatives -- and its Frechet derivative at the mean-flow
jet is taken by the complex-step method (W. Squire and G. Trapp, SIAM Review
40 (1998) 110-112): dR/dj = Im R(j + i h e_j) / h with h = 1e-30, exact to
rounding because no difference is formed.  The derivative with respect to the
q_t, q_x, ... slots is the coefficient of that derivative in the linear
operator, so the tensors of the linear equation

  G q_t + A q_x + B q_y + C q_z + D q
     - Vxx q_xx - Vxy q_xy - Vyy q_yy - Vxz q_xz - Vyz q_yz - Vzz q_zz = 0

follow without any hand linearisation.
"""
from __future__ import annotations

import numpy as np

NV = 5                      # rho, u, v, w, T
SLOTS = ("0", "t", "x", "y", "z", "xx", "xy", "xz", "yy", "yz", "zz")
NJ = NV * len(SLOTS)
GAMMA = 1.4
SUTH_S = 110.4              # Sutherland constant in K (the reference's mattyp = 1 law)


def _slot(name):
    return SLOTS.index(name)


def viscosity(tau, law, Te):
    """(mu, dmu/dtau) in edge units; analytic in tau so complex steps pass."""
    if law == 0:
        one = tau * 0.0 + 1.0
        return one, tau * 0.0
    mu = tau ** 1.5 * (Te + SUTH_S) / (tau * Te + SUTH_S)
    dmu = mu * (1.5 / tau - Te / (tau * Te + SUTH_S))
    return mu, dmu


def residual(jet, Ma, Re, Pr, law, Te, form):
    """Residual of the five equations at one point.  `jet` has NJ entries,
    jet[s*5 + v] = derivative `SLOTS[s]` of variable v.  form 'temporal' or
    'spatial' selects the energy row."""
    g = lambda s, v: jet[_slot(s) * NV + v]
    rho, T = g("0", 0), g("0", 4)
    uvw = [g("0", 1), g("0", 2), g("0", 3)]
    d1 = ("x", "y", "z")
    d2 = {("x", "x"): "xx", ("x", "y"): "xy", ("x", "z"): "xz", ("y", "x"): "xy", ("y", "y"): "yy",
          ("y", "z"): "yz", ("z", "x"): "xz", ("z", "y"): "yz", ("z", "z"): "zz"}
    du = [[g(d1[j], 1 + i) for j in range(3)] for i in range(3)]          # du[i][j] = u_i,j
    ddu = [[[g(d2[(d1[j], d1[k])], 1 + i) for k in range(3)] for j in range(3)] for i in range(3)]
    drho = [g(d, 0) for d in d1]
    dT = [g(d, 4) for d in d1]
    ddT = [g(d2[(d, d)], 4) for d in d1]
    div = du[0][0] + du[1][1] + du[2][2]
    ddiv = [ddu[0][0][k] + ddu[1][1][k] + ddu[2][2][k] for k in range(3)]     # d_k div u

    mu, dmu = viscosity(T, law, Te)
    lam, dlam = -2.0 / 3.0 * mu, -2.0 / 3.0 * dmu
    kc, dkc = mu, dmu                                   # k / k_e = mu / mu_e at constant Pr
    gm2 = GAMMA * Ma * Ma

    cont = g("t", 0) + sum(uvw[j] * drho[j] for j in range(3)) + rho * div
    mom = []
    for i in range(3):
        acc = g("t", 1 + i) + sum(uvw[j] * du[i][j] for j in range(3))
        dp = (drho[i] * T + rho * dT[i]) / gm2
        dtau = 0.0
        for j in range(3):
            dtau = dtau + dmu * dT[j] * (du[i][j] + du[j][i]) + mu * (ddu[i][j][j] + ddu[j][i][j])
        dtau = dtau + dlam * dT[i] * div + lam * ddiv[i]
        mom.append(rho * acc + dp - dtau / Re)
    cond = dkc * (dT[0] ** 2 + dT[1] ** 2 + dT[2] ** 2) + kc * (ddT[0] + ddT[1] + ddT[2])
    phi = lam * div * div
    for i in range(3):
        for j in range(3):
            phi = phi + mu * (du[i][j] + du[j][i]) * du[i][j]
    dTdt = g("t", 4) + sum(uvw[j] * dT[j] for j in range(3))
    E = (rho * dTdt + (GAMMA - 1.0) * rho * T * div - GAMMA / (Pr * Re) * cond
         - GAMMA * (GAMMA - 1.0) * Ma * Ma * phi / Re)
    if form == "temporal":
        en = (E - (GAMMA - 1.0) * T * cont) / GAMMA
    elif form == "spatial":
        en = E
    else:
        raise ValueError(form)
    return [cont, mom[0], mom[1], mom[2], en]


def mean_jets(vm, d1, deta, d2eta):
    """Jets of a parallel mean flow (functions of y only; v = 0 as the
    reference's getmean forces).  vm[ny, 5]; derivatives d/dy = eta' d/deta,
    d2/dy2 = eta'^2 d2/deta2 + eta'' d/deta with d/deta the collocation D1."""
    ny = vm.shape[0]
    f = vm.copy()
    f[:, 2] = 0.0
    fy = (d1 @ f) * deta[:, None]
    fyy = (d1 @ (d1 @ f)) * (deta ** 2)[:, None] + (d1 @ f) * d2eta[:, None]
    jets = np.zeros((ny, NJ))
    jets[:, _slot("0") * NV:(_slot("0") + 1) * NV] = f
    jets[:, _slot("y") * NV:(_slot("y") + 1) * NV] = fy
    jets[:, _slot("yy") * NV:(_slot("yy") + 1) * NV] = fyy
    return jets


def linear_tensors(jets, Ma, Re, Pr, law, Te, form, h=1e-30):
    """Per-node 5x5 tensors of the linear operator by complex steps.
    Returns dict name -> [ny, 5, 5] in the reference's sign convention
    (G, A, B, C, D = dR/d(q_t, q_x, q_y, q_z, q); Vab = -dR/d(q_ab))."""
    ny = jets.shape[0]
    M = np.zeros((ny, len(SLOTS), NV, NV))
    for i in range(ny):
        base = jets[i].astype(np.complex128)
        for s in range(len(SLOTS)):
            for v in range(NV):
                j = base.copy()
                j[s * NV + v] += 1j * h
                r = residual(j, Ma, Re, Pr, law, Te, form)
                M[i, s, :, v] = [np.imag(x) / h for x in r]
    out = {"G": M[:, _slot("t")], "A": M[:, _slot("x")], "B": M[:, _slot("y")], "C": M[:, _slot("z")],
           "D": M[:, _slot("0")]}
    for nm, s in (("Vxx", "xx"), ("Vxy", "xy"), ("Vyy", "yy"), ("Vxz", "xz"), ("Vyz", "yz"), ("Vzz", "zz")):
        out[nm] = -M[:, _slot(s)]
    return out


def interior_operator(T, d1, deta, d2eta, alpha, beta, omega, d2=None, all_rows=False):
    """Rows of nodes 1..ny-2 (every node with all_rows) of the operator
    L(alpha, beta, omega) applied to q(y) exp(i(alpha x + beta z - omega t)),
    assembled directly from the linear equation (d_t -> -i omega,
    d_x -> i alpha, d_z -> i beta, d_y -> eta' D1, d_yy -> eta'^2 D2 + eta'' D1
    with D2 = D1 D1 unless given).  Returns [5 (ny-2), 5 ny] ([5 ny, 5 ny])."""
    ny = d1.shape[0]
    Dy = deta[:, None] * d1
    Dyy = (deta ** 2)[:, None] * (d1 @ d1 if d2 is None else d2) + d2eta[:, None] * d1
    ia, ib, iw = 1j * alpha, 1j * beta, -1j * omega
    diag = (T["D"] + iw * T["G"] + ia * T["A"] + ib * T["C"]
            - (ia * ia) * T["Vxx"] - (ia * ib) * T["Vxz"] - (ib * ib) * T["Vzz"])
    dy = T["B"] - ia * T["Vxy"] - ib * T["Vyz"]
    dyy = -T["Vyy"]
    L = np.zeros((5 * ny, 5 * ny), dtype=np.complex128)
    for i in (range(ny) if all_rows else range(1, ny - 1)):
        for p in range(5):
            for q in range(5):
                L[5 * i + p, q::5] += dy[i, p, q] * Dy[i] + dyy[i, p, q] * Dyy[i]
                L[5 * i + p, 5 * i + q] += diag[i, p, q]
    return L if all_rows else L[5:5 * (ny - 1)]
