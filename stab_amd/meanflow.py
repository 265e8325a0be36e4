"""Compressible Blasius similarity solver writing stab ``profile.<ind>`` files.

The reference reads its mean flow from files produced by external solvers
(``fsc`` / ``compbl``, run from thesis/TStest/run.sh:3-8 with the inputs of
thesis/TStest/blasius.inp); neither is part of the reference.  This module is
an independent solver of the same similarity problem (zero pressure gradient,
parallel flow), written from the textbook equations:

    (C f'')' + f f'' = 0
    (C g'/Pr)' + f g' + (gamma-1) M^2 C f''^2 = 0
    f(0) = f'(0) = 0,  g'(0) = 0 (adiabatic) or g(0) = Tw/Te,  f'(inf) = g(inf) = 1

with g = T/Te, the Chapman-Rubesin factor C = rho mu / (rho_e mu_e) and the
Howarth-Dorodnitsyn transformation dy = g d(eta).  eta uses the factor-2
convention (f''(0) = 0.4696 for C = 1, thesis/TStest/blasius.inp:5).  The
profile is written in displacement-thickness units, columns ``y rho u v w T``
normalised by edge values, as getmean.f90:70-77 reads them.

This runs on the host: it prepares input files, as the external solvers do.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np
from scipy.integrate import solve_ivp


@dataclasses.dataclass
class BlasiusSpec:
    Ma: float = 0.5
    Pr: float = 0.72
    gamma: float = 1.4
    law: str = "sutherland"      # "sutherland" | "chapman" (C = 1) | "constant" (mu = mu_e)
    Te: float = 288.15           # edge temperature in K for the Sutherland law
    S: float = 110.4             # Sutherland constant in K (input.f90:29-31 uses 110.4)
    Tw_ratio: float | None = None    # None: adiabatic wall; else Tw/Te
    Tw_aw: float | None = None       # if set: isothermal wall at Tw = Tw_aw * Taw (adiabatic wall temperature)
    eta_max: float = 20.0
    nrows: int = 2000
    tol: float = 1e-12


def _mu_ratio(spec: BlasiusSpec, g):
    if spec.law == "sutherland":
        s = spec.S / spec.Te
        return g ** 1.5 * (1.0 + s) / (g + s)
    if spec.law == "constant":
        return np.ones_like(g) if isinstance(g, np.ndarray) else 1.0
    return g            # mu ~ T  => C = 1


def _chapman(spec: BlasiusSpec, g):
    if spec.law == "chapman":
        return np.ones_like(g) if isinstance(g, np.ndarray) else 1.0
    return _mu_ratio(spec, g) / g


def _rhs(spec: BlasiusSpec):
    gm1M2 = (spec.gamma - 1.0) * spec.Ma ** 2
    Pr = spec.Pr

    def f(eta, s):
        # s = [f, f', F = C f'', g, Q = C g'/Pr, y = int g, d = int (g - f')]
        ff, fp, F, g, Q = s[:5]
        C = _chapman(spec, g)
        fpp = F / C
        gp = Pr * Q / C
        return [fp, fpp, -ff * fpp, gp, -ff * gp - gm1M2 * C * fpp * fpp, g, g - fp]

    return f


def _integrate(spec: BlasiusSpec, F0, gw_or_Q0, dense=False):
    if spec.Tw_ratio is None:
        g0, Q0 = gw_or_Q0, 0.0
    else:
        g0, Q0 = spec.Tw_ratio, gw_or_Q0
    s0 = [0.0, 0.0, F0, g0, Q0, 0.0, 0.0]
    sol = solve_ivp(_rhs(spec), (0.0, spec.eta_max), s0, method="DOP853",
                    rtol=spec.tol, atol=spec.tol * 1e-2, dense_output=dense)
    if not sol.success:
        raise RuntimeError("similarity integration failed: " + sol.message)
    return sol


def _newton(spec: BlasiusSpec, x):
    for _ in range(60):
        sol = _integrate(spec, x[0], x[1])
        r = np.array([sol.y[1, -1] - 1.0, sol.y[3, -1] - 1.0])
        if np.max(np.abs(r)) < 1e-13:
            return x
        J = np.empty((2, 2))
        for k in range(2):
            h = 1e-7 * max(1.0, abs(x[k]))
            xp = x.copy(); xp[k] += h
            sp = _integrate(spec, xp[0], xp[1])
            J[:, k] = (np.array([sp.y[1, -1] - 1.0, sp.y[3, -1] - 1.0]) - r) / h
        x = x - np.linalg.solve(J, r)
    raise RuntimeError("shooting did not converge")


def solve(spec: BlasiusSpec):
    """Newton shooting on (f''(0), g(0)) [adiabatic] or (f''(0), g'(0)) [isothermal]
    so that f'(eta_max) = 1 and g(eta_max) = 1.  Returns the dense solution and
    the wall values.  With Tw_aw set, the adiabatic problem is solved first and
    the wall temperature is stepped from Taw to Tw_aw * Taw, each isothermal
    Newton solve starting from the previous one."""
    # Crocco-Busemann starting guess for the wall temperature
    taw = 1.0 + spec.Pr ** 0.5 * 0.5 * (spec.gamma - 1.0) * spec.Ma ** 2
    if spec.Tw_aw is not None:
        ad = dataclasses.replace(spec, Tw_aw=None, Tw_ratio=None)
        xa = _newton(ad, np.array([0.4696, taw]))
        g_aw = xa[1]
        x = np.array([xa[0], 0.0])
        steps = max(1, int(math.ceil(abs(spec.Tw_aw - 1.0) / 0.1)))
        for k in range(1, steps + 1):
            gw = g_aw * (1.0 + (spec.Tw_aw - 1.0) * k / steps)
            x = _newton(dataclasses.replace(spec, Tw_aw=None, Tw_ratio=gw), x)
        iso = dataclasses.replace(spec, Tw_aw=None, Tw_ratio=g_aw * spec.Tw_aw)
        return _integrate(iso, x[0], x[1], dense=True), x
    x = np.array([0.4696, taw if spec.Tw_ratio is None else 0.0])
    x = _newton(spec, x)
    return _integrate(spec, x[0], x[1], dense=True), x


def profile(spec: BlasiusSpec):
    """Return (y/delta*, rho, u, v, w, T) on nrows uniform-eta rows; edge values 1."""
    sol, _ = solve(spec)
    eta = np.linspace(0.0, spec.eta_max, spec.nrows)
    s = sol.sol(eta)
    fp, g, y = s[1], s[3], s[5]
    dstar = sol.y[6, -1]        # g - f' has decayed to rounding level at eta_max
    out = np.zeros((spec.nrows, 6))
    out[:, 0] = y / dstar
    out[:, 1] = 1.0 / g
    out[:, 2] = fp
    out[:, 5] = g
    return out


def write_profile(path: str, table: np.ndarray, header: str | None = None):
    """Write ``y rho u v w T`` rows (getmean.f90:70-77 format: list-directed)."""
    with open(path, "w") as f:
        if header:
            for line in header.splitlines():
                f.write("# " + line + "\n")
        for row in table:
            f.write(" ".join("%22.15E" % v for v in row) + "\n")
