"""Pins of the compressible Blasius generator (stab_amd/meanflow.py) on the
Sutherland and Pr != 1 branches the benchmark uses, by invariants of the
similarity equations that do not depend on how they are integrated:

  * energy integral: with an adiabatic wall the total-enthalpy defect flux
    vanishes, int f' (g + r f'^2 - 1 - r) d eta = 0, r = (gamma-1) M^2 / 2,
    for any viscosity law and any Pr (it follows from the momentum and energy
    equations and the boundary conditions alone);
  * Crocco-Busemann: at Pr = 1 and an adiabatic wall the total enthalpy is
    constant, g + r f'^2 = 1 + r at every eta, for any viscosity law;
  * momentum integral: C_w f''(0) = int f' (1 - f') d eta;
  * recovery factor at C = 1 (Chapman law), Pr = 0.72: Pohlhausen's quadrature
    r_f = 2 Pr int_0^inf f''^Pr int_0^eta f''^(2-Pr) ds d eta over the
    incompressible Blasius f'' (computed here by a separate shooting), equal
    to (g_w - 1) / r.
"""

import numpy as np
import pytest
from scipy.integrate import solve_ivp, simpson

from stab_amd import meanflow as mf


def _dense(spec, n=20001):
    sol, x = mf.solve(spec)
    eta = np.linspace(0.0, spec.eta_max, n)
    s = sol.sol(eta)
    return eta, s, x


@pytest.mark.parametrize("Ma,Pr", [(0.5, 0.72), (2.0, 0.72), (1.5, 1.0)])
def test_energy_integral_sutherland(Ma, Pr):
    spec = mf.BlasiusSpec(Ma=Ma, Pr=Pr, law="sutherland", Te=288.15, eta_max=20.0)
    eta, s, _ = _dense(spec)
    fp, g = s[1], s[3]
    r = 0.5 * (spec.gamma - 1.0) * Ma * Ma
    integral = simpson(fp * (g + r * fp * fp - 1.0 - r), x=eta)
    scale = simpson(fp * (g - 1.0), x=eta)          # one of its (non-vanishing) parts
    assert abs(scale) > 1e-3
    assert abs(integral) < 1e-9 * max(1.0, abs(scale))


@pytest.mark.parametrize("law", ["sutherland", "constant"])
def test_crocco_busemann_pr1(law):
    spec = mf.BlasiusSpec(Ma=1.5, Pr=1.0, law=law, Te=288.15, eta_max=20.0)
    eta, s, _ = _dense(spec, 2001)
    fp, g = s[1], s[3]
    r = 0.5 * (spec.gamma - 1.0) * spec.Ma ** 2
    assert np.max(np.abs(g + r * fp * fp - 1.0 - r)) < 1e-9


@pytest.mark.parametrize("Ma,Pr", [(0.5, 0.72), (2.5, 0.72)])
def test_momentum_integral_sutherland(Ma, Pr):
    spec = mf.BlasiusSpec(Ma=Ma, Pr=Pr, law="sutherland", Te=288.15, eta_max=20.0)
    eta, s, _ = _dense(spec)
    fp, F = s[1], s[2]            # F = C f''
    assert abs(F[0] - simpson(fp * (1.0 - fp), x=eta)) < 1e-9


def _blasius_fpp(eta):
    """f'' of f''' + f f'' = 0, f(0) = f'(0) = 0, f'(inf) = 1, by secant
    shooting on f''(0)."""
    def end(a):
        s = solve_ivp(lambda t, y: [y[1], y[2], -y[0] * y[2]], (0, eta[-1]), [0, 0, a],
                      rtol=1e-13, atol=1e-15, method="DOP853")
        return s.y[1, -1] - 1.0
    a0, a1 = 0.45, 0.48
    f0, f1 = end(a0), end(a1)
    for _ in range(40):
        a2 = a1 - f1 * (a1 - a0) / (f1 - f0)
        a0, f0, a1, f1 = a1, f1, a2, end(a2)
        if abs(f1) < 1e-14:
            break
    s = solve_ivp(lambda t, y: [y[1], y[2], -y[0] * y[2]], (0, eta[-1]), [0, 0, a1], t_eval=eta,
                  rtol=1e-13, atol=1e-15, method="DOP853")
    return s.y[2], a1


def test_recovery_factor_chapman_pr072():
    Pr, Ma = 0.72, 0.5
    spec = mf.BlasiusSpec(Ma=Ma, Pr=Pr, law="chapman", eta_max=20.0)
    _, _, x = _dense(spec, 11)
    gw = x[1]
    eta = np.linspace(0.0, 20.0, 40001)
    fpp, fpp0 = _blasius_fpp(eta)
    fpp = np.maximum(fpp, 0.0)             # rounding-level negatives far out
    assert abs(fpp0 - 0.4696) < 1e-4                       # thesis/TStest/blasius.inp:5
    inner = np.concatenate([[0.0], np.cumsum(0.5 * (fpp[1:] ** (2 - Pr) + fpp[:-1] ** (2 - Pr)) * np.diff(eta))])
    rf = 2.0 * Pr * simpson(fpp ** Pr * inner, x=eta)
    r = 0.5 * (spec.gamma - 1.0) * Ma * Ma
    assert abs((gw - 1.0) / r - rf) < 1e-6
    assert 0.84 < rf < 0.86                                 # close to sqrt(Pr) = 0.849


def test_sutherland_law_in_edge_units():
    """C = rho mu / (rho_e mu_e) = mu(T)/mu(Te) / g for mu = mu_r (T/T_r)^1.5 (T_r + S)/(T + S)."""
    spec = mf.BlasiusSpec(law="sutherland", Te=250.0)
    g = np.linspace(0.5, 3.0, 11)
    mu = lambda T: (T / 273.0) ** 1.5 * (273.0 + 110.4) / (T + 110.4)
    assert np.allclose(mf._chapman(spec, g), mu(g * 250.0) / mu(250.0) / g, rtol=1e-14)
