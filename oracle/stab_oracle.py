"""CPU oracle for stab's Chebyshev-collocation stability path.

TEST INFRASTRUCTURE ONLY.  Only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` may import this module, and only as the
checker (or as the timed CPU baseline).  The product path (``stab_amd``) never
imports it and fails loudly when its HIP library is missing.

This is a NumPy/SciPy restatement, written from scratch, of the reference's
algorithm (sscollis/stab, Fortran).  Every function cites the reference
file:line it follows.  LAPACK calls of the reference (ZGESV, ZGEEV, ZGETRF,
ZGETRS) are replaced by SciPy's LAPACK bindings of the same routines; the QZ
judge is ZGGEV (``scipy.linalg.eig(A, B)``).

Pinning (see tests/test_oracle.py):
  * sgengrid      -- closed form and the y column of thesis/TStest/time.ref
  * spline        -- exactness on quadratics (cantilever end condition),
                     brute-force tridiagonal solve on tiny inputs
  * chebyd        -- exact differentiation of polynomials of degree <= N,
                     row sums, the textbook corner entry (2N^2+1)/6
  * getmat        -- Sutherland law closed form and finite differences
  * pencils       -- end-to-end: thesis/TStest/time.ref and space.ref
                     eigenvalues and eigenfunctions (ndiff -abserr 1e-8),
                     TStest/README.md and test/space.1 eigenvalues, and the
                     residual of the reference eigenfunction in our pencil
  * sort/select/normalise/format -- the fixture text files above.
  * finite-difference temporal (a15, tests/test_fd_oracle.py) -- stencil
                     rows exact on polynomials to their orders, ogengrid
                     closed form, rows against an independent linearisation,
                     fourth-order convergence to the thesis/TStest golden
                     with an isothermal wall row.
"""
from __future__ import annotations

import dataclasses
import io
import math
import os
import struct
from typing import Optional

import numpy as np
import scipy.linalg as sla

# Constants of module `stuff` (stuff.f90:29-47).
PT5 = 0.5
PT66 = 6.6666666666666666666e-1
GAMMA = 1.4
GAMMA1 = 0.4
CP = 1003.1
NDOF = 5
BIG = 1.0e98


# ----------------------------------------------------------------------------
# a1: input semantics (input.f90:2-165, stab.f90:42-55)
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Params:
    mattyp: int = 0
    T0: float = 0.0
    Ma: float = 0.0
    Re: float = 0.0
    Pr: float = 0.0
    ny: int = 1
    yi: float = 0.0
    ymax: float = 0.0
    ievec: int = 1
    ider: bool = True
    top: int = 0
    wall: int = 0
    wallt: int = 0
    curve: int = 0
    itype: int = 1
    alpha: complex = 0j
    beta: complex = 0j
    omega: complex = 0j
    ind: int = 0
    x: float = 0.0
    # derived edge state (input.f90:15-44)
    Te: float = 0.0
    datmat: tuple = (1.0, 0.0, 0.0)
    rmue: float = 0.0
    rlme: float = 0.0
    cone: float = 0.0

    @property
    def navier(self) -> bool:
        # temporal.f90:88-91, spatial.f90:89-92
        return not (self.Re >= BIG or self.Re == 0.0)


def _tokens(line: str):
    """List-directed read: fields split by blanks/commas, '!' comments dropped."""
    line = line.split("!")[0]
    return [t for t in line.replace(",", " ").split() if t]


def parse_input(text: str) -> Params:
    """Parse an ``*.inp`` stream exactly in the order input.f90 reads it.

    input.f90:14-44 (mattyp, T0 only if mattyp=1, Ma Re Pr), :46-59 (ny yi ymax,
    ievec, ider, top wall wallt curve), :72-128 (itype and alpha/beta or
    omega/beta), stab.f90:46-55 (ind, and x for spatial).
    Quirk kept: Te is computed before Ma is read, so Te = T0 (input.f90:24-25).
    """
    lines = [ln for ln in text.splitlines()]
    it = iter(lines)

    def nxt():
        while True:
            toks = _tokens(next(it))
            if toks:
                return toks

    p = Params()
    p.mattyp = int(nxt()[0])
    if p.mattyp == 1:
        p.T0 = float(nxt()[0])
        p.Te = p.T0 / (1.0 + PT5 * GAMMA1 * 0.0 ** 2)   # Ma still 0 here
        p.datmat = (1.715336725523065e-05, 273.0, 110.4)
    else:
        p.datmat = (1.0, 0.0, 0.0)
    t = nxt()
    p.Ma, p.Re, p.Pr = float(t[0]), float(t[1]), float(t[2])
    rmu, rlm, con, *_ = sgetmat(p, p.Te)
    p.rmue, p.rlme, p.cone = rmu, rlm, con
    t = nxt()
    p.ny, p.yi, p.ymax = int(t[0]), float(t[1]), float(t[2])
    p.ievec = int(nxt()[0])
    p.ider = int(nxt()[0]) != 0
    t = nxt()
    p.top, p.wall, p.wallt, p.curve = (int(v) for v in t[:4])
    p.itype = int(nxt()[0])
    if p.itype in (1, 3):
        t = nxt(); p.alpha = complex(float(t[0]), float(t[1]))
        t = nxt(); p.beta = complex(float(t[0]), float(t[1]))
    elif p.itype in (2, 4):
        t = nxt(); p.omega = complex(float(t[0]), float(t[1]))
        t = nxt(); p.beta = complex(float(t[0]), float(t[1]))
    else:
        raise ValueError(f"itype {p.itype} is outside the Chebyshev path")
    p.ind = int(nxt()[0])
    if p.itype in (2, 4):
        p.x = float(nxt()[0])
    return p


# ----------------------------------------------------------------------------
# a5: material law (getmat.f90:2-75)
# ----------------------------------------------------------------------------
def getmat(p: Params, t):
    """Array material law getmat.f90:2-37: returns
    (rmu, rlm, con, dmu, d2mu, dlm, d2lm, dcon, d2con)."""
    t = np.asarray(t, dtype=np.float64)
    d1, d2, d3 = p.datmat
    if p.mattyp == 0:
        rmu = np.full_like(t, d1)
        dmu = np.zeros_like(t)
        d2mu = np.zeros_like(t)
    else:
        rmu = d1 * (t) / d2 * np.sqrt((t) / d2) * (d2 + d3) / (t + d3)
        dmu = (d1 * (3.0 * d3 + t) * (d3 + d2) * np.sqrt(t / d2)) / (2.0 * (d3 + t) ** 2 * d2)
        d2mu = (d1 * (3.0 * d3 ** 2 - 6.0 * d3 * t - t ** 2) * (d3 + d2)) / (
            4.0 * (d3 + t) ** 3 * np.sqrt(t / d2) * d2 ** 2)
    con = rmu * CP / p.Pr
    dcon = dmu * CP / p.Pr
    d2con = d2mu * CP / p.Pr
    rlm = -PT66 * rmu
    dlm = -PT66 * dmu
    d2lm = -PT66 * d2mu
    return rmu, rlm, con, dmu, d2mu, dlm, d2lm, dcon, d2con


def sgetmat(p: Params, t: float):
    """Scalar variant getmat.f90:40-75 (as called from input.f90:43-44)."""
    return tuple(float(v) for v in getmat(p, np.array([t])) for v in [v[0]])


# ----------------------------------------------------------------------------
# a2: grid and mapping (sgengrid.f90:2-45)
# ----------------------------------------------------------------------------
def sgengrid(p: Params):
    """Gauss-Lobatto eta and the algebraic mappings of sgengrid.f90:15-45.
    Returns (y, eta, deta, d2eta); node 0 is eta=1 (top), node N-1 the wall."""
    ny = p.ny
    dth = math.pi / float(ny - 1)
    eta = np.cos(np.arange(ny, dtype=np.float64) * dth)
    if p.yi == 0.0:
        raise ValueError("tanh mapping (Yi=0) is interactive and outside this path")
    if p.ymax == 0.0:
        lmap = p.yi
        deta = (eta - 1.0) ** 2 / (2.0 * lmap)
        d2eta = (eta - 1.0) ** 3 / (2.0 * lmap ** 2)
        y = np.empty(ny)
        for i in range(ny):
            y[i] = lmap * (1.0 + eta[i]) / (1.0 - eta[i]) if eta[i] != 1.0 else 1.0e99
    else:
        yi, ymax = p.yi, p.ymax
        y = ymax * yi * (1.0 + eta) / (1.0 + 2.0 * yi - eta)
        deta = (2.0 * yi + 1.0 - eta) ** 2 / (2.0 * ymax * yi * (yi + 1.0))
        d2eta = -PT5 * (2.0 * yi + 1.0 - eta) ** 3 / (ymax * yi * (yi + 1.0)) ** 2
    return y, eta, deta, d2eta


# ----------------------------------------------------------------------------
# a3: mean-flow profile and cubic spline (getmean.f90:2-122, spline.f90:2-75)
# ----------------------------------------------------------------------------
def read_profile(path: str):
    """Rows of `y rho u v w T`, lines starting with '#' skipped, v forced to
    zero (getmean.f90:40-80).  Returns (ym[nym], vt[nym,5])."""
    rows = []
    with open(path) as f:
        for line in f:
            if line[:1] == "#":
                continue
            toks = _tokens(line)
            if not toks:
                continue
            rows.append([float(v.replace("D", "E").replace("d", "e")) for v in toks[:6]])
    arr = np.array(rows, dtype=np.float64)
    ym = arr[:, 0].copy()
    vt = arr[:, 1:6].copy()
    vt[:, 2] = 0.0
    return ym, vt


def spline(x, y):
    """Second derivatives of the cubic spline with the cantilever end condition
    (lambda=1): spline.f90:2-49 (tridiagonal elimination, then back substitution)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    n = x.size
    a = np.zeros(n); b = np.zeros(n); c = np.zeros(n); r = np.zeros(n)
    fdp = np.zeros(n)
    lam = 1.0
    c[0] = x[1] - x[0]
    for i in range(1, n - 1):
        c[i] = x[i + 1] - x[i]
        a[i] = c[i - 1]
        b[i] = 2.0 * (a[i] + c[i])
        r[i] = 6.0 * ((y[i + 1] - y[i]) / c[i] - (y[i] - y[i - 1]) / c[i - 1])
    b[1] = b[1] + lam * c[0]
    b[n - 2] = b[n - 2] + lam * c[n - 2]
    for i in range(2, n - 1):
        t = a[i] / b[i - 1]
        b[i] = b[i] - t * c[i - 1]
        r[i] = r[i] - t * r[i - 1]
    fdp[n - 2] = r[n - 2] / b[n - 2]
    for k in range(2, n - 1):          # i = n-3 ... 1
        i = n - 1 - k
        fdp[i] = (r[i] - c[i] * fdp[i + 1]) / b[i]
    fdp[0] = lam * fdp[1]
    fdp[n - 1] = lam * fdp[n - 2]
    return fdp


def speval(x, y, fdp, xx):
    """Evaluate the spline at xx: first interval with xx <= x[i+1]
    (spline.f90:52-75)."""
    n = len(x)
    i = 0
    while i < n - 1 and not (xx <= x[i + 1]):
        i += 1
    dxm = xx - x[i]
    dxp = x[i + 1] - xx
    de = x[i + 1] - x[i]
    return (fdp[i] * dxp * (dxp * dxp / de - de) / 6.0
            + fdp[i + 1] * dxm * (dxm * dxm / de - de) / 6.0
            + y[i] * dxp / de + y[i + 1] * dxm / de)


def getmean(y, ym, vt):
    """Spline the profile onto the collocation grid; above the last profile
    row use that row's values (getmean.f90:82-112).  Returns vm[ny,5]."""
    ny = len(y)
    vm = np.zeros((ny, NDOF))
    fdps = [spline(ym, vt[:, k]) for k in range(NDOF)]
    ymaxm = ym[-1]
    for j in range(ny):
        if y[j] <= ymaxm:
            for k in range(NDOF):
                vm[j, k] = speval(ym, vt[:, k], fdps[k], y[j])
        else:
            vm[j, :] = vt[-1, :]
    return vm


# ----------------------------------------------------------------------------
# a4: Chebyshev derivative matrix (chebyd.f90:2-60)
# ----------------------------------------------------------------------------
def chebyd(n: int):
    """Brute-force Lagrange derivative matrix at x_j = cos(pi j/n), j=0..n
    (chebyd.f90:25-57): a_j = prod_{k!=j}(x_j-x_k), D_jj = sum 1/(x_j-x_k),
    D_jk = a_j/(a_k (x_j-x_k))."""
    pi = math.acos(-1.0)
    x = np.array([math.cos(pi * float(j) / float(n)) for j in range(n + 1)])
    a = np.ones(n + 1)
    for j in range(n + 1):
        for k in range(n + 1):
            if k != j:
                a[j] = a[j] * (x[j] - x[k])
    d = np.zeros((n + 1, n + 1))
    for j in range(n + 1):
        s = 0.0
        for k in range(n + 1):
            if k != j:
                s = s + 1.0 / (x[j] - x[k])
        d[j, j] = s
    for j in range(n + 1):
        for k in range(n + 1):
            if k != j:
                d[j, k] = a[j] / (a[k] * (x[j] - x[k]))
    return d


def derivative_operators(p: Params):
    """D1, D2 = D1 D1, and the temperature operators Dt1/Dt2
    (temporal.f90:135-144, spatial.f90:156-165)."""
    d1 = chebyd(p.ny - 1)
    d2 = d1 @ d1
    if p.wallt == 2:
        dt1 = d1.copy()
        dt1[p.ny - 1, :] = 0.0
        dt2 = d1 @ dt1
    else:
        dt1, dt2 = d1, d2
    return d1, d2, dt1, dt2


# ----------------------------------------------------------------------------
# a6 / a10: per-node 5x5 coefficient tensors
# ----------------------------------------------------------------------------
def _mean_state(p: Params, vm, d1, d2, deta, d2eta, spatial: bool, raw=None):
    """Mean-field gradients in physical space and material properties
    (temporal.f90:126-307; spatial.f90:147-352 with curve=0).  `raw`, when
    given, holds the eta-space first and second derivatives of vm (the
    finite-difference path's grad calls, fd_temporal.f90:131-145) in place of
    d1 @ vm and d2 @ vm."""
    ny = p.ny
    z = np.zeros(ny)
    rhom = vm[:, 0].copy(); u1m = vm[:, 1].copy()
    u2m = z.copy() if spatial else vm[:, 2].copy()
    u3m = vm[:, 3].copy(); tm = vm[:, 4].copy()
    if raw is None:
        g2vm = d1 @ vm
        g22vm = d2 @ vm
    else:
        g2vm, g22vm = raw[0].copy(), raw[1].copy()
    g22vm = g22vm * (deta ** 2)[:, None] + g2vm * d2eta[:, None]
    g2vm = g2vm * deta[:, None]
    zero5 = np.zeros((ny, NDOF))
    g1vm = zero5; g3vm = zero5
    g11vm = zero5; g12vm = zero5; g13vm = zero5
    g21vm = zero5; g23vm = zero5
    g31vm = zero5; g32vm = zero5; g33vm = zero5
    gum = np.zeros((ny, 3, 3))
    gum[:, 0, 0] = g1vm[:, 1]; gum[:, 0, 1] = g2vm[:, 1]; gum[:, 0, 2] = g3vm[:, 1]
    gum[:, 1, 0] = g1vm[:, 2]; gum[:, 1, 1] = g2vm[:, 2]; gum[:, 1, 2] = g3vm[:, 2]
    gum[:, 2, 0] = g1vm[:, 3]; gum[:, 2, 1] = g2vm[:, 3]; gum[:, 2, 2] = g3vm[:, 3]
    grhom = np.stack([g1vm[:, 0], g2vm[:, 0], g3vm[:, 0]], axis=1)
    gtm = np.stack([g1vm[:, 4], g2vm[:, 4], g3vm[:, 4]], axis=1)
    fact = 1.0 / (GAMMA * p.Ma ** 2)
    gpm = fact * (grhom * tm[:, None] + rhom[:, None] * gtm)
    st = dict(rhom=rhom, u1m=u1m, u2m=u2m, u3m=u3m, tm=tm, gum=gum, grhom=grhom,
              gtm=gtm, gpm=gpm, g2vm=g2vm, g22vm=g22vm, g11vm=g11vm, g12vm=g12vm,
              g13vm=g13vm, g21vm=g21vm, g23vm=g23vm, g31vm=g31vm, g32vm=g32vm, g33vm=g33vm)
    if spatial:
        # curve = 0: h = 1, all metric derivatives zero (spatial.f90:119-125)
        h = np.ones(ny); dhds = z; dhdr = z; dhdsr = z; dhdrr = z
        st.update(h=h)
        st["divum"] = (gum[:, 0, 0] + u2m * dhdr) / h + gum[:, 1, 1] + gum[:, 2, 2]
        st["g1divum"] = (-dhds / h ** 3 * (gum[:, 0, 0] + u2m * dhdr)
                         + 1.0 / h ** 2 * (g11vm[:, 1] + gum[:, 1, 0] * dhdr + u2m * dhdsr)
                         + 1.0 / h * (g12vm[:, 2] + g13vm[:, 3]))
        st["g2divum"] = (-dhdr / h ** 2 * (gum[:, 0, 0] + u2m * dhdr)
                         + 1.0 / h * (g12vm[:, 1] + gum[:, 1, 1] * dhdr + u2m * dhdrr)
                         + (g22vm[:, 2] + g23vm[:, 3]))
        st["g3divum"] = 1.0 / h * (g13vm[:, 1] + gum[:, 1, 2] * dhdr) + g23vm[:, 2] + g33vm[:, 3]
        S = np.zeros((ny, 3, 3))
        S[:, 0, 0] = (gum[:, 0, 0] + u2m * dhdr) / h
        S[:, 0, 1] = PT5 * ((gum[:, 1, 0] - u1m * dhdr) / h + gum[:, 0, 1])
        S[:, 0, 2] = PT5 * (gum[:, 2, 0] / h + gum[:, 0, 2])
        S[:, 1, 0] = S[:, 0, 1]
        S[:, 1, 1] = gum[:, 1, 1]
        S[:, 1, 2] = PT5 * (gum[:, 2, 1] + gum[:, 1, 2])
        S[:, 2, 0] = S[:, 0, 2]
        S[:, 2, 1] = S[:, 1, 2]
        S[:, 2, 2] = gum[:, 2, 2]
        st["S"] = S
        st["S1jj"] = (-PT5 * (dhdr ** 2 + dhdrr * h) / h ** 2 * u1m + PT5 * dhdr * gum[:, 0, 1] / h
                      + PT5 * g22vm[:, 1] - dhds * gum[:, 0, 0] / h ** 3 + g11vm[:, 1] / h ** 2
                      + PT5 * g33vm[:, 1] + (h * dhdsr - dhdr * dhds) / h ** 3 * u2m
                      + 3.0 * dhdr * gum[:, 1, 0] / (2.0 * h ** 2)
                      + PT5 * g12vm[:, 2] / h + PT5 * g13vm[:, 3] / h)
        st["S2jj"] = (PT5 * (dhdr * dhds - h * dhdsr) / h ** 3 * u1m
                      - 3.0 * dhdr * gum[:, 0, 0] / (2.0 * h ** 2)
                      + PT5 * g12vm[:, 1] / h - dhdr ** 2 * u2m / h ** 2
                      + dhdr * gum[:, 1, 1] / h + g22vm[:, 2]
                      - PT5 * dhds * gum[:, 1, 0] / h ** 3 + PT5 * g11vm[:, 2] / h ** 2
                      + PT5 * g33vm[:, 2] + PT5 * g23vm[:, 3])
        st["S3jj"] = (PT5 * g13vm[:, 1] / h + PT5 * g23vm[:, 2]
                      + PT5 * dhdr * gum[:, 1, 2] / h + PT5 * dhdr * gum[:, 2, 1] / h
                      + PT5 * g22vm[:, 3] + PT5 * g11vm[:, 3] / h ** 2 + g33vm[:, 3]
                      - PT5 * dhds * gum[:, 2, 0] / h ** 3)
        st["Lapt"] = 1.0 / h * (-dhds / h ** 2 * gtm[:, 0] + 1.0 / h * g11vm[:, 4]
                                + h * g22vm[:, 4] + gtm[:, 1] * dhdr + h * g33vm[:, 4])
        st.update(dhds=dhds, dhdr=dhdr, dhdsr=dhdsr, dhdrr=dhdrr)
    else:
        st["divum"] = gum[:, 0, 0] + gum[:, 1, 1] + gum[:, 2, 2]
        st["g1divum"] = g11vm[:, 1] + g12vm[:, 2]
        st["g2divum"] = g12vm[:, 1] + g22vm[:, 2]
        st["g3divum"] = z.copy()
        S = np.zeros((ny, 3, 3))
        for i in range(3):
            for j in range(3):
                S[:, i, j] = PT5 * (gum[:, i, j] + gum[:, j, i])
        st["S"] = S
        st["S1jj"] = PT5 * (g11vm[:, 1] + g11vm[:, 1] + g22vm[:, 1] + g12vm[:, 2] + g33vm[:, 1] + g13vm[:, 3])
        st["S2jj"] = PT5 * (g11vm[:, 2] + g21vm[:, 1] + g22vm[:, 2] + g22vm[:, 2] + g33vm[:, 2] + g23vm[:, 3])
        st["S3jj"] = PT5 * (g11vm[:, 3] + g31vm[:, 1] + g22vm[:, 3] + g32vm[:, 2] + g33vm[:, 3] + g33vm[:, 3])
    rmu, rlm, con, dmu, d2mu, dlm, d2lm, dcon, d2con = getmat(p, tm * p.Te)
    Te = p.Te
    rmu = rmu / p.rmue; dmu = dmu * Te / p.rmue; d2mu = d2mu * Te ** 2 / p.rmue
    con = con / p.cone; dcon = dcon * Te / p.cone; d2con = d2con * Te ** 2 / p.cone
    rlm = rlm / p.rlme; dlm = dlm * Te / p.rlme; d2lm = d2lm * Te ** 2 / p.rlme
    st.update(rmu=rmu, dmu=dmu, d2mu=d2mu, con=con, dcon=dcon, d2con=d2con,
              rlm=rlm, dlm=dlm, d2lm=d2lm)
    for nm, base, dbase in (("mu", dmu, d2mu), ("con", dcon, d2con), ("lm", dlm, d2lm)):
        st["g%s" % nm] = base[:, None] * gtm          # g1mu,g2mu,g3mu -> [:,0..2]
        st["gd%s" % nm] = dbase[:, None] * gtm        # g1dmu ...
    return st


def temporal_coefficients(p: Params, st):
    """Per-node 5x5 tensors G,A,B,C,D,Vxx,Vxy,Vyy,Vxz,Vyz,Vzz of the temporal
    formulation (temporal.f90:309-598).  Arrays are [node, row, col], 0-based."""
    ny = p.ny
    T = {k: np.zeros((ny, 5, 5)) for k in ("G", "A", "B", "C", "D", "Vxx", "Vxy", "Vyy", "Vxz", "Vyz", "Vzz")}
    G, A, B, C, D = T["G"], T["A"], T["B"], T["C"], T["D"]
    Vxx, Vxy, Vyy, Vxz, Vyz, Vzz = T["Vxx"], T["Vxy"], T["Vyy"], T["Vxz"], T["Vyz"], T["Vzz"]
    rhom, u1m, u2m, u3m, tm = st["rhom"], st["u1m"], st["u2m"], st["u3m"], st["tm"]
    gum, grhom, gtm, gpm, S = st["gum"], st["grhom"], st["gtm"], st["gpm"], st["S"]
    divum = st["divum"]
    g1divum, g2divum, g3divum = st["g1divum"], st["g2divum"], st["g3divum"]
    rmu, dmu, rlm, dlm, con, dcon = st["rmu"], st["dmu"], st["rlm"], st["dlm"], st["con"], st["dcon"]
    gmu, gdmu, glm, gdlm, gcon, gdcon = st["gmu"], st["gdmu"], st["glm"], st["gdlm"], st["gcon"], st["gdcon"]
    g11vm, g22vm, g33vm = st["g11vm"], st["g22vm"], st["g33vm"]
    gM2 = GAMMA * p.Ma ** 2
    Re, Pr, Ma = p.Re, p.Pr, p.Ma
    nav = p.navier

    # continuity (temporal.f90:313-327)
    G[:, 0, 0] = 1.0
    A[:, 0, 0] = u1m; A[:, 0, 1] = rhom
    B[:, 0, 0] = u2m; B[:, 0, 2] = rhom
    C[:, 0, 0] = u3m; C[:, 0, 3] = rhom
    D[:, 0, 0] = divum; D[:, 0, 1] = grhom[:, 0]; D[:, 0, 2] = grhom[:, 1]; D[:, 0, 3] = grhom[:, 2]

    # three momentum components (temporal.f90:331-526)
    for c in range(3):
        r = 1 + c
        G[:, r, r] = rhom
        conv = [A, B, C]
        for d in range(3):
            conv[d][:, r, r] = rhom * (u1m, u2m, u3m)[d]
        conv[c][:, r, 0] = tm / gM2
        conv[c][:, r, 4] = rhom / gM2
        D[:, r, 0] = u1m * gum[:, c, 0] + u2m * gum[:, c, 1] + u3m * gum[:, c, 2] + gtm[:, c] / gM2
        D[:, r, 1] = rhom * gum[:, c, 0]
        D[:, r, 2] = rhom * gum[:, c, 1]
        D[:, r, 3] = rhom * gum[:, c, 2]
        D[:, r, 4] = grhom[:, c] / gM2
    if nav:
        gdiv = (g1divum, g2divum, g3divum)
        S_jj = (st["S1jj"], st["S2jj"], st["S3jj"])
        # second-viscosity (lambda) terms
        f = p.rlme / (p.rmue * Re)
        # x-momentum (temporal.f90:352-367)
        A[:, 1, 1] -= f * glm[:, 0]
        A[:, 1, 4] -= f * dlm * divum
        B[:, 1, 2] -= f * glm[:, 0]
        C[:, 1, 3] -= f * glm[:, 0]
        D[:, 1, 4] -= f * (gdlm[:, 0] * divum + dlm * g1divum)
        Vxx[:, 1, 1] = f * rlm; Vxy[:, 1, 2] = f * rlm; Vxz[:, 1, 3] = f * rlm
        # shear-viscosity (mu) terms (temporal.f90:371-392)
        f = 1.0 / Re
        A[:, 1, 1] -= f * 2.0 * gmu[:, 0]
        A[:, 1, 2] -= f * gmu[:, 1]
        A[:, 1, 3] -= f * gmu[:, 2]
        A[:, 1, 4] -= f * dmu * 2.0 * S[:, 0, 0]
        B[:, 1, 1] -= f * gmu[:, 1]
        B[:, 1, 4] -= f * dmu * 2.0 * S[:, 0, 1]
        C[:, 1, 1] -= f * gmu[:, 2]
        C[:, 1, 4] -= f * dmu * 2.0 * S[:, 0, 2]
        D[:, 1, 4] -= f * 2.0 * (gdmu[:, 0] * S[:, 0, 0] + gdmu[:, 1] * S[:, 0, 1]
                                 + gdmu[:, 2] * S[:, 0, 2] + dmu * S_jj[0])
        Vxx[:, 1, 1] += f * 2.0 * rmu; Vxy[:, 1, 2] += f * rmu; Vyy[:, 1, 1] += f * rmu
        Vxz[:, 1, 3] += f * rmu; Vzz[:, 1, 1] += f * rmu
        # y-momentum (temporal.f90:417-458)
        f = p.rlme / (p.rmue * Re)
        A[:, 2, 1] -= f * glm[:, 1]
        B[:, 2, 2] -= f * glm[:, 1]
        B[:, 2, 4] -= f * dlm * divum
        C[:, 2, 3] -= f * glm[:, 1]
        D[:, 2, 4] -= f * (gdlm[:, 1] * divum + dlm * g2divum)
        Vxy[:, 2, 1] = f * rlm; Vyy[:, 2, 2] = f * rlm; Vyz[:, 2, 3] = f * rlm
        f = 1.0 / Re
        A[:, 2, 2] -= f * gmu[:, 0]
        A[:, 2, 4] -= f * dmu * 2.0 * S[:, 1, 0]
        B[:, 2, 1] -= f * gmu[:, 0]
        B[:, 2, 2] -= f * 2.0 * gmu[:, 1]
        B[:, 2, 3] -= f * gmu[:, 2]
        B[:, 2, 4] -= f * dmu * 2.0 * S[:, 1, 1]
        C[:, 2, 2] -= f * gmu[:, 2]
        C[:, 2, 4] -= f * dmu * 2.0 * S[:, 1, 2]
        D[:, 2, 4] -= f * 2.0 * (gdmu[:, 0] * S[:, 1, 0] + gdmu[:, 1] * S[:, 1, 1]
                                 + gdmu[:, 2] * S[:, 1, 2] + dmu * S_jj[1])
        Vxx[:, 2, 2] += f * rmu; Vxy[:, 2, 1] += f * rmu; Vyy[:, 2, 2] += f * 2.0 * rmu
        Vyz[:, 2, 3] += f * rmu; Vzz[:, 2, 2] += f * rmu
        # z-momentum (temporal.f90:483-524)
        f = p.rlme / (p.rmue * Re)
        A[:, 3, 1] -= f * glm[:, 2]
        B[:, 3, 2] -= f * glm[:, 2]
        C[:, 3, 3] -= f * glm[:, 2]
        C[:, 3, 4] -= f * dlm * divum
        D[:, 3, 4] -= f * (gdlm[:, 2] * divum + dlm * g3divum)
        Vxz[:, 3, 1] = f * rlm; Vyz[:, 3, 2] = f * rlm; Vzz[:, 3, 3] = f * rlm
        f = 1.0 / Re
        A[:, 3, 3] -= f * gmu[:, 0]
        A[:, 3, 4] -= f * dmu * 2.0 * S[:, 2, 0]
        B[:, 3, 3] -= f * gmu[:, 1]
        B[:, 3, 4] -= f * dmu * 2.0 * S[:, 2, 1]
        C[:, 3, 1] -= f * gmu[:, 0]
        C[:, 3, 2] -= f * gmu[:, 1]
        C[:, 3, 3] -= f * 2.0 * gmu[:, 2]
        C[:, 3, 4] -= f * dmu * 2.0 * S[:, 2, 2]
        D[:, 3, 4] -= f * 2.0 * (gdmu[:, 0] * S[:, 2, 0] + gdmu[:, 1] * S[:, 2, 1]
                                 + gdmu[:, 2] * S[:, 2, 2] + dmu * S_jj[2])
        Vxx[:, 3, 3] += f * rmu; Vyy[:, 3, 3] += f * rmu; Vxz[:, 3, 1] += f * rmu
        Vyz[:, 3, 2] += f * rmu; Vzz[:, 3, 3] += f * 2.0 * rmu

    # energy (temporal.f90:530-598)
    G[:, 4, 0] = -GAMMA1 * tm / GAMMA
    G[:, 4, 4] = rhom / GAMMA
    A[:, 4, 0] = -GAMMA1 * u1m * tm / GAMMA; A[:, 4, 4] = rhom * u1m / GAMMA
    B[:, 4, 0] = -GAMMA1 * u2m * tm / GAMMA; B[:, 4, 4] = rhom * u2m / GAMMA
    C[:, 4, 0] = -GAMMA1 * u3m * tm / GAMMA; C[:, 4, 4] = rhom * u3m / GAMMA
    D[:, 4, 0] = 1.0 / GAMMA * (u1m * gtm[:, 0] + u2m * gtm[:, 1] + u3m * gtm[:, 2])
    D[:, 4, 1] = rhom * gtm[:, 0] - GAMMA1 * Ma ** 2 * gpm[:, 0]
    D[:, 4, 2] = rhom * gtm[:, 1] - GAMMA1 * Ma ** 2 * gpm[:, 1]
    D[:, 4, 3] = rhom * gtm[:, 2] - GAMMA1 * Ma ** 2 * gpm[:, 2]
    D[:, 4, 4] = -GAMMA1 / GAMMA * (u1m * grhom[:, 0] + u2m * grhom[:, 1] + u3m * grhom[:, 2])
    if nav:
        f = 1.0 / (Pr * Re)
        A[:, 4, 4] -= f * (gcon[:, 0] + dcon * gtm[:, 0])
        B[:, 4, 4] -= f * (gcon[:, 1] + dcon * gtm[:, 1])
        C[:, 4, 4] -= f * (gcon[:, 2] + dcon * gtm[:, 2])
        D[:, 4, 4] -= f * (gdcon[:, 0] * gtm[:, 0] + gdcon[:, 1] * gtm[:, 1] + gdcon[:, 2] * gtm[:, 2]
                           + dcon * (g11vm[:, 4] + g22vm[:, 4] + g33vm[:, 4]))
        Vxx[:, 4, 4] = f * con; Vyy[:, 4, 4] = f * con; Vzz[:, 4, 4] = f * con
        f = GAMMA1 * Ma ** 2 * p.rlme / (Re * p.rmue)
        A[:, 4, 1] -= f * 2.0 * rlm * divum
        B[:, 4, 2] -= f * 2.0 * rlm * divum
        C[:, 4, 3] -= f * 2.0 * rlm * divum
        D[:, 4, 4] -= f * dlm * divum * divum
        f = 2.0 * GAMMA1 * Ma ** 2 / Re
        A[:, 4, 1] -= f * 2.0 * rmu * S[:, 0, 0]
        A[:, 4, 2] -= f * 2.0 * rmu * S[:, 1, 0]
        A[:, 4, 3] -= f * 2.0 * rmu * S[:, 2, 0]
        B[:, 4, 1] -= f * 2.0 * rmu * S[:, 0, 1]
        B[:, 4, 2] -= f * 2.0 * rmu * S[:, 1, 1]
        B[:, 4, 3] -= f * 2.0 * rmu * S[:, 2, 1]
        C[:, 4, 1] -= f * 2.0 * rmu * S[:, 0, 2]
        C[:, 4, 2] -= f * 2.0 * rmu * S[:, 1, 2]
        C[:, 4, 3] -= f * 2.0 * rmu * S[:, 2, 2]
        D[:, 4, 4] -= f * dmu * (S[:, 0, 0] ** 2 + S[:, 0, 1] ** 2 + S[:, 0, 2] ** 2 + S[:, 1, 0] ** 2
                                 + S[:, 1, 1] ** 2 + S[:, 1, 2] ** 2 + S[:, 2, 0] ** 2
                                 + S[:, 2, 1] ** 2 + S[:, 2, 2] ** 2)
    return T


def spatial_coefficients(p: Params, st):
    """Per-node tensors of the spatial formulation with curve=0, h=1
    (spatial.f90:356-673).  Terms multiplying metric derivatives vanish."""
    ny = p.ny
    T = {k: np.zeros((ny, 5, 5)) for k in ("G", "A", "B", "C", "D", "Vxx", "Vxy", "Vyy", "Vxz", "Vyz", "Vzz")}
    G, A, B, C, D = T["G"], T["A"], T["B"], T["C"], T["D"]
    Vxx, Vxy, Vyy, Vxz, Vyz, Vzz = T["Vxx"], T["Vxy"], T["Vyy"], T["Vxz"], T["Vyz"], T["Vzz"]
    rhom, u1m, u2m, u3m, tm = st["rhom"], st["u1m"], st["u2m"], st["u3m"], st["tm"]
    gum, grhom, gtm, S = st["gum"], st["grhom"], st["gtm"], st["S"]
    h, dhds, dhdr, dhdsr, dhdrr = st["h"], st["dhds"], st["dhdr"], st["dhdsr"], st["dhdrr"]
    divum = st["divum"]
    g1divum, g2divum, g3divum = st["g1divum"], st["g2divum"], st["g3divum"]
    rmu, dmu, rlm, dlm, con, dcon = st["rmu"], st["dmu"], st["rlm"], st["dlm"], st["con"], st["dcon"]
    gmu, gdmu, glm, gdlm, gcon, gdcon = st["gmu"], st["gdmu"], st["glm"], st["gdlm"], st["gcon"], st["gdcon"]
    Lapt = st["Lapt"]
    gM2 = GAMMA * p.Ma ** 2
    Re, Pr, Ma = p.Re, p.Pr, p.Ma
    nav = p.navier

    # continuity (spatial.f90:358-372)
    G[:, 0, 0] = 1.0
    A[:, 0, 0] = u1m / h; A[:, 0, 1] = rhom / h
    B[:, 0, 0] = u2m; B[:, 0, 2] = rhom
    C[:, 0, 0] = u3m; C[:, 0, 3] = rhom
    D[:, 0, 0] = divum; D[:, 0, 1] = grhom[:, 0] / h
    D[:, 0, 2] = grhom[:, 1] + rhom * dhdr / h; D[:, 0, 3] = grhom[:, 2]

    # x-momentum (spatial.f90:376-450)
    G[:, 1, 1] = rhom
    A[:, 1, 0] = tm / (h * gM2); A[:, 1, 1] = rhom * u1m / h; A[:, 1, 4] = rhom / (h * gM2)
    B[:, 1, 1] = rhom * u2m
    C[:, 1, 1] = rhom * u3m
    D[:, 1, 0] = u1m / h * (gum[:, 0, 0] + u2m * dhdr) + u2m * gum[:, 0, 1] + u3m * gum[:, 0, 2] + gtm[:, 0] / (h * gM2)
    D[:, 1, 1] = rhom * (gum[:, 0, 0] + u2m * dhdr) / h
    D[:, 1, 2] = rhom * (gum[:, 0, 1] + u1m * dhdr / h)
    D[:, 1, 3] = rhom * gum[:, 0, 2]
    D[:, 1, 4] = grhom[:, 0] / (h * gM2)
    if nav:
        f = p.rlme / (p.rmue * Re)
        A[:, 1, 1] -= f * (glm[:, 0] / h ** 2 - rlm / h ** 3 * dhds)
        A[:, 1, 2] -= f * rlm / h ** 2 * dhdr
        A[:, 1, 4] -= f * dlm * divum / h
        B[:, 1, 2] -= f * (glm[:, 0] / h)
        C[:, 1, 3] -= f * (glm[:, 0] / h)
        D[:, 1, 2] -= f * (glm[:, 0] * dhdr / h ** 2 - rlm / h ** 3 * dhds * dhdr + rlm / h ** 2 * dhdsr)
        D[:, 1, 4] -= f * (gdlm[:, 0] * divum / h + dlm * g1divum)
        Vxx[:, 1, 1] = f * rlm / h ** 2; Vxy[:, 1, 2] = f * rlm / h; Vxz[:, 1, 3] = f * rlm / h
        f = 1.0 / Re
        A[:, 1, 1] -= f * (2.0 * gmu[:, 0] / h ** 2 - 2.0 * rmu * dhds / h ** 3)
        A[:, 1, 2] -= f * (gmu[:, 1] / h + rmu * 3.0 * dhdr / h ** 2)
        A[:, 1, 3] -= f * gmu[:, 2] / h
        A[:, 1, 4] -= f * dmu * 2.0 * S[:, 0, 0] / h
        B[:, 1, 1] -= f * (gmu[:, 1] + rmu * dhdr / h)
        B[:, 1, 4] -= f * dmu * 2.0 * S[:, 0, 1]
        C[:, 1, 1] -= f * gmu[:, 2]
        C[:, 1, 4] -= f * dmu * 2.0 * S[:, 0, 2]
        D[:, 1, 1] -= f * (gmu[:, 1] / h * (-dhdr) - rmu * (dhdr ** 2 + dhdrr * h) / h ** 2)
        D[:, 1, 2] -= f * (2.0 * gmu[:, 0] / h ** 2 * dhdr + 2.0 * rmu * (dhdsr * h - dhds * dhdr) / h ** 3)
        D[:, 1, 4] -= f * 2.0 * (gdmu[:, 0] / h * S[:, 0, 0] + gdmu[:, 1] * S[:, 0, 1]
                                 + gdmu[:, 2] * S[:, 0, 2] + dmu * st["S1jj"])
        Vxx[:, 1, 1] += f * 2.0 * rmu / h ** 2; Vxy[:, 1, 2] += f * rmu / h; Vyy[:, 1, 1] += f * rmu
        Vxz[:, 1, 3] += f * rmu / h; Vzz[:, 1, 1] += f * rmu

    # y-momentum (spatial.f90:454-526)
    G[:, 2, 2] = rhom
    A[:, 2, 2] = rhom * u1m / h
    B[:, 2, 0] = tm / gM2; B[:, 2, 2] = rhom * u2m; B[:, 2, 4] = rhom / gM2
    C[:, 2, 2] = rhom * u3m
    D[:, 2, 0] = u1m / h * (gum[:, 1, 0] - u1m * dhdr) + u2m * gum[:, 1, 1] + u3m * gum[:, 1, 2] + gtm[:, 1] / gM2
    D[:, 2, 1] = rhom * (gum[:, 1, 0] - 2.0 * u1m * dhdr) / h
    D[:, 2, 2] = rhom * gum[:, 1, 1]
    D[:, 2, 3] = rhom * gum[:, 1, 2]
    D[:, 2, 4] = grhom[:, 1] / gM2
    if nav:
        f = p.rlme / (p.rmue * Re)
        A[:, 2, 1] -= f * (glm[:, 1] / h - rlm * dhdr / h ** 2)
        B[:, 2, 2] -= f * (glm[:, 1] + rlm * dhdr / h)
        B[:, 2, 4] -= f * dlm * divum
        C[:, 2, 3] -= f * (glm[:, 1])
        D[:, 2, 2] -= f * (glm[:, 1] / h * dhdr - rlm * dhdr / h ** 2 * dhdr + rlm / h * dhdrr)
        D[:, 2, 4] -= f * (gdlm[:, 1] * divum + dlm * g2divum)
        Vxy[:, 2, 1] = f * rlm / h; Vyy[:, 2, 2] = f * rlm; Vyz[:, 2, 3] = f * rlm
        f = 1.0 / Re
        A[:, 2, 1] += f * rmu * 3.0 * dhdr / h ** 2
        A[:, 2, 2] -= f * (gmu[:, 0] / h ** 2 - rmu * dhds / h ** 3)
        A[:, 2, 4] -= f * dmu * 2.0 * S[:, 1, 0] / h
        B[:, 2, 1] -= f * gmu[:, 0] / h
        B[:, 2, 2] -= f * (2.0 * gmu[:, 1] + 2.0 * rmu * dhdr / h)
        B[:, 2, 3] -= f * gmu[:, 2]
        B[:, 2, 4] -= f * dmu * 2.0 * S[:, 1, 1]
        C[:, 2, 2] -= f * gmu[:, 2]
        C[:, 2, 4] -= f * dmu * 2.0 * S[:, 1, 2]
        D[:, 2, 1] -= f * (gmu[:, 0] / h ** 2 * (-dhdr) + rmu * (dhds * dhdr - h * dhdsr) / h ** 3)
        D[:, 2, 2] += f * 2.0 * rmu * dhdr ** 2 / h ** 2
        D[:, 2, 4] -= f * 2.0 * (gdmu[:, 0] / h * S[:, 1, 0] + gdmu[:, 1] * S[:, 1, 1]
                                 + gdmu[:, 2] * S[:, 1, 2] + dmu * st["S2jj"])
        Vxx[:, 2, 2] += f * rmu / h ** 2; Vxy[:, 2, 1] += f * rmu / h; Vyy[:, 2, 2] += f * 2.0 * rmu
        Vyz[:, 2, 3] += f * rmu; Vzz[:, 2, 2] += f * rmu

    # z-momentum (spatial.f90:530-595)
    G[:, 3, 3] = rhom
    A[:, 3, 3] = rhom * u1m / h
    B[:, 3, 3] = rhom * u2m
    C[:, 3, 0] = tm / gM2; C[:, 3, 3] = rhom * u3m; C[:, 3, 4] = rhom / gM2
    D[:, 3, 0] = u1m * gum[:, 2, 0] / h + u2m * gum[:, 2, 1] + u3m * gum[:, 2, 2] + gtm[:, 2] / gM2
    D[:, 3, 1] = rhom * gum[:, 2, 0] / h
    D[:, 3, 2] = rhom * gum[:, 2, 1]
    D[:, 3, 3] = rhom * gum[:, 2, 2]
    D[:, 3, 4] = grhom[:, 2] / gM2
    if nav:
        f = p.rlme / (p.rmue * Re)
        A[:, 3, 1] -= f * glm[:, 2] / h
        B[:, 3, 2] -= f * glm[:, 2]
        C[:, 3, 3] -= f * glm[:, 2]
        C[:, 3, 2] -= f * rlm / h * dhdr
        C[:, 3, 4] -= f * dlm * divum
        D[:, 3, 2] -= f * (glm[:, 1] / h * dhdr)
        D[:, 3, 4] -= f * (gdlm[:, 2] * divum + dlm * g3divum)
        Vxz[:, 3, 1] = f * rlm / h; Vyz[:, 3, 2] = f * rlm; Vzz[:, 3, 3] = f * rlm
        f = 1.0 / Re
        A[:, 3, 3] -= f * (gmu[:, 0] / h ** 2 - rmu * dhds / h ** 3)
        A[:, 3, 4] -= f * dmu * 2.0 * S[:, 2, 0] / h
        B[:, 3, 3] -= f * (gmu[:, 1] + rmu * dhdr / h)
        B[:, 3, 4] -= f * dmu * 2.0 * S[:, 2, 1]
        C[:, 3, 1] -= f * gmu[:, 0] / h
        C[:, 3, 2] -= f * (gmu[:, 1] + rmu * dhdr / h)
        C[:, 3, 3] -= f * 2.0 * gmu[:, 2]
        C[:, 3, 4] -= f * dmu * 2.0 * S[:, 2, 2]
        D[:, 3, 4] -= f * 2.0 * (gdmu[:, 0] / h * S[:, 2, 0] + gdmu[:, 1] * S[:, 2, 1]
                                 + gdmu[:, 2] * S[:, 2, 2] + dmu * st["S3jj"])
        Vxx[:, 3, 3] += f * rmu / h ** 2; Vyy[:, 3, 3] += f * rmu; Vxz[:, 3, 1] += f * rmu / h
        Vyz[:, 3, 2] += f * rmu; Vzz[:, 3, 3] += f * 2.0 * rmu

    # energy (spatial.f90:599-673)
    G[:, 4, 4] = rhom
    A[:, 4, 1] = rhom * GAMMA1 * tm / h; A[:, 4, 4] = rhom * u1m / h
    B[:, 4, 2] = rhom * GAMMA1 * tm; B[:, 4, 4] = rhom * u2m
    C[:, 4, 3] = rhom * GAMMA1 * tm; C[:, 4, 4] = rhom * u3m
    D[:, 4, 0] = u1m / h * gtm[:, 0] + u2m * gtm[:, 1] + u3m * gtm[:, 2] + GAMMA1 * tm * divum
    D[:, 4, 1] = rhom * gtm[:, 0] / h
    D[:, 4, 2] = rhom * gtm[:, 1] + rhom * GAMMA1 * tm * dhdr / h
    D[:, 4, 3] = rhom * gtm[:, 2]
    D[:, 4, 4] = rhom * GAMMA1 * divum
    if nav:
        f = GAMMA / (Pr * Re)
        A[:, 4, 4] -= f * (gcon[:, 0] / h ** 2 + dcon * gtm[:, 0] / h ** 2 - con * dhds / h ** 3)
        B[:, 4, 4] -= f * (gcon[:, 1] + dcon * gtm[:, 1] + con * dhdr / h)
        C[:, 4, 4] -= f * (gcon[:, 2] + dcon * gtm[:, 2])
        D[:, 4, 4] -= f * (gdcon[:, 0] * gtm[:, 0] / h ** 2 + gdcon[:, 1] * gtm[:, 1]
                           + gdcon[:, 2] * gtm[:, 2] + dcon * Lapt)
        Vxx[:, 4, 4] = f * con / h ** 2; Vyy[:, 4, 4] = f * con; Vzz[:, 4, 4] = f * con
        f = GAMMA * GAMMA1 * Ma ** 2 * p.rlme / (Re * p.rmue)
        A[:, 4, 1] -= f * 2.0 * rlm * divum / h
        B[:, 4, 2] -= f * 2.0 * rlm * divum
        C[:, 4, 3] -= f * 2.0 * rlm * divum
        D[:, 4, 2] -= f * 2.0 * rlm * divum * dhdr / h
        D[:, 4, 4] -= f * dlm * divum * divum
        f = GAMMA * GAMMA1 * Ma ** 2 / Re
        A[:, 4, 1] -= f * 4.0 * rmu * S[:, 0, 0] / h
        A[:, 4, 2] -= f * 4.0 * rmu * S[:, 1, 0] / h
        A[:, 4, 3] -= f * 4.0 * rmu * S[:, 2, 0] / h
        B[:, 4, 1] -= f * 4.0 * rmu * S[:, 0, 1]
        B[:, 4, 2] -= f * 4.0 * rmu * S[:, 1, 1]
        B[:, 4, 3] -= f * 4.0 * rmu * S[:, 2, 1]
        C[:, 4, 1] -= f * 4.0 * rmu * S[:, 0, 2]
        C[:, 4, 2] -= f * 4.0 * rmu * S[:, 1, 2]
        C[:, 4, 3] -= f * 4.0 * rmu * S[:, 2, 2]
        D[:, 4, 1] += f * 4.0 * rmu * S[:, 1, 0] * dhdr / h
        D[:, 4, 2] -= f * 4.0 * rmu * S[:, 0, 0] * dhdr / h
        D[:, 4, 4] -= f * 2.0 * dmu * (S[:, 0, 0] ** 2 + S[:, 0, 1] ** 2 + S[:, 0, 2] ** 2 + S[:, 1, 0] ** 2
                                       + S[:, 1, 1] ** 2 + S[:, 1, 2] ** 2 + S[:, 2, 0] ** 2
                                       + S[:, 2, 1] ** 2 + S[:, 2, 2] ** 2)
    return T


# ----------------------------------------------------------------------------
# a7: temporal pencil (temporal.f90:600-752)
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Problem:
    p: Params
    y: np.ndarray
    eta: np.ndarray
    deta: np.ndarray
    d2eta: np.ndarray
    vm: np.ndarray
    d1: np.ndarray
    d2: np.ndarray
    dt1: np.ndarray
    dt2: np.ndarray
    T: dict
    dg: Optional[np.ndarray] = None    # finite-difference path: d/deta of the mean flow (grad.f90)


def setup(p: Params, ym, vt, spatial: bool) -> Problem:
    """Grid, mean flow, operators and coefficient tensors for one point."""
    if not p.ider:
        raise ValueError("ider=0 (external derivative files) is outside this path")
    y, eta, deta, d2eta = sgengrid(p)
    vm = getmean(y, ym, vt)
    d1, d2, dt1, dt2 = derivative_operators(p)
    st = _mean_state(p, vm, d1, d2, deta, d2eta, spatial)
    T = spatial_coefficients(p, st) if spatial else temporal_coefficients(p, st)
    return Problem(p, y, eta, deta, d2eta, vm, d1, d2, dt1, dt2, T)


def _block_rows(Bh, Vyy, d1, d2, Dh, i):
    """Row block i of A0: Bh_i (x) D1(i,:) - Vyy_i (x) D2(i,:) + delta_ij Dh_i."""
    ny = d1.shape[0]
    blk = (Bh[i][:, None, :] * d1[i][None, :, None]) - (Vyy[i][:, None, :] * d2[i][None, :, None])
    blk = blk.astype(np.complex128)
    blk[:, i, :] += Dh[i]
    return blk.reshape(5, ny * 5)


def temporal_pencil(pb: Problem):
    """Dense A0 and B0 with A0 x = omega B0 x (temporal.f90:604-752)."""
    p, T = pb.p, pb.T
    ny = p.ny
    n = NDOF * ny
    al, be = p.alpha, p.beta
    im = 1j
    Dh = (T["D"] + im * al * T["A"] + im * be * T["C"]
          + al ** 2 * T["Vxx"] + al * be * T["Vxz"] + be ** 2 * T["Vzz"])
    Bh = T["B"] - im * al * T["Vxy"] - im * be * T["Vyz"]
    Bh = Bh * pb.deta[:, None, None] - T["Vyy"] * pb.d2eta[:, None, None]
    Vyy = T["Vyy"] * (pb.deta ** 2)[:, None, None]
    A0 = np.zeros((n, n), dtype=np.complex128)
    B0 = np.zeros((n, n), dtype=np.complex128)
    # B0: top node identity, interior i*G, wall rows (temporal.f90:630-662)
    for k in range(5):
        B0[k, k] = im
    for i in range(1, ny - 1):
        B0[i * 5:(i + 1) * 5, i * 5:(i + 1) * 5] = im * T["G"][i]
    w = (ny - 1) * 5
    for k in range(4):
        B0[w + k, w + k] = im
    if p.wallt == 0:
        B0[w + 4, w + 4] = im
    elif p.wallt == 2:
        B0[w + 4, w:w + 5] = im * T["G"][ny - 1, 4, :]
    else:
        raise ValueError(f"Illegal value of wallt: {p.wallt}")
    # interior rows (temporal.f90:683-703)
    for i in range(1, ny - 1):
        A0[i * 5:(i + 1) * 5, :] = _block_rows(Bh, Vyy, pb.d1, pb.d2, Dh, i)
    # continuity rows at the top and at the wall: D1 term only (:666-681, :707-722)
    for i in (0, ny - 1):
        row = (Bh[i][0][None, :] * pb.d1[i][:, None]).astype(np.complex128)   # [j, q]
        row[i, :] += Dh[i][0]
        A0[i * 5, :] = row.reshape(-1)
    if p.wallt == 2:
        # wall energy row: D1 used for both terms in cols 1-4 (temporal.f90:729-741)
        i = ny - 1
        row = np.zeros((ny, 5), dtype=np.complex128)
        for q in range(4):
            row[:, q] = Bh[i][4, q] * pb.d1[i] - Vyy[i][4, q] * pb.d1[i]
        row[:, 4] = Bh[i][4, 4] * pb.dt1[i] - Vyy[i][4, 4] * pb.dt2[i]
        row[i, :] += Dh[i][4]
        A0[i * 5 + 4, :] = row.reshape(-1)
    return A0, B0


# ----------------------------------------------------------------------------
# a11: spatial pencil L(alpha) = C0 + alpha C1 + alpha^2 C2 (spatial.f90:681-959)
# ----------------------------------------------------------------------------
def spatial_pencil(pb: Problem):
    p, T = pb.p, pb.T
    ny = p.ny
    n = NDOF * ny
    im = 1j
    be, om = p.beta, p.omega
    Dh = T["D"] + im * be * T["C"] + be ** 2 * T["Vzz"]
    Bh = T["B"] - im * be * T["Vyz"]
    Bh = Bh * pb.deta[:, None, None] - T["Vyy"] * pb.d2eta[:, None, None]
    Vyy = T["Vyy"] * (pb.deta ** 2)[:, None, None]
    C0 = np.zeros((n, n), dtype=np.complex128)
    C1 = np.zeros((n, n), dtype=np.complex128)
    C2 = np.zeros((n, n), dtype=np.complex128)
    z = np.zeros_like(Vyy)

    def cont_row(M, Bhm, Dhm, i):
        row = (Bhm[i][0][None, :] * pb.d1[i][:, None]).astype(np.complex128)
        row[i, :] += Dhm[i][0]
        M[i * 5, :] = row.reshape(-1)

    def wall_energy(M, Bhm, Vyym, Dhm, use_d2):
        i = ny - 1
        row = np.zeros((ny, 5), dtype=np.complex128)
        for q in range(4):
            row[:, q] = Bhm[i][4, q] * pb.d1[i] - (Vyym[i][4, q] * pb.d2[i] if use_d2 else 0.0)
        row[:, 4] = Bhm[i][4, 4] * pb.dt1[i] - (Vyym[i][4, 4] * pb.dt2[i] if use_d2 else 0.0)
        row[i, :] += Dhm[i][4]
        M[i * 5 + 4, :] = row.reshape(-1)

    # C0 (spatial.f90:702-788)
    if p.top == 1:
        cont_row(C0, Bh, Dh, 0)
    for i in range(1, ny - 1):
        C0[i * 5:(i + 1) * 5, :] = _block_rows(Bh, Vyy, pb.d1, pb.d2, Dh, i)
    cont_row(C0, Bh, Dh, ny - 1)
    if p.wallt == 2:
        wall_energy(C0, Bh, Vyy, Dh, True)
    # time term and boundary rows (spatial.f90:792-835)
    if p.top == 1:
        C0[0, 0] = C0[0, 0] - im * om
    else:
        C0[0, 0] = -1.0
    for k in range(1, 5):
        C0[k, k] = -1.0
    for i in range(1, ny - 1):
        C0[i * 5:(i + 1) * 5, i * 5:(i + 1) * 5] -= im * om * T["G"][i]
    w = (ny - 1) * 5
    C0[w, w] -= im * om
    for k in range(1, 4):
        C0[w + k, w + k] -= 1.0
    if p.wallt == 0:
        C0[w + 4, w + 4] -= 1.0
    elif p.wallt == 2:
        C0[w + 4, w:w + 5] -= im * om * T["G"][ny - 1, 4, :]
    else:
        raise ValueError(f"Illegal value of wallt: {p.wallt}")
    # C1 (spatial.f90:839-938)
    Dh1 = im * T["A"] + be * T["Vxz"]
    Bh1 = (-im * T["Vxy"]) * pb.deta[:, None, None]
    if p.top == 1:
        cont_row(C1, Bh1, Dh1, 0)
    for i in range(1, ny - 1):
        C1[i * 5:(i + 1) * 5, :] = _block_rows(Bh1, z, pb.d1, pb.d2, Dh1, i)
    cont_row(C1, Bh1, Dh1, ny - 1)
    if p.wallt == 2:
        wall_energy(C1, Bh1, z, Dh1, False)
    # C2 (spatial.f90:942-959)
    for i in range(1, ny - 1):
        C2[i * 5:(i + 1) * 5, i * 5:(i + 1) * 5] = T["Vxx"][i]
    if p.wallt == 2:
        C2[w + 4, w:w + 5] = T["Vxx"][ny - 1, 4, :]
    return C0, C1, C2


def spatial_linearization(C0, C1, C2):
    """Generalised form of the reference companion (spatial.f90:982-1016):
    [[C0,0],[0,I]] z = alpha [[-C1,-C2],[I,0]] z, z = [phi; alpha phi]."""
    n = C0.shape[0]
    I = np.eye(n, dtype=np.complex128)
    Z = np.zeros((n, n), dtype=np.complex128)
    Aa = np.block([[C0, Z], [Z, I]])
    Bb = np.block([[-C1, -C2], [I, Z]])
    return Aa, Bb


# ----------------------------------------------------------------------------
# a8 / a9 / a12: eigen-solves and sort
# ----------------------------------------------------------------------------
def sort_by_imag(vals, vecs=None):
    """Ascending sort by imaginary part carrying the permutation; equal keys
    keep their input order (a straight-insertion sort is stable).  Behaviour of
    the external PIKSR2 call at temporal.f90:844-855 / spatial.f90:1073-1084."""
    idx = np.argsort(vals.imag, kind="stable")
    return vals[idx], (vecs[:, idx] if vecs is not None else None)


def scale_columns(vecs, rows: Optional[int] = None):
    """Divide each column by its first entry of largest modulus
    (temporal.f90:867-879, getevec.f90:179-189)."""
    out = vecs.copy()
    m = out.shape[0] if rows is None else rows
    for j in range(out.shape[1]):
        col = out[:m, j]
        k = int(np.argmax(np.abs(col)))      # first maximum
        s = col[k]
        if s != 0:
            out[:, j] = out[:, j] / s
    return out


def temporal_solve_reference(A0, B0):
    """Reference method: ZGESV(B0, A0) then ZGEEV('N','V') (temporal.f90:771-810),
    sort by Im omega, then max-modulus scaling (:863-879)."""
    M = sla.solve(B0, A0)
    w, v = sla.eig(M, left=False, right=True, check_finite=False)
    w, v = sort_by_imag(w, v)
    v = scale_columns(v)
    return w, v


def qz(A, B, right=True):
    """QZ judge: ZGGEV on the same FP64 pencil."""
    if right:
        w, v = sla.eig(A, B, left=False, right=True, check_finite=False)
        return w, v
    return sla.eig(A, B, left=False, right=False, check_finite=False), None


def spatial_solve_reference(C0, C1, C2, ievec=1):
    """Reference method: ZGETRF/ZGETRS, companion, ZGEEV, alpha = 1/lambda
    (0 if lambda = 0), sort by Im alpha (spatial.f90:972-1084)."""
    n = C0.shape[0]
    lu = sla.lu_factor(C0)
    X1 = sla.lu_solve(lu, -C1)
    X2 = sla.lu_solve(lu, -C2)
    comp = np.zeros((2 * n, 2 * n), dtype=np.complex128)
    comp[:n, :n] = X1
    comp[:n, n:] = X2
    comp[n:, :n] = np.eye(n)
    if ievec:
        lam, v = sla.eig(comp, left=False, right=True, check_finite=False)
    else:
        lam = sla.eig(comp, left=False, right=False, check_finite=False)
        v = None
    alp = np.where(lam != 0, 1.0 / np.where(lam != 0, lam, 1.0), 0.0)
    alp, v = sort_by_imag(alp, v)
    return alp, v


# ----------------------------------------------------------------------------
# a13: getevec selection, normalisation and text output (getevec.f90:160-227)
# ----------------------------------------------------------------------------
def select_nearest(evals, guess):
    """First index minimising |guess - eval| (getevec.f90:160-173)."""
    d = np.abs(guess - evals)
    return int(np.argmin(d))


def normalize_evec(vec, nrows):
    """Max-modulus normalisation over the first 5N components (getevec.f90:179-189)."""
    col = vec[:nrows]
    k = int(np.argmax(np.abs(col)))
    s = col[k]
    return vec / s if s != 0 else vec.copy()


def fmt_e(v: float) -> str:
    """Fortran 1pe21.13E3 field."""
    if v == 0.0:
        mant, ex = 0.0, 0
    else:
        ex = int(math.floor(math.log10(abs(v))))
        mant = v / 10.0 ** ex
        s = "%.13f" % abs(mant)
        if s.startswith("10."):
            ex += 1
    s = "%.13E" % v
    m, e = s.split("E")
    e = int(e)
    return ("%s%s%03d" % (m, "E+" if e >= 0 else "E-", abs(e))).rjust(21)


def eigfunction_table(itype, Re, Ma, Pr, eval_sel, alpha, beta, omega, y, vec, ny):
    """Text of time.N / space.N (getevec.f90:193-227)."""
    out = io.StringIO()
    out.write("# Re = %s, Ma = %s, Pr = %s\n" % (_f13(Re), _f13(Ma), _f13(Pr)))
    if itype == 1:
        om, al = eval_sel, alpha
    else:
        om, al = omega, eval_sel
    out.write("# Omega = (%s,%s)\n" % (fmt_e(om.real), fmt_e(om.imag)))
    out.write("# Alpha = (%s,%s)\n" % (fmt_e(al.real), fmt_e(al.imag)))
    out.write("# Beta  = (%s,%s)\n" % (fmt_e(beta.real), fmt_e(beta.imag)))
    for i in range(ny - 1, -1, -1):
        vals = [y[i]]
        for k in range(5):
            z = vec[i * 5 + k]
            vals += [z.real, z.imag]
        out.write(" ".join(fmt_e(v) for v in vals) + " \n")
    return out.getvalue()


def _f13(v):
    s = "%.6E" % v
    m, e = s.split("E")
    return ("%sE%+03d" % (m, int(e))).rjust(13)


def parse_table(text):
    """Parse a time.N/space.N file: header eigen values and the numeric table."""
    hdr = {}
    rows = []
    for line in text.splitlines():
        if line.startswith("#"):
            if "=" in line and "(" in line:
                key = line[1:].split("=")[0].strip()
                inner = line.split("(")[1].split(")")[0]
                a, b = inner.split(",")
                hdr[key] = complex(float(a), float(b))
            continue
        if line.strip():
            rows.append([float(v) for v in line.split()])
    return hdr, np.array(rows)


# ----------------------------------------------------------------------------
# Leading discrete mode.  The reference leaves the choice of mode to the user
# (sorted spectrum, temporal.f90:844-879; getevec -v with a guess,
# getevec.f90:160-173, thesis/TStest/run.sh:13-17).  The rule restated here:
# a mode is discrete when its normalised eigenvector stays below 0.1 on the
# nodes of the outer eighth of the domain and its phase speed is more than
# 0.01 away from the edge velocity (vorticity/entropy continuum); spatial
# modes must also travel downstream (alpha_r > 0).  The leading mode is the
# discrete one with the largest Im(omega) (temporal) or smallest Im(alpha)
# (spatial).
# ----------------------------------------------------------------------------
FARFIELD_MAX = 0.1
CONTINUUM_BAND = 0.01


def farfield_ratio(vec, ny, nodes=None):
    """max |x| over the farfield nodes (5 variables each) / max |x|.  The
    farfield is the first ceil(ny/8) nodes (Chebyshev ordering, eta near 1)
    unless `nodes` lists others (fd_farfield_nodes)."""
    a = np.abs(np.asarray(vec)[:5 * ny])
    if nodes is None:
        nodes = np.arange((ny + 7) // 8)
    m = a.max()
    far = a.reshape(ny, 5)[np.asarray(nodes)]
    return float(far.max() / m) if m > 0 else 1.0


def fd_farfield_nodes(y, ymax):
    """Farfield of a finite-difference point (wall at node 0): the nodes of
    the outer half of the domain, y >= ymax/2 (for yi = 2/46, ymax = 50 the
    Chebyshev window above starts at y = 25.7 too)."""
    return np.nonzero(np.asarray(y) >= 0.5 * ymax)[0]


def is_discrete(itype, ev, ff, alpha_r=0.0, omega_r=0.0):
    if not ff < FARFIELD_MAX:
        return False
    if itype == 1:
        if alpha_r == 0.0:
            return True
        c = ev.real / alpha_r
    else:
        if not ev.real > 0.0:
            return False
        c = omega_r / ev.real
    return abs(c - 1.0) > CONTINUUM_BAND


def leading_mode(itype, w, V, ny, alpha_r=0.0, omega_r=0.0, candidates=None, nodes=None):
    """Index of the leading discrete mode among `candidates` (default: all
    finite, nonzero eigenvalues), or -1; `nodes` as for farfield_ratio."""
    idx = candidates if candidates is not None else [j for j in range(len(w))
                                                     if np.isfinite(w[j]) and abs(w[j]) > 1e-9]
    best, key = -1, -np.inf
    for j in idx:
        ff = farfield_ratio(V[:, j], ny, nodes)
        if not is_discrete(itype, w[j], ff, alpha_r, omega_r):
            continue
        k = w[j].imag if itype == 1 else -w[j].imag
        if k > key:
            best, key = j, k
    return best


# ----------------------------------------------------------------------------
# a14: evec.dat (gfortran sequential unformatted; temporal.f90:883-890)
# ----------------------------------------------------------------------------
def read_evec_dat(path):
    """Read the 5-record file written by temporal/spatial (4-byte markers)."""
    with open(path, "rb") as f:
        data = f.read()
    recs = []
    pos = 0
    while pos < len(data):
        (ln,) = struct.unpack_from("<i", data, pos)
        recs.append(data[pos + 4:pos + 4 + ln])
        (ln2,) = struct.unpack_from("<i", data, pos + 4 + ln)
        assert ln2 == ln
        pos += 8 + ln
    ints = struct.unpack("<10i", recs[0])
    ind, ny, ndof, itype, ievec, curve, top, wall, wallt, ider = ints
    om, al, be = (complex(*struct.unpack_from("<2d", recs[1], 16 * k)) for k in range(3))
    Re, Ma, Pr = struct.unpack_from("<3d", recs[1], 48)
    r3 = np.frombuffer(recs[2], dtype="<f8")
    x = r3[0]
    y = r3[1:1 + ny]
    nmax = ndof * ny if itype == 1 else 2 * ndof * ny
    ev = np.frombuffer(recs[3], dtype="<c16")
    out = dict(ind=ind, ny=ny, ndof=ndof, itype=itype, ievec=ievec, curve=curve, top=top,
               wall=wall, wallt=wallt, ider=ider, omega=om, alpha=al, beta=be, Re=Re, Ma=Ma,
               Pr=Pr, x=x, y=y, eta=r3[1 + ny:1 + 2 * ny], deta=r3[1 + 2 * ny:1 + 3 * ny],
               d2eta=r3[1 + 3 * ny:1 + 4 * ny], yi=r3[1 + 4 * ny], ymax=r3[2 + 4 * ny], evals=ev)
    if len(recs) > 4:
        out["evec"] = np.frombuffer(recs[4], dtype="<c16").reshape(-1, ev.size, order="F") \
            if len(recs[4]) else None
    return out


# ----------------------------------------------------------------------------
# convenience: full reference-method runs
# ----------------------------------------------------------------------------
# ----------------------------------------------------------------------------
# a15: finite-difference temporal path (fd_temporal.f90, itype 3)
# ----------------------------------------------------------------------------
# Explicit stencils of module `stencils` (stencils.f90:1-128) that fd_temporal
# uses, in units of 1/deta (first derivative) and 1/deta^2 (second).  The
# rational ones are written as fractions.  The third-order boundary closures
# of Carpenter, Gottlieb & Abarbanel (1994) for the fourth-order interior
# scheme (gg, gh, gi, gj) are fixed by a stability condition as well as by
# accuracy, so they are the decimals of stencils.f90:56-90.
FD_GA = (1.0 / 12.0, -2.0 / 3.0, 2.0 / 3.0, -1.0 / 12.0)          # nodes i-2, i-1, i+1, i+2 (:9-12)
FD_GG = (-1.87603205562073772290017534999, 3.18538322557789286703668141674,
         -1.81454567943752757247830550034, 0.59165824105260274421658150053,
         -0.101052068000505624644095417028, 0.0145883364282753087693133500917)     # row 1, nodes 1-6
FD_GH = (-0.384254232677925402043233692192, -0.290638947767348681070279077173,
         0.67176478451531541138011989728, 0.071081659837399872713651693126,
         -0.073630718761724245070378308432, 0.0056774548542830440901194873933)     # row 2, nodes 1-6
FD_GI = (0.182885278686826206583549109115, -1.08001475417455516433881053795,
         0.65787289649662525818641772732, 0.1776170486891931456381189546,
         0.076779836395827558602005515075, -0.0151403060939170046712807681567)     # row 3, nodes 1-6
FD_GJ = (-0.0341837103365291857798396404953, 0.224829025740103121729831613399,
         -0.89081233292845396245426338198, 0.165299947710035014782196870489,
         0.61343955208752529977826815383, -0.078572482272680288056193615249)       # row 4, nodes 1-6
FD_DA = (-1.0 / 12.0, 4.0 / 3.0, -5.0 / 2.0, 4.0 / 3.0, -1.0 / 12.0)                # (:97-101)
FD_DB = (11.0 / 12.0, -5.0 / 3.0, 1.0 / 2.0, 1.0 / 3.0, -1.0 / 12.0)                # (:105-109)
FD_DD = (35.0 / 12.0, -26.0 / 3.0, 19.0 / 2.0, -14.0 / 3.0, 11.0 / 12.0)            # (:120-124)


def ogengrid(p: Params):
    """Uniform eta = (j-1)/(ny-1) with node 0 at the wall and the algebraic
    map y = aa eta/(bb - eta), aa = ymax yi/(ymax - 2 yi), bb = 1 + aa/ymax
    (ogengrid.f90:17-48; half the nodes lie below y = yi).  The yi = 0 branch
    (ogengrid.f90:49-73) reads its parameters from stdin and is not restated.
    Returns (y, eta, deta, d2eta, dy)."""
    ny = p.ny
    if p.yi == 0.0:
        raise ValueError("yi = 0 selects the interactive Mahesh map (ogengrid.f90:49-73)")
    if not p.ymax > 2.0 * p.yi:
        raise ValueError("ogengrid needs ymax > 2 yi")
    dy = 1.0 / float(ny - 1)
    aa = p.ymax * p.yi / (p.ymax - 2.0 * p.yi)
    bb = 1.0 + aa / p.ymax
    eta = np.arange(ny, dtype=np.float64) * dy
    y = aa * eta / (bb - eta)
    deta = (bb - eta) ** 2 / (aa * bb)
    d2eta = -2.0 * (bb - eta) ** 3 / (aa * bb) ** 2
    return y, eta, deta, d2eta, dy


def fd_grad_matrix(ny: int, dy: float):
    """d/deta of the mean flow as a matrix: grad.f90:47-85 (the y part of
    subroutine grad), fourth-order explicit, five-point one-sided and biased
    rows at both ends, the ones at the far end mirrored and negated."""
    G = np.zeros((ny, ny))
    w = 12.0 * dy
    G[0, 0:5] = np.array([-25.0, 48.0, -36.0, 16.0, -3.0]) / w
    G[1, 0:5] = np.array([-3.0, -10.0, 18.0, -6.0, 1.0]) / w
    for i in range(2, ny - 2):
        G[i, i - 2:i + 3] = np.array([1.0, -8.0, 0.0, 8.0, -1.0]) / w
    G[ny - 2, ny - 5:ny] = -np.array([1.0, -6.0, 18.0, -10.0, -3.0]) / w
    G[ny - 1, ny - 5:ny] = -np.array([-3.0, 16.0, -36.0, 48.0, -25.0]) / w
    return G


def fd_operators(ny: int, dy: float):
    """First and second eta-derivative rows of the finite-difference pencil,
    as used by fd_temporal.f90:630-890 (node 0 the wall, node ny-1 the top,
    whose rows are Dirichlet and stay zero here).  Kept as the reference
    writes them, including the sign pattern of the second-derivative rows at
    nodes ny-4 and ny-3 (fd_temporal.f90:826-878: the entries right of the
    diagonal enter with the opposite sign, except the last one of row ny-3)."""
    if ny < 10:
        raise ValueError("the finite-difference closures need ny >= 10")
    D1 = np.zeros((ny, ny))
    D2 = np.zeros((ny, ny))
    h1, h2 = 1.0 / dy, 1.0 / (dy * dy)
    D1[0, 0:6] = np.array(FD_GG) * h1
    D2[0, 0:5] = np.array(FD_DD) * h2                   # wall energy row (:666-700)
    D1[1, 0:6] = np.array(FD_GH) * h1
    D2[1, 0:5] = np.array(FD_DB) * h2                   # :705-733
    D1[2, 0:6] = np.array(FD_GI) * h1
    D2[2, 0:5] = np.array(FD_DA) * h2                   # :735-766
    D1[3, 0:6] = np.array(FD_GJ) * h1
    D2[3, 1:6] = np.array(FD_DA) * h2                   # :768-793
    ga = np.array([FD_GA[0], FD_GA[1], 0.0, FD_GA[2], FD_GA[3]])
    for i in range(4, ny - 4):                          # :795-812
        D1[i, i - 2:i + 3] = ga * h1
        D2[i, i - 2:i + 3] = np.array(FD_DA) * h2
    r = ny - 4                                          # :814-838
    D1[r, r - 2:r + 4] = -np.array(FD_GJ[::-1]) * h1
    D2[r, r - 2:r + 3] = np.array(FD_DA) * np.array([1.0, 1.0, 1.0, -1.0, -1.0]) * h2
    r = ny - 3                                          # :840-864
    D1[r, r - 3:r + 3] = -np.array(FD_GI[::-1]) * h1
    D2[r, r - 2:r + 3] = np.array(FD_DA) * np.array([1.0, 1.0, 1.0, -1.0, 1.0]) * h2
    r = ny - 2                                          # :866-890
    D1[r, r - 4:r + 2] = -np.array(FD_GH[::-1]) * h1
    D2[r, r - 3:r + 2] = np.array(FD_DB[::-1]) * h2
    return D1, D2


def fd_spatial_operators(ny: int, dy: float):
    """The derivative rows of the finite-difference spatial pencil
    (fd_spatial.f90:669-960): rows 0-3 as fd_operators, the central rows on
    nodes 4 .. ny-3, row ny-2 the second-to-last row of grad.f90 (D1) with the
    mirrored biased D2 row, row ny-1 (the top continuity row when top = 1,
    :881-903) the last row of grad.f90 and no D2."""
    if ny < 10:
        raise ValueError("the finite-difference closures need ny >= 10")
    D1, D2 = fd_operators(ny, dy)
    h1, h2 = 1.0 / dy, 1.0 / (dy * dy)
    ga = np.array([FD_GA[0], FD_GA[1], 0.0, FD_GA[2], FD_GA[3]])
    for i in range(ny - 4, ny - 2):                      # :834-852
        D1[i, :] = 0.0
        D2[i, :] = 0.0
        D1[i, i - 2:i + 3] = ga * h1
        D2[i, i - 2:i + 3] = np.array(FD_DA) * h2
    G = fd_grad_matrix(ny, dy)
    D1[ny - 2, :] = G[ny - 2]                            # :854-879
    D1[ny - 1, :] = G[ny - 1]                            # :881-903
    return D1, D2


def fd_setup(p: Params, ym, vt, spatial: bool = False) -> Problem:
    """Grid, mean flow, derivative rows and coefficient tensors of one
    finite-difference point (fd_temporal.f90:87-578, fd_spatial.f90:95-648):
    the mean-flow derivatives are the grad stencils applied twice in eta, then
    mapped to y (fd_temporal.f90:131-151); the tensors are those of
    temporal.f90:309-598 / spatial.f90:354-673 (each pair of files forms them
    identically)."""
    if not p.ider:
        raise ValueError("ider=0 (external derivative files) is outside this path")
    y, eta, deta, d2eta, dy = ogengrid(p)
    vm = getmean(y, ym, vt)
    dg = fd_grad_matrix(p.ny, dy)
    g2 = dg @ vm
    g22 = dg @ g2
    st = _mean_state(p, vm, None, None, deta, d2eta, spatial, raw=(g2, g22))
    T = spatial_coefficients(p, st) if spatial else temporal_coefficients(p, st)
    d1, d2 = fd_spatial_operators(p.ny, dy) if spatial else fd_operators(p.ny, dy)
    return Problem(p, y, eta, deta, d2eta, vm, d1, d2, d1, d2, T, dg)


def fd_temporal_pencil(pb: Problem):
    """Dense A0, B0 of the finite-difference temporal problem, A0 x = omega B0 x
    (fd_temporal.f90:582-890).  Rows: node 0 (wall) continuity (D1 term only)
    and energy (columns 0-3 the full row, column 4 without the D1 term),
    wall momentum rows and all top-node rows Dirichlet (A0 zero, B0 = i);
    every other row Bh D1 - Vyy D2 + Dh on the rows of fd_operators."""
    p, T = pb.p, pb.T
    ny = p.ny
    n = NDOF * ny
    al, be = p.alpha, p.beta
    im = 1j
    Dh = (T["D"] + im * al * T["A"] + im * be * T["C"]
          + al ** 2 * T["Vxx"] + al * be * T["Vxz"] + be ** 2 * T["Vzz"])
    Bh = T["B"] - im * al * T["Vxy"] - im * be * T["Vyz"]
    Bh = Bh * pb.deta[:, None, None] - T["Vyy"] * pb.d2eta[:, None, None]
    Vyy = T["Vyy"] * (pb.deta ** 2)[:, None, None]
    d1, d2 = pb.d1, pb.d2
    A0 = np.zeros((n, n), dtype=np.complex128)
    B0 = np.zeros((n, n), dtype=np.complex128)
    for k in range(4):                                    # :606-608
        B0[k, k] = im
    B0[4, 0:5] = im * T["G"][0, 4, :]                     # :610-615
    for i in range(1, ny - 1):                            # :617-624
        B0[i * 5:(i + 1) * 5, i * 5:(i + 1) * 5] = im * T["G"][i]
    w = (ny - 1) * 5
    for k in range(5):                                    # :626-629
        B0[w + k, w + k] = im
    # wall continuity row (:633-656)
    row = (Bh[0][0][None, :] * d1[0][:, None]).astype(np.complex128)
    row[0, :] += Dh[0][0]
    A0[0, :] = row.reshape(-1)
    # wall energy row (:658-700)
    row = np.zeros((ny, 5), dtype=np.complex128)
    for q in range(4):
        row[:, q] = Bh[0][4, q] * d1[0] - Vyy[0][4, q] * d2[0]
    row[:, 4] = -Vyy[0][4, 4] * d2[0]
    row[0, :] += Dh[0][4]
    A0[4, :] = row.reshape(-1)
    for i in range(1, ny - 1):                            # :703-890
        A0[i * 5:(i + 1) * 5, :] = _block_rows(Bh, Vyy, d1, d2, Dh, i)
    return A0, B0


def fd_spatial_pencil(pb: Problem):
    """C0, C1, C2 of the finite-difference spatial problem (fd_spatial.f90:650-
    1190), L(alpha) = C0 + alpha C1 + alpha^2 C2.  Rows: wall continuity (D1
    term only); wall energy with wallt = 2 as in fd_temporal (columns 0-3 the
    full row, column 4 without the D1 term; C1 column 4 Dh1 only), Dirichlet
    (C0 = -1) with wallt = 0; wall momentum rows Dirichlet; top continuity
    generic on the last grad row when top = 1, else Dirichlet, the other top
    rows Dirichlet; every other row Bh D1 - Vyy D2 + Dh.  The wall energy row
    takes its time term with G of the node the assembly loops ended on (node
    ny-2, or ny-1 with top = 1: :883-904 reuse the loop index i)."""
    p, T = pb.p, pb.T
    ny = p.ny
    n = NDOF * ny
    im = 1j
    be, om = p.beta, p.omega
    Dh = T["D"] + im * be * T["C"] + be ** 2 * T["Vzz"]
    Bh = (T["B"] - im * be * T["Vyz"]) * pb.deta[:, None, None] - T["Vyy"] * pb.d2eta[:, None, None]
    Vyy = T["Vyy"] * (pb.deta ** 2)[:, None, None]
    d1, d2 = pb.d1, pb.d2
    C0 = np.zeros((n, n), dtype=np.complex128)
    C1 = np.zeros((n, n), dtype=np.complex128)
    C2 = np.zeros((n, n), dtype=np.complex128)
    z = np.zeros_like(Vyy)

    def cont_row(M, Bhm, Dhm, i):
        row = (Bhm[i][0][None, :] * d1[i][:, None]).astype(np.complex128)
        row[i, :] += Dhm[i][0]
        M[i * 5, :] = row.reshape(-1)

    def wall_energy(M, Bhm, Vyym, Dhm):
        row = np.zeros((ny, 5), dtype=np.complex128)
        for q in range(4):
            row[:, q] = Bhm[0][4, q] * d1[0] - Vyym[0][4, q] * d2[0]
        row[:, 4] = -Vyym[0][4, 4] * d2[0]
        row[0, :] += Dhm[0][4]
        M[4, :] = row.reshape(-1)

    if p.wallt not in (0, 2):
        raise ValueError(f"Illegal value of wallt: {p.wallt}")
    # C0 (:669-923)
    cont_row(C0, Bh, Dh, 0)
    if p.wallt == 2:
        wall_energy(C0, Bh, Vyy, Dh)
    for i in range(1, ny - 1):
        C0[i * 5:(i + 1) * 5, :] = _block_rows(Bh, Vyy, d1, d2, Dh, i)
    if p.top == 1:
        cont_row(C0, Bh, Dh, ny - 1)
    C0[0, 0] -= im * om
    for k in range(1, 4):
        C0[k, k] = -1.0
    if p.wallt == 0:
        C0[4, 4] = -1.0
    else:
        iq = ny - 1 if p.top == 1 else ny - 2
        C0[4, 0:5] -= im * om * T["G"][iq, 4, :]
    for i in range(1, ny - 1):
        C0[i * 5:(i + 1) * 5, i * 5:(i + 1) * 5] -= im * om * T["G"][i]
    w = (ny - 1) * 5
    if p.top == 1:
        C0[w, w] -= im * om
    else:
        C0[w, w] = -1.0
    for k in range(1, 5):
        C0[w + k, w + k] = -1.0
    # C1 (:951-1160)
    Dh1 = im * T["A"] + be * T["Vxz"]
    Bh1 = (-im * T["Vxy"]) * pb.deta[:, None, None]
    cont_row(C1, Bh1, Dh1, 0)
    if p.wallt == 2:
        wall_energy(C1, Bh1, z, Dh1)
    for i in range(1, ny - 1):
        C1[i * 5:(i + 1) * 5, :] = _block_rows(Bh1, z, d1, d2, Dh1, i)
    if p.top == 1:
        cont_row(C1, Bh1, Dh1, ny - 1)
    # C2 (:1162-1180)
    if p.wallt == 2:
        C2[4, 0:5] = T["Vxx"][0, 4, :]
    for i in range(1, ny - 1):
        C2[i * 5:(i + 1) * 5, i * 5:(i + 1) * 5] = T["Vxx"][i]
    return C0, C1, C2


def run_fd_spatial(p: Params, ym, vt):
    """fd_spatial end to end: the companion of spatial.f90 (ZGETRF, ZGETRS,
    ZGEEV, alpha = 1/lambda, sort by Im alpha; fd_spatial.f90:1192-1290)."""
    pb = fd_setup(p, ym, vt, spatial=True)
    C = fd_spatial_pencil(pb)
    alp, v = spatial_solve_reference(*C, p.ievec)
    return pb, C, alp, v


def run_fd_temporal(p: Params, ym, vt):
    """fd_temporal end to end: ZGESV(B0, A0), ZGEEV, sort by Im omega, scale
    (fd_temporal.f90:972-1048, the same solve as temporal.f90:771-879)."""
    pb = fd_setup(p, ym, vt)
    A0, B0 = fd_temporal_pencil(pb)
    w, v = temporal_solve_reference(A0, B0)
    return pb, A0, B0, w, v


def run_temporal(p: Params, ym, vt):
    pb = setup(p, ym, vt, spatial=False)
    A0, B0 = temporal_pencil(pb)
    w, v = temporal_solve_reference(A0, B0)
    return pb, A0, B0, w, v


def run_spatial(p: Params, ym, vt):
    pb = setup(p, ym, vt, spatial=True)
    C0, C1, C2 = spatial_pencil(pb)
    alp, v = spatial_solve_reference(C0, C1, C2, p.ievec)
    return pb, (C0, C1, C2), alp, v
