"""Mode tracking along parameter sweeps.

The reference follows one eigenvalue through a sequence of eigensystem files
written by its sweep drivers (eig.<iver>, mtemporal.f90:37 / mspatial.f90:90):

  getalpha.f90:60-146   first file: the user's pick; then the eigenvalue
                        nearest the previous one;
  getab.f90:119-198     first two files: the eigenvalue nearest the user's
                        value; then the one nearest the linear extrapolation
                        e1 + (e2 - e1) / (p2 - p1) (p - p1) of the last two
                        (e1, p1 the latest) in the sweep parameter p (beta).

`select_getalpha` / `select_getab` restate that selection on given spectra.
`track` applies the same rule on the GPU without computing any spectrum: the
value the rule would search near becomes the shift of one batched
stab_solve_leading(STAB_TRACK) call per step, which returns the eigenvalue
nearest it; many curves advance together.
"""
from __future__ import annotations

from typing import Optional, Sequence

import numpy as np

from . import _lib


def nearest(evals, target) -> int:
    """First index minimising |e - target| (strict <, as getab.f90:160-170)."""
    ev = np.asarray(evals)
    best, err = 0, np.inf
    for j, e in enumerate(ev):
        d = abs(e - target)
        if d < err:
            best, err = j, d
    return best


def extrapolate(e1, p1, e2, p2, p):
    """getab.f90:190: e1 + (e2 - e1) / (p2 - p1) * (p - p1), (e1, p1) the latest."""
    return e1 + (e2 - e1) / (p2 - p1) * (p - p1)


def select_getab(spectra: Sequence, params: Sequence[float], guess0, guess1):
    """Indices chosen by getab's rule over a sequence of spectra."""
    out, e1 = [], None
    for k, (w, p) in enumerate(zip(spectra, params)):
        if k < 2:
            j = nearest(w, guess0 if k == 0 else guess1)
        else:
            j = nearest(w, extrapolate(e1, p1, e2, p2, p))
        if k >= 1:
            e2, p2 = e1, p1
        e1, p1 = w[j], p
        out.append(j)
    return out


def select_getalpha(spectra: Sequence, guess0):
    """Indices chosen by getalpha's rule (first pick by value)."""
    out, prev = [], guess0
    for w in spectra:
        j = nearest(w, prev)
        prev = w[j]
        out.append(j)
    return out


def track(solver, curves: Sequence[Sequence], params: np.ndarray, guess0, guess1=None, rule: str = "getab",
          mode: int = _lib.MODE_MIXED, maxit: int = 30):
    """Follow one eigenvalue along every curve.

    curves  ncurves lists of nsteps points (stab_point), step k of every curve
            solved in one batched call
    params  (ncurves, nsteps) sweep parameter of each point (getab's p)
    guess0, guess1
            per-curve starting values (guess1: getab's second pick; with
            rule "getalpha" only guess0 is used)
    Returns (eigenvalues (ncurves, nsteps) complex, info (ncurves, nsteps)).
    """
    nc, ns = len(curves), len(curves[0])
    assert all(len(c) == ns for c in curves)
    params = np.asarray(params, dtype=float)
    g0 = np.broadcast_to(np.asarray(guess0, dtype=complex), (nc,))
    g1 = None if guess1 is None else np.broadcast_to(np.asarray(guess1, dtype=complex), (nc,))
    ev = np.zeros((nc, ns), complex)
    info = np.zeros((nc, ns), np.int32)
    for k in range(ns):
        if k == 0:
            sh = g0
        elif rule == "getab" and k == 1 and g1 is not None:
            sh = g1
        elif rule == "getab" and k >= 2:
            sh = extrapolate(ev[:, k - 1], params[:, k - 1], ev[:, k - 2], params[:, k - 2], params[:, k])
        else:
            sh = ev[:, k - 1]
        r = solver.solve_leading([c[k] for c in curves], mode, _lib.TRACK, maxit=maxit, fallback_shifts=sh,
                                 shifts=sh)
        ev[:, k] = r["evals"]
        info[:, k] = r["info"]
    return ev, info
