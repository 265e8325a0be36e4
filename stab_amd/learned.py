"""Learned shift for the leading-mode solve.

A small MLP (stab_mlp_* in libstab_hip, trained and evaluated by the
library's own HIP kernels) maps the parameters of a temporal stability point
of the compressible Blasius family to its leading eigenvalue; the prediction
is the shift of the residual inverse iteration (stab_solve_leading with
STAB_LEAD_INVIT), with the plain FP64 Krylov-Schur path as the fallback and as
the source of every training label.  The reference has no such component: its
user supplies the guess (getevec -v, thesis/TStest/run.sh:13-17).

Inputs  x = (alpha, beta, log Re, M, Tw/Taw), each mapped to [-1, 1] over the
        family's ranges.
Targets t = (omega_r / alpha, omega_i / alpha), the complex phase speed of the
        leading mode, standardised by the training set's mean and spread.

Mean-flow family: compressible Blasius layers (Sutherland viscosity,
Te = 288.15 K, Pr = 0.72, gamma = 1.4) on a grid of Mach numbers and wall
temperatures Tw = (Tw/Taw) Taw; each grid node is one profile slot of the
solver context, and a point takes the (M, Tw/Taw) of its node.
"""
from __future__ import annotations

import argparse
import ctypes as C
import dataclasses
import json
import math
import os
import sys
import time
from typing import Sequence

import numpy as np

from . import _lib
from . import meanflow as mf
from . import solver as sv

NIN, NOUT = 5, 2
RANGES = {"alpha": (0.05, 0.4), "beta": (0.0, 0.2), "logRe": (math.log(300.0), math.log(3000.0)),
          "M": (0.1, 2.5), "Tw": (0.5, 1.5)}
# shift_mlp.npz: width 256, 400,000 labels on a 13 x 5 (M, Tw/Taw) grid (the benchmark's network);
# shift_mlp_config3.npz: width 128, 4 layers, 200,000 labels on a 26 x 11 grid (the held-out family test)
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "shift_mlp.npz")
WEIGHTS_CONFIG3 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "shift_mlp_config3.npz")
# spatial points (complex alpha at real omega): inputs (omega, beta, log Re, M, Tw/Taw), targets the
# leading mode's alpha / omega (inverse complex phase speed)
RANGES_SPATIAL = {"omega": (0.02, 0.15), "beta": (0.0, 0.2), "logRe": (math.log(500.0), math.log(2500.0)),
                  "M": (1.5, 2.5), "Tw": (0.5, 1.5)}
WEIGHTS_ALPHA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "alpha_mlp.npz")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


class ShiftMLP:
    """Device-resident network of one solver context (stab_mlp_*)."""

    def __init__(self, solver: sv.Solver, width: int = 128, depth: int = 3, seed: int = 1, ranges=None):
        self.S, self.lib = solver, solver.lib
        self.width, self.depth = width, depth
        self.h = C.c_void_p()
        self.S._check(self.lib.stab_mlp_create(solver.ctx, NIN, width, depth, NOUT, seed, C.byref(self.h)))
        self.np = self.lib.stab_mlp_nparams(self.h)
        self.scaling = None
        rg = (ranges or RANGES).values()
        in_shift = np.array([0.5 * (a + b) for a, b in rg], np.float32)
        in_scale = np.array([2.0 / (b - a) for a, b in (ranges or RANGES).values()], np.float32)
        self.set_scaling(in_shift, in_scale, np.zeros(NOUT, np.float32), np.ones(NOUT, np.float32))

    def close(self):
        if self.h:
            self.lib.stab_mlp_destroy(self.h)
            self.h = C.c_void_p()

    def set_scaling(self, in_shift, in_scale, out_shift, out_scale):
        arrs = [np.ascontiguousarray(a, dtype=np.float32) for a in (in_shift, in_scale, out_shift, out_scale)]
        self.S._check(self.lib.stab_mlp_set_scaling(self.h, *[_fp(a) for a in arrs]))
        self.scaling = arrs

    def get_params(self):
        p = np.zeros(self.np, np.float32)
        self.S._check(self.lib.stab_mlp_get_params(self.h, _fp(p)))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.shape == (self.np,)
        self.S._check(self.lib.stab_mlp_set_params(self.h, _fp(p)))

    def predict(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((x.shape[0], NOUT), np.float32)
        self.S._check(self.lib.stab_mlp_predict(self.h, x.shape[0], _fp(x), _fp(y)))
        return y

    def loss_grad(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.np, np.float32)
        self.S._check(self.lib.stab_mlp_loss_grad(self.h, x.shape[0], _fp(x), _fp(t), _fp(loss), _fp(g)))
        return float(loss[0]), g

    def train(self, x, t, batch=1024, steps=1000, lr=1e-3, seed=7):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        hist = np.zeros(steps, np.float32)
        self.S._check(self.lib.stab_mlp_train(self.h, x.shape[0], _fp(x), _fp(t), batch, steps, lr, seed, _fp(hist)))
        return hist

    def save(self, path, meta=None):
        os.makedirs(os.path.dirname(path), exist_ok=True)
        np.savez(path, params=self.get_params(), in_shift=self.scaling[0], in_scale=self.scaling[1],
                 out_shift=self.scaling[2], out_scale=self.scaling[3], width=self.width, depth=self.depth,
                 meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, solver: sv.Solver, path: str = WEIGHTS):
        d = np.load(path, allow_pickle=False)
        m = cls(solver, int(d["width"]), int(d["depth"]))
        m.set_scaling(d["in_shift"], d["in_scale"], d["out_shift"], d["out_scale"])
        m.set_params(d["params"])
        m.meta = json.loads(str(d["meta"]))
        return m


# ---------------------------------------------------------------------------
# Mean-flow family and points
# ---------------------------------------------------------------------------
@dataclasses.dataclass
class Family:
    machs: Sequence[float] = (0.01,) + tuple(np.round(np.linspace(0.1, 2.5, 25), 6))
    tws: Sequence[float] = tuple(np.round(np.linspace(0.5, 1.5, 11), 6))
    nrows: int = 1000
    Te: float = 288.15

    def nodes(self):
        return [(float(m), float(w)) for m in self.machs for w in self.tws]


def _profile_job(args):
    m, w, nrows, Te = args
    return mf.profile(mf.BlasiusSpec(Ma=m, Tw_aw=w, Te=Te, nrows=nrows))


def heldout_family(nprof: int = 128, per: int = 64, seed: int = 2025, ny: int = 128, nev: int = 8,
                   workers: int = 8, slot0: int = 0):
    """Held-out points over continuous draws of the mean-flow family: nprof
    draws of M ~ U[0, 2.5] (below 1e-3 moved to 1e-3: M = 0 has no compressible
    pressure scale) and Tw/Taw ~ U[0.5, 1.5], none on the training grid, each
    with `per` points (alpha ~ U[0.05, 0.4], beta ~ U[0, 0.2], log Re ~
    U[log 300, log 3000]).  Returns (tables, points, features, profile of each
    point, machs, tws); profile k goes to slot slot0 + k."""
    rng = np.random.default_rng(seed)
    machs = np.maximum(rng.uniform(0.0, 2.5, nprof), 1e-3)
    tws = rng.uniform(0.5, 1.5, nprof)
    jobs = [(float(m), float(w), 1000, 288.15) for m, w in zip(machs, tws)]
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(min(workers, len(jobs))) as pool:
            tables = pool.map(_profile_job, jobs)
    else:
        tables = [_profile_job(j) for j in jobs]
    n = nprof * per
    prof = np.repeat(np.arange(nprof), per)
    al = rng.uniform(*RANGES["alpha"], n)
    be = rng.uniform(*RANGES["beta"], n)
    re = np.exp(rng.uniform(*RANGES["logRe"], n))
    pts, feats = [], np.zeros((n, NIN), np.float32)
    for k in range(n):
        M, w = machs[prof[k]], tws[prof[k]]
        inp = sv.Inputs(mattyp=1, T0=288.15, Te=288.15, Ma=float(M), Re=float(re[k]), Pr=0.72, ny=ny,
                        yi=2.0 / 46.0, ymax=50.0, itype=1, alpha=complex(al[k], 0.0), beta=complex(be[k], 0.0))
        pts.append(sv.make_point(inp, complex(0.33 * al[k], 0.0), nev, slot0 + int(prof[k])))
        feats[k] = (al[k], be[k], math.log(re[k]), M, w)
    return tables, pts, feats, prof, machs, tws


def build_profiles(fam: Family, workers: int = 8):
    """Profile tables of every family node (host solves, in parallel)."""
    jobs = [(m, w, fam.nrows, fam.Te) for m, w in fam.nodes()]
    if workers > 1:
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(min(workers, len(jobs))) as pool:
            return pool.map(_profile_job, jobs)
    return [_profile_job(j) for j in jobs]


def load_family(S: sv.Solver, fam: Family, tables):
    for slot, tab in enumerate(tables):
        S.set_profile(slot, tab[:, 0], tab[:, 1:6])


def sample_points(fam: Family, n: int, seed: int, ny: int = 96, nev: int = 8):
    """n seeded temporal points over the family; returns (points, features)."""
    rng = np.random.default_rng(seed)
    nodes = fam.nodes()
    node = rng.integers(0, len(nodes), n)
    al = rng.uniform(*RANGES["alpha"], n)
    be = rng.uniform(*RANGES["beta"], n)
    re = np.exp(rng.uniform(*RANGES["logRe"], n))
    pts, feats = [], np.zeros((n, NIN), np.float32)
    for k in range(n):
        M, tw = nodes[node[k]]
        inp = sv.Inputs(mattyp=1, T0=fam.Te, Te=fam.Te, Ma=M, Re=float(re[k]), Pr=0.72, ny=ny, yi=2.0 / 46.0,
                        ymax=50.0, itype=1, alpha=complex(al[k], 0.0), beta=complex(be[k], 0.0))
        pts.append(sv.make_point(inp, complex(0.33 * al[k], 0.0), nev, int(node[k])))
        feats[k] = (al[k], be[k], math.log(re[k]), M, tw)
    return pts, feats


def labels(S: sv.Solver, pts, chunk: int = 8192):
    """Leading eigenvalue of every point by the plain FP64 Krylov-Schur path
    (NaN where no discrete pair converged)."""
    out = np.full(len(pts), np.nan + 1j * np.nan)
    for c0 in range(0, len(pts), chunk):
        r = S.solve_leading(pts[c0:c0 + chunk], _lib.MODE_PLAIN_F64, _lib.LEAD_KRYLOV)
        ok = r["info"] == _lib.INFO_OK
        out[c0:c0 + chunk][ok] = r["evals"][ok]
    return out


def targets(feats, omega):
    a = feats[:, 0].astype(np.float64)
    return np.stack([omega.real / a, omega.imag / a], 1).astype(np.float32)


def shifts(mlp: ShiftMLP, feats):
    """Predicted leading eigenvalues omega = alpha c of the points."""
    c = mlp.predict(feats).astype(np.float64)
    return feats[:, 0] * (c[:, 0] + 1j * c[:, 1])


def with_shifts(pts, om):
    out = []
    for p, s in zip(pts, om):
        q = sv.clone_point(p)
        q.shift.re, q.shift.im = float(s.real), float(s.imag)
        out.append(q)
    return out


def fit(S: sv.Solver, feats, tgt, width=128, depth=3, seed=1, schedule=((6000, 1e-3), (3000, 3e-4), (3000, 1e-4)),
        batch=1024, ranges=None):
    """Train a fresh network on (feats, tgt) with the GPU kernels; returns
    (network, loss history)."""
    mlp = ShiftMLP(S, width, depth, seed, ranges)
    mu, sd = tgt.mean(0), tgt.std(0) + 1e-12
    mlp.set_scaling(mlp.scaling[0], mlp.scaling[1], mu.astype(np.float32), (1.0 / sd).astype(np.float32))
    hist = []
    for k, (steps, lr) in enumerate(schedule):
        hist.append(mlp.train(feats, tgt, batch=batch, steps=steps, lr=lr, seed=seed * 1000 + k))
    return mlp, np.concatenate(hist)


def evaluate(S: sv.Solver, mlp: ShiftMLP, pts, feats, ref, maxit=20, shift_fn=None):
    """Leading-mode solves with predicted shifts (FP32 factors, FP64 fallback
    from the input shift) against reference eigenvalues `ref`."""
    fb = [complex(p.shift.re, p.shift.im) for p in pts]
    t0 = time.perf_counter()
    om = (shift_fn or shifts)(mlp, feats)
    t1 = time.perf_counter()
    r = S.solve_leading(pts, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=maxit, fallback_shifts=fb, shifts=om)
    t2 = time.perf_counter()
    ok = np.isfinite(ref)
    match = ok & (np.abs(r["evals"] - ref) <= 1e-8 * np.abs(ref))
    rel = np.abs(om - ref) / np.abs(ref)
    return dict(n=len(pts), predict_s=t1 - t0, solve_s=t2 - t1, info=np.bincount(r["info"], minlength=4).tolist(),
                match=int(match.sum()), comparable=int(ok.sum()),
                shift_rel_err_median=float(np.nanmedian(rel[ok])) if ok.any() else None,
                shift_rel_err_p95=float(np.nanpercentile(rel[ok], 95)) if ok.any() else None,
                iters_mean=float(r["iters"][r["info"] == 0].mean()) if (r["info"] == 0).any() else None,
                stats=r["stats"], result=r)


# ---------------------------------------------------------------------------
# Spatial points: the leading mode's alpha from (omega, beta, Re, M, Tw/Taw)
# ---------------------------------------------------------------------------
def spatial_input_shift(omega):
    """The input-file shift of a spatial point: alpha = omega / 0.6 - 0.002 i
    (a phase-speed guess of 0.6, slightly damped)."""
    return omega / 0.6 - 0.002j


def sample_spatial(n: int, seed: int, ny: int = 128, Ma: float = 2.0, tw: float = 1.0, slot: int = 0,
                   nev: int = 4, beta: float = 0.0):
    """n seeded spatial points on one mean flow: omega ~ U[0.02, 0.15],
    Re ~ U[500, 2500]; returns (points, features)."""
    rng = np.random.default_rng(seed)
    om = rng.uniform(*RANGES_SPATIAL["omega"], n)
    re = rng.uniform(500.0, 2500.0, n)
    pts, feats = [], np.zeros((n, NIN), np.float32)
    for k in range(n):
        inp = sv.Inputs(mattyp=1, T0=288.15, Te=288.15, Ma=Ma, Re=float(re[k]), Pr=0.72, ny=ny, yi=2.0 / 46.0,
                        ymax=50.0, itype=2, omega=complex(om[k], 0.0), beta=complex(beta, 0.0))
        pts.append(sv.make_point(inp, spatial_input_shift(om[k]), nev, slot))
        feats[k] = (om[k], beta, math.log(re[k]), Ma, tw)
    return pts, feats


def targets_spatial(feats, alpha):
    w = feats[:, 0].astype(np.float64)
    return np.stack([alpha.real / w, alpha.imag / w], 1).astype(np.float32)


def shifts_spatial(mlp: ShiftMLP, feats):
    """Predicted leading eigenvalues alpha = omega (a_r + i a_i) of spatial points."""
    c = mlp.predict(feats).astype(np.float64)
    return feats[:, 0] * (c[:, 0] + 1j * c[:, 1])


def main_spatial(argv=None):
    ap = argparse.ArgumentParser(description="alpha-shift network for spatial points on the Mach 2 Blasius layer: "
                                             "labels from the plain path, GPU training, held-out evaluation")
    ap.add_argument("--train-points", type=int, default=24576)
    ap.add_argument("--eval-points", type=int, default=2048)
    ap.add_argument("--ny", type=int, default=128)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--width", type=int, default=128)
    ap.add_argument("--seed", type=int, default=41)
    ap.add_argument("--out", default=WEIGHTS_ALPHA)
    ap.add_argument("--summary", default=None)
    a = ap.parse_args(argv)
    tab = mf.profile(mf.BlasiusSpec(Ma=2.0, Pr=0.72, law="sutherland", Te=288.15, eta_max=20.0, nrows=2000))
    S = sv.Solver(0)
    S.set_profile(0, tab[:, 0], tab[:, 1:6])
    pts, feats = sample_spatial(a.train_points, a.seed, a.ny)
    t0 = time.perf_counter()
    al = labels(S, pts, chunk=4096)
    tl = time.perf_counter() - t0
    ok = np.isfinite(al)
    print(f"# labels: {int(ok.sum())}/{len(pts)} spatial points with a leading discrete mode, {tl:.1f} s", flush=True)
    s1 = int(a.steps * 0.5)
    sched = ((s1, 1e-3), (int(a.steps * 0.25), 3e-4), (a.steps - s1 - int(a.steps * 0.25), 1e-4))
    t0 = time.perf_counter()
    mlp, hist = fit(S, feats[ok], targets_spatial(feats[ok], al[ok]), width=a.width, seed=a.seed, schedule=sched,
                    ranges=RANGES_SPATIAL)
    tt = time.perf_counter() - t0
    print(f"# trained {a.steps} steps in {tt:.1f} s, loss {hist[:10].mean():.3e} -> {hist[-100:].mean():.3e}",
          flush=True)
    ept, ef = sample_spatial(a.eval_points, a.seed + 1, a.ny)
    ref = labels(S, ept)
    ev = evaluate(S, mlp, ept, ef, ref, shift_fn=shifts_spatial)
    r = ev.pop("result")
    fb = [complex(p.shift.re, p.shift.im) for p in ept]
    base = S.solve_leading(ept, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=fb)
    ev["input_shift"] = dict(info=np.bincount(base["info"], minlength=4).tolist(),
                             match=int((np.isfinite(ref) & (np.abs(base["evals"] - ref) <= 1e-8 * np.abs(ref))).sum()),
                             op_applies=int(base["stats"]["op_applies"]), ms_total=float(base["stats"]["ms_total"]))
    ev["learned_op_applies"] = int(ev["stats"]["op_applies"])
    print(f"# held-out: {json.dumps({k: v for k, v in ev.items() if k != 'stats'})}", flush=True)
    meta = dict(kind="spatial alpha shift", train_points=a.train_points, labelled=int(ok.sum()), ny=a.ny,
                steps=a.steps, width=a.width, seed=a.seed, mean_flow="compressible Blasius M=2, adiabatic",
                loss_final=float(hist[-100:].mean()),
                heldout=dict((k, v) for k, v in ev.items() if k != "stats"))
    mlp.save(a.out, meta)
    if a.summary:
        with open(a.summary, "w") as f:
            json.dump(dict(meta=meta, label_seconds=tl, train_seconds=tt, loss_hist=hist[::100].tolist()), f)
    print(f"# saved {a.out}", flush=True)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "spatial":
        return main_spatial(argv[1:])
    ap = argparse.ArgumentParser(description="generate labels with the plain path, train the shift MLP on the GPU, "
                                             "evaluate it on held-out points, save the weights")
    ap.add_argument("--train-points", type=int, default=200000)
    ap.add_argument("--ny-train", type=int, default=96)
    ap.add_argument("--ny-eval", type=int, default=128)
    ap.add_argument("--steps", type=int, default=120000)
    ap.add_argument("--width", type=int, default=128)
    ap.add_argument("--depth", type=int, default=3)
    ap.add_argument("--seed", type=int, default=11)
    ap.add_argument("--workers", type=int, default=8)
    ap.add_argument("--out", default=WEIGHTS_CONFIG3)
    ap.add_argument("--summary", default=None)
    a = ap.parse_args(argv)
    fam = Family()
    t0 = time.perf_counter()
    tables = build_profiles(fam, a.workers)
    print(f"# {len(tables)} mean flows in {time.perf_counter() - t0:.1f} s", flush=True)
    S = sv.Solver(0)
    load_family(S, fam, tables)
    pts, feats = sample_points(fam, a.train_points, a.seed, a.ny_train)
    t0 = time.perf_counter()
    om = labels(S, pts)
    tl = time.perf_counter() - t0
    ok = np.isfinite(om)
    print(f"# labels: {int(ok.sum())}/{len(pts)} points with a leading mode, {tl:.1f} s", flush=True)
    s1 = int(a.steps * 0.5)
    sched = ((s1, 1e-3), (int(a.steps * 0.25), 3e-4), (a.steps - s1 - int(a.steps * 0.25), 1e-4))
    t0 = time.perf_counter()
    mlp, hist = fit(S, feats[ok], targets(feats[ok], om[ok]), width=a.width, depth=a.depth, seed=a.seed,
                    schedule=sched)
    tt = time.perf_counter() - t0
    print(f"# trained {a.steps} steps in {tt:.1f} s, loss {hist[:10].mean():.3e} -> {hist[-100:].mean():.3e}",
          flush=True)
    # held out: 128 continuous (M, Tw/Taw) draws off the training grid, 64 points each, at N = 128
    htab, ept, ef, _, _, _ = heldout_family(ny=a.ny_eval, workers=a.workers, slot0=len(tables))
    for k, tab in enumerate(htab):
        S.set_profile(len(tables) + k, tab[:, 0], tab[:, 1:6])
    ref = labels(S, ept)
    ev = evaluate(S, mlp, ept, ef, ref)
    ev.pop("result")
    print(f"# held-out: {json.dumps(ev)}", flush=True)
    meta = dict(train_points=a.train_points, labelled=int(ok.sum()), ny_train=a.ny_train, steps=a.steps, width=a.width,
                depth=a.depth, seed=a.seed, family=dict(machs=list(fam.machs), tws=list(fam.tws)),
                loss_final=float(hist[-100:].mean()),
                heldout=dict((k, v) for k, v in ev.items() if k != "stats"))
    mlp.save(a.out, meta)
    if a.summary:
        with open(a.summary, "w") as f:
            json.dump(dict(meta=meta, label_seconds=tl, train_seconds=tt, loss_hist=hist[::100].tolist()), f)
    print(f"# saved {a.out}", flush=True)


if __name__ == "__main__":
    main()
