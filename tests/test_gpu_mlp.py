"""Shift-predictor MLP on the GPU (stab_mlp_*): loss, gradient, Adam step and
prediction against the CPU oracle (oracle/mlp_oracle.py, BF16 operand
rounding mirrored), a training run, and what the trained network gains over
the leading-mode solve without it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def net(gpu_solver):
    from stab_amd import learned
    m = learned.ShiftMLP(gpu_solver, width=64, depth=3, seed=5)
    yield m
    m.close()


def _identity_scaling(m):
    m.set_scaling(np.zeros(5), np.ones(5), np.zeros(2), np.ones(2))


def test_loss_and_gradient_match_oracle(net):
    from oracle import mlp_oracle as mo
    lay = mo.Layout(5, 64, 3, 2)
    assert net.np == lay.nparams
    rng = np.random.default_rng(3)
    p = lay.random_params(rng, 0.4).astype(np.float32)
    net.set_params(p)
    _identity_scaling(net)
    x = rng.uniform(-1, 1, (64, 5)).astype(np.float32)
    t = rng.standard_normal((64, 2)).astype(np.float32)
    loss, g = net.loss_grad(x, t)
    lo, go = mo.loss_grad(lay, p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), rnd=True)
    lf, gf = mo.loss_grad(lay, p.astype(np.float64), x.astype(np.float64), t.astype(np.float64), rnd=False)
    e_rnd = np.linalg.norm(g - go) / np.linalg.norm(go)
    e_f64 = np.linalg.norm(g - gf) / np.linalg.norm(gf)
    _say(f"loss gpu {loss:.6e} oracle(bf16) {lo:.6e} oracle(f64) {lf:.6e}; grad rel err vs bf16 oracle {e_rnd:.2e}, "
         f"vs f64 {e_f64:.2e}")
    assert abs(loss - lo) <= 1e-5 * lo
    assert e_rnd <= 2e-3
    assert e_f64 <= 3e-2          # BF16 operands: percent-level agreement with exact arithmetic
    # padded weights get no gradient
    off, o, i, ob = lay.layers()[0]
    assert np.all(g[off:off + o * i].reshape(o, i)[:, 5:] == 0)


def test_adam_step_matches_oracle(net):
    from oracle import mlp_oracle as mo
    lay = mo.Layout(5, 64, 3, 2)
    rng = np.random.default_rng(4)
    p0 = lay.random_params(rng, 0.4).astype(np.float32)
    net.set_params(p0)
    _identity_scaling(net)
    x = rng.uniform(-1, 1, (32, 5)).astype(np.float32)
    t = rng.standard_normal((32, 2)).astype(np.float32)
    _, g = net.loss_grad(x, t)
    lr = 1e-3
    net.train(x, t, batch=32, steps=1, lr=lr)         # batch == rows: the whole set, in order
    p1 = net.get_params()
    po, _, _ = mo.adam_step(p0.astype(np.float64), g.astype(np.float64), 0.0, 0.0, 1, lr)
    err = np.max(np.abs(p1 - po))
    _say(f"adam step max abs diff {err:.2e} (lr {lr})")
    assert err <= 1e-3 * lr


def test_predict_matches_oracle_with_scaling(net):
    from oracle import mlp_oracle as mo
    lay = mo.Layout(5, 64, 3, 2)
    rng = np.random.default_rng(5)
    p = lay.random_params(rng, 0.4).astype(np.float32)
    net.set_params(p)
    ish = np.array([0.2, 0.1, 7.0, 1.3, 1.0], np.float32)
    isc = np.array([5.7, 10.0, 0.9, 0.8, 2.0], np.float32)
    osh = np.array([0.35, -0.01], np.float32)
    osc = np.array([12.0, 150.0], np.float32)
    net.set_scaling(ish, isc, osh, osc)
    x = (ish + rng.uniform(-1, 1, (100, 5)) / isc).astype(np.float32)     # 100 rows: not a multiple of 16
    y = net.predict(x)
    _, yo = mo.forward(lay, p.astype(np.float64), (x.astype(np.float64) - ish) * isc, rnd=True)
    yo = yo[:, :2] / osc + osh
    err = np.max(np.abs(y - yo) * osc)
    _say(f"predict max scaled err {err:.2e}")
    assert err <= 1e-4


def test_training_reduces_loss(gpu_solver):
    from stab_amd import learned
    rng = np.random.default_rng(6)
    x = rng.uniform(-1, 1, (4096, 5)).astype(np.float32)
    t = np.stack([np.sin(2 * x[:, 0]) + x[:, 1] ** 2, x[:, 2] * x[:, 3] - 0.5 * x[:, 4]], 1).astype(np.float32)
    m = learned.ShiftMLP(gpu_solver, width=64, depth=3, seed=9)
    m.set_scaling(np.zeros(5), np.ones(5), np.zeros(2), np.ones(2))
    h = m.train(x, t, batch=512, steps=1500, lr=3e-3)
    _say(f"synthetic fit: loss {h[:20].mean():.3e} -> {h[-100:].mean():.3e}")
    assert h[-100:].mean() <= 1e-2 * h[:20].mean()
    m.close()


def test_learned_shift_gains_over_the_input_shift(gpu_solver):
    """Train on solves of a small family, then solve held-out points for their
    leading mode three ways: residual inverse iteration from the network's
    shift, the same from the input-file shift (the method without the
    network), and plain FP64 Krylov-Schur (the reference answer).  With the
    network every point returns the Krylov eigenvalue, with far fewer
    fallbacks and solves; from the input shift, inverse iteration can also
    settle on another discrete mode nearer that shift, so it never matches
    more points."""
    from stab_amd import _lib
    from stab_amd import learned
    fam = learned.Family(machs=(0.3, 0.7, 1.1), tws=(0.8, 1.0, 1.2), nrows=600)
    tables = learned.build_profiles(fam, workers=4)
    learned.load_family(gpu_solver, fam, tables)
    pts, feats = learned.sample_points(fam, 3072, seed=21, ny=64)
    om = learned.labels(gpu_solver, pts)
    ok = np.isfinite(om)
    _say(f"gain test: {int(ok.sum())}/{len(pts)} labelled")
    mlp, hist = learned.fit(gpu_solver, feats[ok], learned.targets(feats[ok], om[ok]), width=128, depth=3, seed=3,
                            schedule=((1500, 1e-3), (1000, 3e-4)), batch=512)
    _say(f"  loss {hist[:20].mean():.3e} -> {hist[-100:].mean():.3e}")
    ept, ef = learned.sample_points(fam, 512, seed=22, ny=64)
    ref = learned.labels(gpu_solver, ept)
    keep = np.isfinite(ref)
    ept = [p for p, k in zip(ept, keep) if k]
    ef, ref = ef[keep], ref[keep]
    fb = [complex(p.shift.re, p.shift.im) for p in ept]
    ev = learned.evaluate(gpu_solver, mlp, ept, ef, ref)
    base = gpu_solver.solve_leading(ept, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=fb)
    nb = int((base["info"] == _lib.INFO_FALLBACK).sum())
    nm = ev["info"][_lib.INFO_FALLBACK]
    mb = int((np.abs(base["evals"] - ref) <= 1e-8 * np.abs(ref)).sum())
    _say(f"  held-out {len(ept)}: learned shift: fallbacks {nm}, solves/point {ev['stats']['op_applies'] / len(ept):.1f}, "
         f"match {ev['match']}; input shift: fallbacks {nb}, solves/point {base['stats']['op_applies'] / len(ept):.1f}, "
         f"match {mb}; shift rel err median {ev['shift_rel_err_median']:.2e} p95 {ev['shift_rel_err_p95']:.2e}")
    assert ev["match"] == len(ept)
    assert mb <= ev["match"]
    assert nm <= 0.05 * len(ept)
    assert nb >= 4 * max(nm, 1)
    assert ev["stats"]["op_applies"] < 0.5 * base["stats"]["op_applies"]
    mlp.close()
