"""Leading eigenmode per point through the GPU path (stab_solve_leading):
Krylov-Schur selection against QZ with the oracle's leading-mode rule,
residual inverse iteration from perturbed shifts (FP64 and FP32 factors), the
FP64 fallback, and the thesis goldens reached by inverse iteration."""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

NPTS = 12


def _say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def sweep(gpu_solver):
    import bench
    tab, al, be, re, ny = bench.make_workload(NPTS, 77)
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    pts = bench.make_points(al, be, re, ny, 8)
    return tab, al, be, re, ny, pts


@pytest.fixture(scope="module")
def krylov(gpu_solver, sweep):
    from stab_amd import _lib
    pts = sweep[5]
    return gpu_solver.solve_leading(pts, _lib.MODE_PLAIN_F64, _lib.LEAD_KRYLOV, want_vectors=True)


def _oracle_pencil(tab, a, b, r, n):
    from oracle import stab_oracle as so
    txt = (f"1\n288.15\n0.5 {float(r)!r} 0.72\n{int(n)} {2.0 / 46.0!r} 50\n1\n1\n0 0 0 0\n1\n"
           f"{float(a)!r} 0\n{float(b)!r} 0\n0\n")
    p = so.parse_input(txt)
    pb = so.setup(p, tab[:, 0], tab[:, 1:6].copy(), spatial=False)
    return so.temporal_pencil(pb)


def test_leading_krylov_matches_qz(gpu_solver, sweep, krylov):
    """For each point, the leading discrete pair among the 8 nearest the
    input shift equals the one the oracle's rule picks among the 8 QZ
    eigenvalues nearest the same shift (1e-8), eigenvector to 1e-6."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    tab, al, be, re, ny, pts = sweep
    _say(f"leading krylov: info {krylov['info']} stats {krylov['stats']['nconv']}")
    nfound = 0
    for i in range(NPTS):
        A0, B0 = _oracle_pencil(tab, al[i], be[i], re[i], ny[i])
        w, V = so.qz(A0, B0)
        ok = np.where(np.isfinite(w) & (np.abs(w) > 1e-9))[0]
        sig = complex(pts[i].shift.re, pts[i].shift.im)
        near = ok[np.argsort(np.abs(w[ok] - sig))[:8]]
        jl = so.leading_mode(1, w, V, int(ny[i]), alpha_r=float(al[i]), candidates=list(near))
        ev = krylov["evals"][i]
        _say(f"  pt {i} N={ny[i]} info {krylov['info'][i]} gpu {ev:.10f} oracle {w[jl] if jl >= 0 else None}")
        if jl < 0:
            assert krylov["info"][i] == _lib.INFO_NOCONV
            continue
        assert krylov["info"][i] == _lib.INFO_OK
        assert abs(ev - w[jl]) <= 1e-8 * abs(w[jl])
        assert krylov["resid"][i] < 1e-10
        x = so.normalize_evec(V[:, jl], 5 * int(ny[i]))
        assert np.max(np.abs(krylov["evecs"][i] - x)) <= 1e-6
        nfound += 1
    assert nfound == NPTS          # every point of this seeded sweep has a leading discrete mode


@pytest.mark.parametrize("mode", [0, 1])
def test_leading_invit_from_perturbed_shift(gpu_solver, sweep, krylov, mode):
    """Residual inverse iteration from shifts 1e-3 (relative) off the leading
    eigenvalue reaches it to 1e-8 with an FP64 residual below 1e-10, with FP64
    (mode 0) and equilibrated FP32 (mode 1) factors, without the fallback."""
    from stab_amd import _lib
    from stab_amd import solver as sv
    pts = sweep[5]
    keep = [i for i in range(NPTS) if krylov["info"][i] == _lib.INFO_OK]
    sub = []
    for i in keep:
        p = sv.clone_point(pts[i])
        ev = krylov["evals"][i]
        s = ev * (1.0 + 1e-3 * (0.6 + 0.8j))
        p.shift.re, p.shift.im = s.real, s.imag
        sub.append(p)
    fb = [complex(pts[i].shift.re, pts[i].shift.im) for i in keep]
    res = gpu_solver.solve_leading(sub, mode, _lib.LEAD_INVIT, maxit=30, fallback_shifts=fb, want_vectors=True)
    _say(f"invit mode {mode}: iters {res['iters']} info {res['info']} nfallback {res['stats']['nfallback']}")
    from oracle import stab_oracle as so
    tab, al, be, re, ny, _ = sweep
    for k, i in enumerate(keep):
        A0, B0 = _oracle_pencil(tab, al[i], be[i], re[i], ny[i])
        w, V = so.qz(A0, B0)
        w = np.where(np.isfinite(w), w, np.inf)
        ev = res["evals"][k]
        j = int(np.argmin(np.abs(w - ev)))
        _say(f"  pt {i}: invit {ev:.12f} resid {res['resid'][k]:.1e} | krylov {krylov['evals'][i]:.12f} "
             f"resid {krylov['resid'][i]:.1e} | qz {w[j]:.12f}")
        assert res["info"][k] == _lib.INFO_OK, (k, res["info"][k], res["resid"][k])
        assert abs(ev - w[j]) <= 1e-8 * abs(w[j])
        assert abs(ev - krylov["evals"][i]) <= 1e-7 * abs(w[j])      # the same mode as the Krylov path
        assert res["resid"][k] < 1e-10
        x = so.normalize_evec(V[:, j], 5 * int(ny[i]))
        assert np.max(np.abs(res["evecs"][k] - x)) <= 1e-6
    assert res["stats"]["nfallback"] == 0


def test_leading_second_pass_on_sweep(gpu_solver):
    """The 4096-point sweep with learned shifts (the benchmark workload): no
    point needs the FP64 fallback, and the first six points that went through
    the two-vector second pass (more solves than the first pass allows) return
    the leading discrete mode the oracle's rule picks among the 8 QZ
    eigenvalues nearest sigma = 0.33 alpha (1e-8), with the eigenvector to
    1e-6."""
    import bench
    from oracle import stab_oracle as so
    from stab_amd import _lib, learned
    tab, al, be, re, ny = bench.make_workload(4096, 2024)
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    pts = bench.make_points(al, be, re, ny, 8)
    mlp = learned.ShiftMLP.load(gpu_solver)
    om = learned.shifts(mlp, bench.features(al, be, re))
    res = gpu_solver.solve_leading(pts, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=om,
                                   shifts=om, nev=4, want_vectors=True)
    second = np.where(res["iters"] > 21)[0]
    _say(f"sweep: nfallback {res['stats']['nfallback']}, second pass {len(second)} points")
    assert res["stats"]["nfallback"] == 0
    assert np.all(res["info"] == _lib.INFO_OK)
    assert len(second) >= 6
    for i in second[:6]:
        A0, B0 = _oracle_pencil(tab, al[i], be[i], re[i], ny[i])
        w, V = so.qz(A0, B0)
        ok = np.where(np.isfinite(w) & (np.abs(w) > 1e-9))[0]
        sig = complex(pts[i].shift.re, pts[i].shift.im)
        near = ok[np.argsort(np.abs(w[ok] - sig))[:8]]
        jl = so.leading_mode(1, w, V, int(ny[i]), alpha_r=float(al[i]), candidates=list(near))
        _say(f"  pt {i} N={ny[i]} solves {res['iters'][i]} gpu {res['evals'][i]:.10f} oracle {w[jl]:.10f}")
        assert abs(res["evals"][i] - w[jl]) <= 1e-8 * abs(w[jl])
        x = so.normalize_evec(V[:, jl], 5 * int(ny[i]))
        assert np.max(np.abs(res["evecs"][i] - x)) <= 1e-6


def test_leading_batch_order_does_not_matter(gpu_solver):
    """The LU and the inverse iteration map work-groups to points through a
    largest-matrix-first order, and the start vectors are seeded from a hash
    of each point's own fields, so a point's result depends on the point
    alone: 512 points of the benchmark sweep solved in their own order and in
    reverse return bitwise the same eigenvalues, residuals, solve counts,
    statuses and eigenvectors.  A result stored for the wrong point would
    differ at O(1)."""
    import bench
    from stab_amd import _lib, learned
    tab, al, be, re, ny = bench.make_workload(4096, 2024)
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    idx = np.arange(512)
    pts = bench.make_points(al[idx], be[idx], re[idx], ny[idx], 8)
    mlp = learned.ShiftMLP.load(gpu_solver)
    om = learned.shifts(mlp, bench.features(al[idx], be[idx], re[idx]))
    a = gpu_solver.solve_leading(pts, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=om,
                                 shifts=om, nev=4, want_vectors=True)
    rev = idx[::-1]
    b = gpu_solver.solve_leading([pts[i] for i in rev], _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20,
                                 fallback_shifts=om[rev], shifts=om[rev], nev=4, want_vectors=True)
    back = np.argsort(rev)
    same = {k: bool(np.array_equal(a[k], b[k][back])) for k in ("evals", "resid", "iters", "info")}
    vec = all(np.array_equal(a["evecs"][i], b["evecs"][back[i]]) for i in range(len(idx)))
    _say(f"batch order: equal {same}, eigenvectors equal {vec}, "
         f"max rel diff {np.max(np.abs(a['evals'] - b['evals'][back]) / np.abs(a['evals'])):.1e}")
    assert all(same.values()) and vec
    assert (a["info"] == _lib.INFO_OK).sum() >= 500


def test_leading_invit_falls_back_from_continuum_shift(gpu_solver, sweep, krylov):
    """A shift on the continuous spectrum (omega = alpha - 1e-5 i) lands on a
    non-decaying mode; the point is redone by FP64 Krylov-Schur from the
    fallback shift and reports STAB_INFO_FALLBACK with the Krylov answer."""
    from stab_amd import _lib
    from stab_amd import solver as sv
    pts = sweep[5]
    keep = [i for i in range(NPTS) if krylov["info"][i] == _lib.INFO_OK][:4]
    sub = []
    for i in keep:
        p = sv.clone_point(pts[i])
        p.shift.re, p.shift.im = p.alpha.re, -1e-5
        sub.append(p)
    fb = [complex(pts[i].shift.re, pts[i].shift.im) for i in keep]
    res = gpu_solver.solve_leading(sub, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, fallback_shifts=fb)
    _say(f"continuum shift: info {res['info']} ff {res['farfield']} nfallback {res['stats']['nfallback']}")
    for k, i in enumerate(keep):
        assert res["info"][k] == _lib.INFO_FALLBACK
        assert abs(res["evals"][k] - krylov["evals"][i]) <= 1e-10 * abs(krylov["evals"][i])
    assert res["stats"]["nfallback"] == len(keep)


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("case", ["temporal", "spatial"])
def test_leading_invit_thesis_goldens(gpu_solver, ts_profile, case, mode):
    """thesis/TStest time.ref / space.ref reached by residual inverse iteration
    from a shift 1e-3 off the golden eigenvalue: eigenvalue to 1e-8 and the
    getevec table to 1e-8 (ndiff -abserr semantics)."""
    from oracle import stab_oracle as so
    from stab_amd import _lib
    from stab_amd import solver as sv
    ref = "time.ref" if case == "temporal" else "space.ref"
    hdr, tab = so.parse_table(open(golden("thesis_TStest", ref)).read())
    gold = hdr["Omega"] if case == "temporal" else hdr["Alpha"]
    inp = sv.read_input(open(golden("thesis_TStest", f"{case}.inp")).read())
    gpu_solver.set_profile(0, *ts_profile)
    pt = sv.make_point(inp, gold * (1.0 + 1e-3 * (0.6 + 0.8j)), 4)
    res = gpu_solver.solve_leading([pt], mode, _lib.LEAD_INVIT, maxit=30, want_vectors=True)
    _say(f"thesis {case} mode {mode}: {res['evals'][0]} iters {res['iters'][0]} resid {res['resid'][0]:.2e} "
         f"ff {res['farfield'][0]:.2e} info {res['info'][0]}")
    assert res["info"][0] == _lib.INFO_OK
    assert abs(res["evals"][0] - gold) <= 1e-8 * abs(gold)
    vec = res["evecs"][0]
    g = gpu_solver.grid(pt)
    mine = np.array([[g["y"][i]] + sum(([vec[i * 5 + k].real, vec[i * 5 + k].imag] for k in range(5)), [])
                     for i in range(inp.ny - 1, -1, -1)])
    d = np.abs(mine - tab)
    d[np.isclose(tab, 1e99)] = 0.0
    d[-1, 0] = 0.0
    assert d.max() <= 1e-8
    if case == "spatial":
        n = 5 * inp.ny
        assert np.linalg.norm(vec[n:] - res["evals"][0] * vec[:n]) <= 1e-10 * np.linalg.norm(vec[n:])


def test_leading_fused_second_pass_matches_second_launch(gpu_solver):
    """The second pass run inside the first-pass work-group as soon as a
    point's first pass ends (the default) and run as a launch of its own after
    the whole first pass (STAB_FUSE2=0) apply the same rule to the same
    first-pass results and the same two-vector iteration: on 1024 points of
    the benchmark sweep (some of which take the second pass) they agree on
    every status and solve count, eigenvalues to 1e-12 and eigenvectors to
    1e-10."""
    import os
    import bench
    from stab_amd import _lib, learned
    tab, al, be, re, ny = bench.make_workload(4096, 2024)
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    idx = np.arange(1024)
    pts = bench.make_points(al[idx], be[idx], re[idx], ny[idx], 8)
    mlp = learned.ShiftMLP.load(gpu_solver)
    om = learned.shifts(mlp, bench.features(al[idx], be[idx], re[idx]))
    out = {}
    try:
        for fuse in ("1", "0"):
            os.environ["STAB_FUSE2"] = fuse
            out[fuse] = gpu_solver.solve_leading(pts, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20,
                                                 fallback_shifts=om, shifts=om, nev=4, want_vectors=True)
    finally:
        os.environ.pop("STAB_FUSE2", None)
    a, b = out["1"], out["0"]
    second = int((a["iters"] > 21).sum())
    rel = np.abs(a["evals"] - b["evals"]) / np.abs(b["evals"])
    dv = max(float(np.max(np.abs(u - v))) for u, v in zip(a["evecs"], b["evecs"]))
    bit = bool(np.array_equal(a["evals"], b["evals"]))
    _say(f"fused second pass: {second} points took it; bitwise equal eigenvalues {bit}; "
         f"max rel diff {rel.max():.1e}, eigenvector diff {dv:.1e}")
    assert second >= 5
    assert np.array_equal(a["info"], b["info"]) and np.array_equal(a["iters"], b["iters"])
    assert rel.max() <= 1e-12 and dv <= 1e-10
    assert (a["info"] == _lib.INFO_OK).sum() >= 1000
