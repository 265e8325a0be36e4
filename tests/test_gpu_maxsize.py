"""The largest shapes of the path against the oracle: a temporal point at
N = 256 (n = 1280, the eight-wave LU and FP32 refinement at that order) and a
spatial point at N = 128 on a Mach 2 compressible Blasius profile (a
linearisation of order 1280).  The oracle side is the dense shift-invert
eigenproblem of the oracle's own FP64 pencil: LAPACK's eigenpairs
(scipy.linalg.eig) of (A - sigma B)^-1 B, factorised by scipy.linalg.lu_factor,
mapped back by omega = sigma + 1/theta (the same eigenvectors; a few seconds at
order 1280, where QZ takes minutes).  Every GPU pair with an FP64 residual
below 1e-10 matches an oracle eigenvalue to 1e-8 and its eigenvector to 1e-6
after getevec normalisation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _shift_invert(A, B, sigma):
    import scipy.linalg as sl
    th, V = sl.eig(sl.lu_solve(sl.lu_factor(A - sigma * B), B))
    keep = np.abs(th) > 1e-300                 # theta = 0: infinite eigenvalues of the pencil
    return sigma + 1.0 / th[keep], V[:, keep]


def _text(itype, Ma, Re, ny, a_or_w, beta):
    txt = (f"1\n288.15\n{Ma!r} {Re!r} 0.72\n{ny} {2.0 / 46.0!r} 50\n1\n1\n0 0 0 0\n{itype}\n"
           f"{a_or_w!r} 0\n{beta!r} 0\n0\n")
    return txt + ("0\n" if itype == 2 else "")


def _check(res, w, V, nrows, nev):
    from oracle import stab_oracle as so
    nconv = 0
    for k in range(nev):
        if not res.resid[0, k] < 1e-10:
            continue
        ev = res.evals[0, k]
        q = int(np.argmin(np.abs(w - ev)))
        _say(f"  pair {k}: gpu {ev:.12f} resid {res.resid[0, k]:.1e} oracle {w[q]:.12f}")
        assert abs(w[q] - ev) <= 1e-8 * abs(w[q])
        a = so.normalize_evec(V[:, q], nrows)
        b = so.normalize_evec(res.evecs[0][:, k], nrows)
        assert np.max(np.abs(a - b)) <= 1e-6
        nconv += 1
    return nconv


@pytest.mark.parametrize("mode", [0, 1])
def test_temporal_n256(gpu_solver, mode):
    import bench
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    tab = bench.make_workload(1, 3)[0]
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    a, b, re, ny = 0.2, 0.1, 2000.0, 256
    txt = _text(1, 0.5, re, ny, a, b)
    sigma = complex(0.33 * a, -0.005)
    pt = sv.make_point(sv.read_input(txt), sigma, 4)
    _say(f"temporal N={ny} mode {mode}")
    res = gpu_solver.solve([pt], mode)
    _say(f"  info {res.info[0]} evals {res.evals[0, :4]}")
    pb = so.setup(so.parse_input(txt), tab[:, 0], tab[:, 1:6].copy(), spatial=False)
    A0, B0 = so.temporal_pencil(pb)
    w, V = _shift_invert(A0, B0, sigma)
    assert _check(res, w, V, 5 * ny, 4) >= 3


@pytest.mark.parametrize("mode", [0, 1])
def test_spatial_n128_mach2(gpu_solver, mode):
    from oracle import stab_oracle as so
    from stab_amd import meanflow as mf
    from stab_amd import solver as sv
    tab = mf.profile(mf.BlasiusSpec(Ma=2.0, Pr=0.72, law="sutherland", Te=288.15, eta_max=20.0, nrows=2000))
    gpu_solver.set_profile(0, tab[:, 0], tab[:, 1:6])
    om, b, re, ny = 0.08, 0.0, 1500.0, 128
    txt = _text(2, 2.0, re, ny, om, b)
    sigma = complex(om / 0.6, -0.002)
    pt = sv.make_point(sv.read_input(txt), sigma, 4)
    _say(f"spatial N={ny} M=2 mode {mode}")
    res = gpu_solver.solve([pt], mode)
    _say(f"  info {res.info[0]} evals {res.evals[0, :4]}")
    pb = so.setup(so.parse_input(txt), tab[:, 0], tab[:, 1:6].copy(), spatial=True)
    C0, C1, C2 = so.spatial_pencil(pb)
    Aa, Bb = so.spatial_linearization(C0, C1, C2)
    w, V = _shift_invert(Aa, Bb, sigma)
    assert _check(res, w, V, 5 * ny, 4) >= 3
