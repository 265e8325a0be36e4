"""GPU pencils on the branches the benchmark and the sweeps use (Sutherland,
Pr = 0.72, beta != 0, a crossflow component w != 0, wallt = 2, top = 0/1,
Mach 0.5 and 2), entry by entry against the oracle.  The oracle itself is
pinned on these branches by tests/test_oracle_pinning.py (an independent
linearisation of the Navier-Stokes equations, rotation invariance and the
temporal <-> spatial round trip).

Tolerance: every row of every matrix to 1e-12 of that row's largest entry
(a row-wise bar, so small viscous terms are not hidden behind large ones).
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TE = 288.15


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def flows():
    from stab_amd import meanflow as mf
    out = {}
    for Ma in (0.5, 2.0):
        tab = mf.profile(mf.BlasiusSpec(Ma=Ma, Pr=0.72, law="sutherland", Te=TE, eta_max=20.0, nrows=1200))
        ym, vt = tab[:, 0].copy(), tab[:, 1:6].copy()
        vtc = vt.copy()
        vtc[:, 3] = 0.4 * vt[:, 1] + 0.2 * vt[:, 1] * np.exp(-0.3 * ym)
        out[(Ma, False)] = (ym, vt)
        out[(Ma, True)] = (ym, vtc)
    return out


def _text(itype, ny, Ma, Re, top, wallt, a, b, om):
    head = f"1\n{TE}\n{Ma} {Re} 0.72\n{ny} {2.0 / 46.0!r} 50\n1\n1\n{top} 0 {wallt} 0\n{itype}\n"
    if itype == 1:
        return head + f"{a.real!r} {a.imag!r}\n{b.real!r} {b.imag!r}\n0\n"
    return head + f"{om.real!r} {om.imag!r}\n{b.real!r} {b.imag!r}\n0\n0\n"


CASES = [
    # itype, ny, Ma, crossflow, top, wallt, Re, alpha, beta, omega
    (1, 64, 0.5, False, 0, 0, 1000.0, 0.1, 0.0, 0.0),          # the Blasius point of the CPU-runnable case
    (1, 96, 0.5, True, 0, 0, 1893.0, 0.287, 0.047, 0.0),
    (1, 48, 2.0, True, 0, 2, 1500.0, 0.21 + 0.01j, 0.17, 0.0),
    (2, 64, 2.0, False, 1, 0, 1200.0, 0.0, 0.0, 0.08),
    (2, 48, 0.5, True, 0, 2, 800.0, 0.0, 0.12 - 0.003j, 0.05 + 0.001j),
    (2, 40, 2.0, True, 1, 2, 2000.0, 0.0, 0.2, 0.11),
]


@pytest.mark.parametrize("case", CASES, ids=[f"it{c[0]}_N{c[1]}_M{c[2]}_cf{int(c[3])}_top{c[4]}_wt{c[5]}"
                                             for c in CASES])
def test_pencil_branches_match_oracle(gpu_solver, flows, case):
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    from stab_amd import _lib
    itype, ny, Ma, cf, top, wallt, Re, a, b, om = case
    _say(f"pencil branches: {case}")
    ym, vt = flows[(Ma, cf)]
    txt = _text(itype, ny, Ma, Re, top, wallt, complex(a), complex(b), complex(om))
    gpu_solver.set_profile(0, ym, vt)
    pt = sv.make_point(sv.read_input(txt), 0.1 + 0j, 4)
    p = so.parse_input(txt)
    if itype == 1:
        pb = so.setup(p, ym, vt, spatial=False)
        mats = dict(zip((_lib.MAT_A0, _lib.MAT_B0), so.temporal_pencil(pb)))
    else:
        pb = so.setup(p, ym, vt, spatial=True)
        mats = dict(zip((_lib.MAT_C0, _lib.MAT_C1, _lib.MAT_C2), so.spatial_pencil(pb)))
    for which, ref in mats.items():
        got = gpu_solver.pencil(pt, which)
        rowmax = np.abs(ref).max(axis=1)
        live = rowmax > 0
        err = np.max(np.abs(got - ref)[live] / rowmax[live, None])
        assert np.all(got[~live] == 0)
        _say(f"  matrix {which}: max row-relative err {err:.2e}")
        assert err < 1e-12


@pytest.mark.parametrize("field", ["Re", "Ma", "alpha", "shift"])
def test_nonfinite_point_field_is_rejected(gpu_solver, flows, field):
    """A point with a NaN or infinite field is refused with STAB_E_ARG before
    any kernel runs (no normalisation ever sees an all-NaN vector from it)."""
    from stab_amd import solver as sv
    from stab_amd import _lib
    _say(f"argument check: non-finite {field}")
    ym, vt = flows[(0.5, False)]
    gpu_solver.set_profile(0, ym, vt)
    txt = _text(1, 32, 0.5, 1000.0, 0, 0, 0.1 + 0j, 0j, 0j)
    pt = sv.make_point(sv.read_input(txt), 0.03 + 0j, 4)
    if field in ("Re", "Ma"):
        setattr(pt, field, float("nan"))
    else:
        setattr(pt, field, _lib.Z(float("inf"), 0.0))
    with pytest.raises(sv.StabError, match="not finite"):
        gpu_solver.solve([pt], _lib.MODE_PLAIN_F64)
    with pytest.raises(sv.StabError, match="not finite"):
        gpu_solver.solve_leading([pt], _lib.MODE_MIXED, _lib.LEAD_INVIT)
    # the context stays usable
    ok = sv.make_point(sv.read_input(txt), 0.03 + 0j, 2)
    res = gpu_solver.solve([ok], _lib.MODE_PLAIN_F64, want_vectors=True)
    assert res.info[0] == _lib.INFO_OK


def test_unused_eigenvector_columns_are_zero(gpu_solver, flows):
    """stab_solve returns zeros in the columns past a point's nev, also after
    an earlier call filled the pooled device buffer with more columns."""
    import ctypes as C
    from stab_amd import solver as sv
    from stab_amd import _lib
    ym, vt = flows[(0.5, False)]
    gpu_solver.set_profile(0, ym, vt)
    txt = _text(1, 32, 0.5, 1000.0, 0, 0, 0.1 + 0j, 0j, 0j)
    gpu_solver.solve([sv.make_point(sv.read_input(txt), 0.03 + 0j, 8)], _lib.MODE_PLAIN_F64)
    pts = (_lib.Point * 1)(sv.make_point(sv.read_input(txt), 0.03 + 0j, 2))
    lib = gpu_solver.lib
    nbytes = lib.stab_evec_bytes(1, pts)
    buf = np.full(nbytes // 16, 7.0 + 7.0j)
    ev = np.zeros(_lib.STAB_MAX_NEV, dtype=np.complex128)
    rs = np.zeros(_lib.STAB_MAX_NEV)
    info = np.zeros(1, dtype=np.int32)
    zp = C.POINTER(_lib.Z)
    gpu_solver._check(lib.stab_solve(gpu_solver.ctx, 1, pts, _lib.MODE_PLAIN_F64, ev.ctypes.data_as(zp),
                                     buf.ctypes.data_as(zp), rs.ctypes.data_as(C.POINTER(C.c_double)),
                                     info.ctypes.data_as(C.POINTER(C.c_int32)), None))
    nv = 5 * 32
    blk = buf.reshape(_lib.STAB_MAX_NEV, nv).T
    assert info[0] == _lib.INFO_OK
    assert np.all(np.abs(blk[:, :2]).max(axis=0) > 0)
    assert np.all(blk[:, 2:] == 0)
    assert np.all(ev[2:] == 0) and np.all(rs[2:] == 0)
