"""stab_solve_leading_noalloc against stab_solve_leading through the C ABI.

The allocation-free entry point carves every device buffer from the library's
static scratch array; it runs the same kernels on the same inputs, so its
results must equal the allocating entry point's bit for bit.  Two calls in a
row, the second with fewer and smaller points (the array keeps the first
call's contents), and a batch above the scratch size that must be refused
before any work.
"""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TE = 288.15


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    from stab_amd import meanflow as mf
    from stab_amd.solver import Solver
    s = Solver(0)
    tab = mf.profile(mf.BlasiusSpec(Ma=0.5, Pr=0.72, law="sutherland", Te=TE, eta_max=20.0, nrows=1200))
    s.set_profile(0, tab[:, 0], tab[:, 1:6])
    yield s
    s.close()


def _points(n, seed, nys):
    from stab_amd import solver as sv
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(n):
        a, b, re = rng.uniform(0.05, 0.4), rng.uniform(0.0, 0.2), rng.uniform(300.0, 3000.0)
        ny = int(nys[k % len(nys)])
        inp = sv.Inputs(mattyp=1, T0=TE, Te=TE, Ma=0.5, Re=float(re), Pr=0.72, ny=ny, yi=2.0 / 46.0,
                        ymax=50.0, itype=1, alpha=complex(a, 0.0), beta=complex(b, 0.0))
        pts.append(sv.make_point(inp, complex(0.33 * a, 0.0), 4, 0))
    return pts


def _call(s, fn, pts, mode, method, maxit=20):
    from stab_amd import _lib
    n = len(pts)
    arr = (_lib.Point * n)(*pts)
    ev = np.zeros(n, dtype=np.complex128)
    nv = [5 * p.ny for p in pts]
    vec = np.zeros(sum(nv), dtype=np.complex128)
    rs, ff = np.zeros(n), np.zeros(n)
    it, info = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)
    zp, ip, dp = C.POINTER(_lib.Z), C.POINTER(C.c_int32), C.POINTER(C.c_double)
    rc = fn(s.ctx, n, arr, mode, method, maxit, None, ev.ctypes.data_as(zp), vec.ctypes.data_as(zp),
            rs.ctypes.data_as(dp), ff.ctypes.data_as(dp), it.ctypes.data_as(ip), info.ctypes.data_as(ip), None)
    return rc, dict(ev=ev, vec=vec, rs=rs, ff=ff, it=it, info=info)


@pytest.mark.parametrize("mode,method", [(1, 1), (0, 0)], ids=["mixed_invit", "plain_krylov"])
def test_noalloc_matches_allocating_entry(ctx, mode, method):
    from stab_amd import _lib
    lib = ctx.lib
    for call, (n, seed, nys) in enumerate([(6, 11, (64, 96, 128)), (3, 12, (48, 64))]):
        pts = _points(n, seed, nys)
        _say(f"no-alloc call {call}: {n} points, N {sorted(set(p.ny for p in pts))}, mode {mode}, method {method}")
        rc0, ref = _call(ctx, lib.stab_solve_leading, pts, mode, method)
        assert rc0 == 0, lib.stab_last_error(ctx.ctx)
        rc1, got = _call(ctx, lib.stab_solve_leading_noalloc, pts, mode, method)
        assert rc1 == 0, lib.stab_last_error(ctx.ctx)
        assert np.all(ref["info"] != _lib.INFO_NOCONV)
        for k in ref:
            assert np.array_equal(got[k], ref[k]), k


def test_noalloc_refuses_batches_above_the_scratch(ctx):
    from stab_amd import _lib
    lib = ctx.lib
    pts = _points(40, 13, (128,))
    _say("no-alloc refusal: 40 points at N = 128")
    rc, _ = _call(ctx, lib.stab_solve_leading_noalloc, pts, _lib.MODE_MIXED, _lib.LEAD_INVIT)
    assert rc == -1
    assert b"scratch" in lib.stab_last_error(ctx.ctx)
    # and the context still solves afterwards, allocation-free
    rc, got = _call(ctx, lib.stab_solve_leading_noalloc, pts[:2], _lib.MODE_MIXED, _lib.LEAD_INVIT)
    assert rc == 0 and np.all(got["info"] != _lib.INFO_NOCONV)
