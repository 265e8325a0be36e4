"""The leading-mode path's first pass as a captured HIP graph, and the LU
launched per run of points that share a work-group shape (two streams).

A graph is captured on the first call of a batch layout and replayed on the
next calls with that layout; the point descriptors (shifts included) are
copied before each launch, so a replay with new shifts must give what a direct
launch gives.  Each replay is checked against a direct run (STAB_GRAPH=0), with
other direct work on the same context between the replays.
"""
import os
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
    os.environ.pop("STAB_GRAPH", None)


def _points(n, seed, scale=1.0):
    from stab_amd import solver as sv
    rng = np.random.default_rng(seed)
    pts = []
    for k in range(n):
        a, b, re = rng.uniform(0.05, 0.4), rng.uniform(0.0, 0.2), rng.uniform(300.0, 3000.0)
        ny = (64, 96, 128)[k % 3]
        inp = sv.Inputs(mattyp=1, T0=TE, Te=TE, Ma=0.5, Re=float(re), Pr=0.72, ny=ny, yi=2.0 / 46.0, ymax=50.0,
                        itype=1, alpha=complex(a, 0.0), beta=complex(b, 0.0))
        pts.append(sv.make_point(inp, complex(0.33 * a * scale, 0.0), 4, 0))
    return pts


def _run(ctx, pts, graph):
    from stab_amd import _lib
    os.environ["STAB_GRAPH"] = "1" if graph else "0"
    return ctx.solve_leading(pts, _lib.MODE_MIXED, _lib.LEAD_INVIT, maxit=20, want_vectors=True)


def _same(a, b):
    for k in ("evals", "resid", "iters", "info", "farfield"):
        assert np.array_equal(a[k], b[k]), k
    for u, v in zip(a["evecs"], b["evecs"]):
        assert np.array_equal(u, v)


def test_graph_replays_match_direct_runs(ctx):
    pts = _points(48, 3)
    _say("graph: direct run of 48 points")
    ref = _run(ctx, pts, False)
    assert ref["stats"]["graph"] == 0
    assert ref["stats"]["lu_launches"] >= 1
    _say(f"   LU launches {ref['stats']['lu_launches']}")
    first = _run(ctx, pts, True)
    _say(f"   first graph call: graph={first['stats']['graph']}")
    assert first["stats"]["graph"] in (1, 2)
    _same(first, ref)
    for rep in range(3):
        _run(ctx, _points(5, 100 + rep), False)            # other direct work on the context in between
        again = _run(ctx, pts, True)
        _say(f"   replay {rep}: graph={again['stats']['graph']}")
        assert again["stats"]["graph"] == 2
        _same(again, ref)
    # same layout, new shifts: the replay reads the freshly copied point descriptors
    moved = _points(48, 3, scale=1.03)
    direct = _run(ctx, moved, False)
    replay = _run(ctx, moved, True)
    assert replay["stats"]["graph"] == 2
    _same(replay, direct)
    assert not np.array_equal(direct["iters"], ref["iters"]) or not np.array_equal(direct["evals"], ref["evals"])


def test_lu_segments_match_single_launch(ctx):
    """The segmented two-stream LU gives the factors of one launch: the
    leading modes equal those of a run with every point in one segment shape."""
    pts = _points(30, 9)
    os.environ["STAB_LU_SEGMENTS"] = "0"                    # one launch shaped by the widest matrix
    try:
        one = _run(ctx, pts, False)
    finally:
        os.environ.pop("STAB_LU_SEGMENTS", None)
    seg = _run(ctx, pts, False)
    _say(f"lu segments: {seg['stats']['lu_launches']} launches vs {one['stats']['lu_launches']}")
    assert one["stats"]["lu_launches"] == 1
    from stab_amd import _lib
    ok = (one["info"] == _lib.INFO_OK) & (seg["info"] == _lib.INFO_OK)
    assert ok.all()
    assert np.max(np.abs(one["evals"] - seg["evals"]) / np.abs(seg["evals"])) <= 1e-9
    assert np.all(seg["resid"] < 1e-13)


def test_lu_streams_give_bitwise_equal_results(ctx):
    """The LU's segments split further and spread over 1, 2 or 4 streams
    (STAB_LU_STREAMS): every point's factorisation is the same arithmetic in
    any grouping, so the leading modes, vectors, residuals and solve counts
    are bitwise equal."""
    pts = _points(60, 21)
    runs = {}
    try:
        for ns in (1, 2, 4):
            os.environ["STAB_LU_STREAMS"] = str(ns)
            runs[ns] = _run(ctx, pts, False)
            _say(f"lu streams {ns}: {runs[ns]['stats']['lu_launches']} segments")
    finally:
        os.environ.pop("STAB_LU_STREAMS", None)
    assert runs[4]["stats"]["lu_launches"] >= runs[1]["stats"]["lu_launches"]
    for ns in (2, 4):
        _same(runs[ns], runs[1])
