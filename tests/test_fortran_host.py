"""The Fortran host driver (stab_amd/fortran, ISO_C_BINDING onto the C ABI):
the reference's input stream in, evec.dat out, checked against the thesis
golden eigenfunction files the way thesis/TStest/run.sh does it."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import golden, ROOT

BIN = os.path.join(ROOT, "stab_amd", "stab_gpu")


def _have_gpu():
    return os.path.exists("/dev/kfd")


def _run(tmp_path, inp_text, extra=""):
    shutil.copy(golden("TStest", "profile.0"), tmp_path / "profile.0")
    r = subprocess.run([BIN], input=inp_text + extra, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    return r


@pytest.mark.skipif(not os.path.exists(BIN), reason="Fortran driver not built (no Fortran compiler)")
@pytest.mark.skipif(_have_gpu(), reason="checks the no-device path")
def test_driver_reports_missing_device(tmp_path):
    r = _run(tmp_path, open(golden("thesis_TStest", "temporal.inp")).read())
    assert r.returncode == 3
    assert "no usable HIP device" in r.stdout


@pytest.mark.skipif(not os.path.exists(BIN), reason="Fortran driver not built (no Fortran compiler)")
def test_driver_reports_missing_profile(tmp_path):
    r = subprocess.run([BIN], input=open(golden("thesis_TStest", "temporal.inp")).read(), capture_output=True,
                       text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 1
    assert "Error reading profile" in r.stdout


@pytest.mark.gpu
@pytest.mark.parametrize("case,ref,key", [("temporal", "time.ref", "Omega"), ("spatial", "space.ref", "Alpha")])
def test_driver_reproduces_thesis_goldens(tmp_path, case, ref, key):
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    print(f"fortran driver {case}", flush=True)
    hdr, tab = so.parse_table(open(golden("thesis_TStest", ref)).read())
    guess = hdr[key]
    r = _run(tmp_path, open(golden("thesis_TStest", f"{case}.inp")).read(), f"{guess.real!r} {guess.imag!r} 4\n")
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    d = sv.read_evec_dat(str(tmp_path / "evec.dat"))
    n = 5 * d["ny"]
    nmax = n if case == "temporal" else 2 * n
    assert d["evals"].size == nmax and d["evec"].shape == (nmax, nmax)   # reference record lengths
    j = sv.select_nearest(d["evals"], guess)
    assert abs(d["evals"][j] - guess) <= 1e-8 * abs(guess)
    text = sv.eigenfunction_text(d, j)
    _, mine = so.parse_table(text)
    diff = np.abs(mine - tab)
    diff[np.isclose(tab, 1e99)] = 0.0
    diff[-1, 0] = 0.0           # wall y: 0 against the reference's round-off
    assert diff.max() <= 1e-8


@pytest.mark.gpu
def test_driver_fd_temporal(tmp_path):
    """itype 3 (stab.f90:56-60): thesis/TStest/fd_temporal.inp through
    fd_temporal on the GPU; evec.dat carries itype 1 as the reference's does,
    the ogengrid grid (wall first) and the oracle's ZGEEV eigenvalue nearest
    the guess, and getevec's table lists the nodes from the top down."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    print("fortran driver fd_temporal", flush=True)
    txt = open(golden("thesis_TStest", "fd_temporal.inp")).read()
    guess = 1.1467e-1 + 2.35e-3j
    r = _run(tmp_path, txt, f"{guess.real!r} {guess.imag!r} 4\n")
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    d = sv.read_evec_dat(str(tmp_path / "evec.dat"))
    assert d["itype"] == 1 and d["ny"] == 128
    p = so.parse_input(txt)
    pb, _, _, w, _ = so.run_fd_temporal(p, *so.read_profile(golden("TStest", "profile.0")))
    assert np.allclose(d["y"], pb.y, rtol=1e-14, atol=0) and d["y"][0] == 0.0
    j = sv.select_nearest(d["evals"], guess)
    ref = w[np.argmin(np.abs(w - guess))]
    assert abs(d["evals"][j] - ref) <= 1e-9 * abs(ref)
    _, tab = so.parse_table(sv.eigenfunction_text(d, j))
    assert tab[0, 0] == pytest.approx(50.0, rel=1e-14) and tab[-1, 0] == 0.0


@pytest.mark.gpu
def test_driver_fd_spatial(tmp_path):
    """itype 4 (stab.f90:61-67): thesis/TStest/fd_spatial.inp through
    fd_spatial on the GPU; evec.dat carries itype 2 as the reference's does
    and the oracle's companion eigenvalue nearest the guess."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    print("fortran driver fd_spatial", flush=True)
    txt = open(golden("thesis_TStest", "fd_spatial.inp")).read().replace("128 1.0 50.0", "96 1.0 50.0")
    guess = 2.2805e-1 - 6.5e-3j
    r = _run(tmp_path, txt, f"{guess.real!r} {guess.imag!r} 1\n")
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    d = sv.read_evec_dat(str(tmp_path / "evec.dat"))
    assert d["itype"] == 2 and d["ny"] == 96 and d["evals"].size == 10 * 96
    p = so.parse_input(txt)
    _, _, alp, _ = so.run_fd_spatial(p, *so.read_profile(golden("TStest", "profile.0")))
    j = sv.select_nearest(d["evals"], guess)
    ref = alp[np.argmin(np.abs(alp - guess))]
    assert abs(d["evals"][j] - ref) <= 1e-9 * abs(ref)


@pytest.mark.gpu
def test_driver_multi_temporal_sweep(tmp_path):
    """itype 7 (mtemporal): a 3 x 2 (alpha, beta) sweep solved in one batched
    call writes eig.1 .. eig.6 (alpha outer, beta inner, mtemporal.f90:25-38)
    with the evec.dat records; each file's eigenvalue nearest the phase-speed
    guess equals the QZ eigenvalue nearest it for the same pencil."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    base = open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", "48 1.0 0.0")
    lines = base.splitlines()
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("1") and "Temporal" in l)
    sweep = lines[:k] + ["7\t\t! multiple temporal"] + ["0.28 0.34 0.02", "0.0 0.02 0.01", "0.372 0.006 4"]
    r = _run(tmp_path, "\n".join(sweep) + "\n")
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    ym, vt = so.read_profile(golden("TStest", "profile.0"))
    iver = 0
    for a in (0.28, 0.30, 0.32):
        for b in (0.0, 0.01):
            iver += 1
            d = sv.read_evec_dat(str(tmp_path / f"eig.{iver}"))
            assert d["itype"] == 1 and abs(d["alpha"] - a) < 1e-12 and abs(d["beta"] - b) < 1e-12
            assert d["evals"].size == 5 * 48 and d["evec"].shape == (240, 240)
            guess = a * (0.372 + 0.006j)
            j = sv.select_nearest(d["evals"], guess)
            p = so.parse_input(base)
            p.alpha, p.beta = complex(a, 0.0), complex(b, 0.0)
            A0, B0 = so.temporal_pencil(so.setup(p, ym, vt, spatial=False))
            w, _ = so.qz(A0, B0, right=False)
            w = w[np.isfinite(w)]
            wq = w[np.argmin(np.abs(w - guess))]
            print(f"  eig.{iver}: alpha {a} beta {b}: {d['evals'][j]:.10f} qz {wq:.10f}", flush=True)
            assert abs(d["evals"][j] - wq) <= 1e-8 * abs(wq)


GETEVEC = os.path.join(ROOT, "stab_amd", "getevec")


@pytest.mark.skipif(not os.path.exists(GETEVEC), reason="getevec not built (no Fortran compiler)")
@pytest.mark.parametrize("case,ref,key", [("temporal", "time.ref", "Omega"), ("spatial", "space.ref", "Alpha")])
def test_getevec_program_reproduces_thesis_tables(tmp_path, case, ref, key):
    """The Fortran getevec (getevec.f90:2-239 semantics) on an eigensystem file
    holding the oracle's whole reference-method spectrum (ZGESV+ZGEEV, sorted by
    Im as PIKSR2 does): -v with the thesis guess writes the thesis table
    (thesis/TStest/run.sh:13-17, ndiff -abserr 1e-8); the index of that pair
    selects the same table."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    ym, vt = so.read_profile(golden("TStest", "profile.0"))
    text = open(golden("thesis_TStest", f"{case}.inp")).read()
    p = so.parse_input(text)
    inp = sv.read_input(text)
    if case == "temporal":
        pb, A0, B0, w, v = so.run_temporal(p, ym, vt)
    else:
        pb, Cs, w, v = so.run_spatial(p, ym, vt)
    grid = dict(y=pb.y, eta=pb.eta, deta=pb.deta, d2eta=pb.d2eta)
    sv.write_evec_dat(str(tmp_path / "evec.dat"), inp, grid, w, v)
    hdr, tab = so.parse_table(open(golden("thesis_TStest", ref)).read())
    guess = hdr[key]
    r = subprocess.run([GETEVEC, "-v", "evec.dat"], input=f"{guess.real!r} {guess.imag!r}\n0 0\n",
                       capture_output=True, text=True, cwd=tmp_path, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    base = "time" if case == "temporal" else "space"
    h2, mine = so.parse_table(open(tmp_path / f"{base}.1").read())
    assert abs(h2[key] - guess) <= 1e-8 * abs(guess)
    diff = np.abs(mine - tab)
    diff[np.isclose(tab, 1e99)] = 0.0
    diff[-1, 0] = 0.0
    assert diff.max() <= 1e-8
    j = int(np.argmin(np.abs(w - guess))) + 1
    r = subprocess.run([GETEVEC, "-i", "evec.dat"], input=f"{j}\n0\n", capture_output=True, text=True,
                       cwd=tmp_path, timeout=60)
    assert r.returncode == 0
    _, again = so.parse_table(open(tmp_path / f"{base}.1").read())
    assert np.array_equal(again, mine)


EXAMPLE = os.path.join(ROOT, "stab_amd", "alpha_march")


@pytest.mark.gpu
def test_driver_multi_spatial_stations_fd(tmp_path):
    """itype 8 with dtype 1 (mspatial.f90:85-89 calls fd_spatial): the
    finite-difference spatial solve at each station's Yi = 2 delta on the
    ogengrid grid (ymax 50), one batched call; eig.<k> holds the alpha
    nearest omega / 0.35, equal to the oracle's companion eigenvalue."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    print("fortran driver: multiple spatial, finite differences", flush=True)
    for k in (1, 2):
        shutil.copy(golden("TStest", "profile.0"), tmp_path / f"profile.{k}")
    (tmp_path / "delta.dat").write_text("# x delta\n100.0 0.5\n200.0 0.6\n")
    base = open(golden("thesis_TStest", "spatial.inp")).read().replace("96 1.0 0.0", "48 1.0 50.0")
    lines = base.splitlines()
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("2") and "Spatial" in l)
    sweep = lines[:k] + ["8\t\t! multiple spatial", "0.08 0.08 0.01", "0.0 0.0 0.0", "1 2 1", "1",
                         "0.35 0.0 1"]
    r = subprocess.run([BIN], input="\n".join(sweep) + "\n", capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    ym, vt = so.read_profile(golden("TStest", "profile.0"))
    for st, (xs, dl) in enumerate(((100.0, 0.5), (200.0, 0.6)), start=1):
        d = sv.read_evec_dat(str(tmp_path / f"eig.{st}"))
        assert d["itype"] == 2 and d["ind"] == st and abs(d["yi"] - 2 * dl) < 1e-12
        assert abs(d["y"][0]) == 0.0 and abs(d["y"][-1] - 50.0) < 1e-12     # ogengrid, wall first
        ev = d["evals"][np.abs(d["evals"].real) < 1e299]
        assert ev.size == 1
        p = so.parse_input(base)
        p.itype, p.omega, p.yi = 4, complex(0.08, 0.0), 2 * dl
        alp, _ = so.spatial_solve_reference(*so.fd_spatial_pencil(so.fd_setup(p, ym, vt, spatial=True)), 0)
        ref = alp[np.argmin(np.abs(alp - ev[0]))]
        assert abs(ev[0] - ref) <= 1e-9 * abs(ref), (st, ev[0], ref)
        assert abs(ev[0] - 0.08 / 0.35) < 0.05


@pytest.mark.gpu
def test_reference_style_caller_links_and_runs(tmp_path):
    """A caller written against the reference's interface only (module stuff,
    input, temporal(name, ind), makename) runs unchanged on this build: each
    eig.<k> holds the QZ eigenvalues nearest 0.35 alpha, the computed pairs in
    the last slots in ascending Im (PIKSR2's order)."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    print("reference-style caller: alpha_march", flush=True)
    shutil.copy(golden("TStest", "profile.0"), tmp_path / "profile.0")
    base = open(golden("thesis_TStest", "temporal.inp")).read().replace("96 1.0 0.0", "48 1.0 0.0")
    lines = base.splitlines()
    k = next(i for i, l in enumerate(lines) if "Profile index" in l)
    text = "\n".join(lines[:k + 1] + ["0.34 3"]) + "\n"
    r = subprocess.run([EXAMPLE], input=text, capture_output=True, text=True, cwd=tmp_path, timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    ym, vt = so.read_profile(golden("TStest", "profile.0"))
    for k, a in enumerate((0.30862069, 0.32431034, 0.34), start=1):
        d = sv.read_evec_dat(str(tmp_path / f"eig.{k}"))
        assert abs(d["alpha"].real - a) < 1e-7
        ev = d["evals"]
        live = np.abs(ev.real) < 1e299
        assert live.sum() == 8 and np.all(live[-8:])                 # computed pairs in the last 8 slots
        assert np.all(np.diff(ev[-8:].imag) >= 0)                    # ascending Im
        p = so.parse_input(base)
        p.alpha = complex(d["alpha"])
        A0, B0 = so.temporal_pencil(so.setup(p, ym, vt, spatial=False))
        w, _ = so.qz(A0, B0, right=False)
        w = w[np.isfinite(w) & (np.abs(w) > 1e-9)]
        for e in ev[-8:]:
            assert np.min(np.abs(w - e)) <= 1e-8 * abs(e)


@pytest.mark.gpu
def test_driver_multi_spatial_stations(tmp_path):
    """itype 8 (mspatial.f90:31-96): stations 1..2 of delta.dat, Yi = 2 delta,
    profile.<station>, omega x beta per station, one batched call; eig.<k>
    holds each point's eigenvalues nearest omega / 0.35, equal to the QZ
    eigenvalues of the oracle's linearised pencil at that station's Yi."""
    from oracle import stab_oracle as so
    from stab_amd import solver as sv
    print("fortran driver: multiple spatial (itype 8)", flush=True)
    for k in (1, 2):
        shutil.copy(golden("TStest", "profile.0"), tmp_path / f"profile.{k}")
    (tmp_path / "delta.dat").write_text("# x delta\n100.0 0.5\n200.0 0.6\n")
    base = open(golden("thesis_TStest", "spatial.inp")).read().replace("96 1.0 0.0", "40 1.0 0.0")
    lines = base.splitlines()
    k = next(i for i, l in enumerate(lines) if l.strip().startswith("2") and "Spatial" in l)
    sweep = lines[:k] + ["8\t\t! multiple spatial", "0.07 0.09 0.01", "0.0 0.0 0.0", "1 2 1", "0",
                         "0.35 0.0 4"]
    r = subprocess.run([BIN], input="\n".join(sweep) + "\n", capture_output=True, text=True, cwd=tmp_path,
                       timeout=300)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    ym, vt = so.read_profile(golden("TStest", "profile.0"))
    iver = 0
    for st, (xs, dl) in enumerate(((100.0, 0.5), (200.0, 0.6)), start=1):
        for om in (0.07, 0.08, 0.09):
            iver += 1
            d = sv.read_evec_dat(str(tmp_path / f"eig.{iver}"))
            assert d["itype"] == 2 and d["ind"] == st and abs(d["x"] - xs) < 1e-12
            assert abs(d["yi"] - 2 * dl) < 1e-12 and abs(d["omega"] - om) < 1e-12
            assert d["evals"].size == 10 * 40
            ev = d["evals"][np.abs(d["evals"].real) < 1e299]
            assert ev.size == 4
            p = so.parse_input(base)
            p.omega, p.yi = complex(om, 0.0), 2 * dl
            C0, C1, C2 = so.spatial_pencil(so.setup(p, ym, vt, spatial=True))
            Aa, Bb = so.spatial_linearization(C0, C1, C2)
            w = so.sla.eigvals(Aa, Bb)
            w = w[np.isfinite(w)]
            for e in ev:
                assert np.min(np.abs(w - e)) <= 1e-8 * abs(e), (iver, e)
