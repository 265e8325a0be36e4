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
