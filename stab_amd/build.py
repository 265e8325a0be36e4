"""Build libstab_hip.so in-tree with hipcc for gfx950."""
from __future__ import annotations

import os
import shutil
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(HERE, "csrc", "api.hip")
OUT = os.path.join(HERE, "libstab_hip.so")
DEPS = ("api.hip", "prep.hip", "assemble.hip", "lu.hip", "eig.hip", "pencil.h", "common.h")

HIPCC_FLAGS = ["-O3", "--offload-arch=gfx950", "-fPIC", "-shared", "-std=c++17",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


FORTRAN_DIR = os.path.join(HERE, "fortran")
FORTRAN_BIN = os.path.join(HERE, "stab_gpu")
GETEVEC_BIN = os.path.join(HERE, "getevec")
EXAMPLE_BIN = os.path.join(HERE, "alpha_march")


def fortran_compiler():
    for cand in (shutil.which("amdflang"), "/opt/rocm/bin/amdflang", shutil.which("flang"), shutil.which("gfortran")):
        if cand and os.path.exists(cand):
            return cand
    return None


def _fc(fc, work, sources, out, link_lib=True):
    tmp = out + ".tmp"
    cmd = [fc, "-O2", "-J", work, *sources]
    if link_lib:
        cmd += ["-L" + HERE, "-lstab_hip", "-Wl,-rpath,$ORIGIN"]
    cmd += ["-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True, cwd=work)
    if r.returncode != 0:
        raise RuntimeError("Fortran build failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    os.replace(tmp, out)


def build_fortran() -> str | None:
    """Compile the Fortran layer (ISO_C_BINDING onto the C ABI) and the programs
    over it next to the library: stab_gpu (the main program), getevec (host
    only) and alpha_march (a caller written against the reference's interface).
    Returns the stab_gpu path, or None when no Fortran compiler exists."""
    fc = fortran_compiler()
    if fc is None:
        return None
    work = os.path.join(ROOT, "_build", "fortran")
    os.makedirs(work, exist_ok=True)
    layer = [os.path.join(FORTRAN_DIR, "stab_hip_iface.f90"), os.path.join(FORTRAN_DIR, "stab_path.f90")]
    _fc(fc, work, layer + [os.path.join(FORTRAN_DIR, "stab_gpu.f90")], FORTRAN_BIN)
    _fc(fc, work, layer + [os.path.join(FORTRAN_DIR, "examples", "alpha_march.f90")], EXAMPLE_BIN)
    _fc(fc, work, [os.path.join(FORTRAN_DIR, "getevec.f90")], GETEVEC_BIN, link_lib=False)
    return FORTRAN_BIN


def build(verbose: bool = False) -> str:
    """Compile the HIP library (always recompiles) and return its path."""
    tmp = OUT + ".tmp"
    cmd = [hipcc(), *HIPCC_FLAGS, "-I" + os.path.join(ROOT, "include"), SRC, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stdout + r.stderr)
    if verbose:
        print(r.stderr)
    os.replace(tmp, OUT)
    build_fortran()
    return OUT


if __name__ == "__main__":
    print(build(verbose=False))
