"""CPU checks of the device-buffer layouts: every index a kernel forms stays
inside the allocation the host makes for it (restated from the kernels'
index expressions), for all supported orders."""
import os

import pytest

from conftest import ROOT

BS = 64          # SOLVE_BS
LU_ALIGN = 32


def npad_of(ny):
    n = 5 * ny
    return -(-n // LU_ALIGN) * LU_ALIGN


def blk_elems(npad, bs=BS):
    nblk = -(-npad // bs)
    return 2 * nblk * bs * bs


@pytest.mark.parametrize("ny", list(range(4, 257)))
def test_block_inverse_layout_fits(ny):
    """k_blkinv writes full BS x BS blocks for inv(L) at b BS^2 and inv(U) at
    (nblk + b) BS^2; lu_solve reads inside the first bs x bs of each."""
    npad = npad_of(ny)
    nblk = -(-npad // BS)
    last_write = (nblk + nblk - 1) * BS * BS + (BS - 1) * BS + (BS - 1)
    assert last_write < blk_elems(npad)


@pytest.mark.parametrize("ny", list(range(4, 257)))
def test_lu_strip_widths_divide(ny):
    """k_lu: npad, the panel width and the 16-column strips tile exactly."""
    npad = npad_of(ny)
    for nb in (16, 32):
        assert npad % nb == 0
        for k0 in range(0, npad, nb):
            assert (npad - k0 - nb) % 16 == 0 and k0 % 16 == 0


@pytest.mark.parametrize("ny", list(range(4, 257)))
def test_solve_gather_buffers_fit(ny):
    """lu_solve gathers npad entries through `work`: z (compact layout, sized
    max(nvec, npad)) or X1/X2 (full layout, 10 ny entries)."""
    npad = npad_of(ny)
    n = 5 * ny
    assert max(n, npad) >= npad
    assert 10 * ny >= npad
    assert npad <= 6 * 256      # lu_solve keeps at most 6 rows per thread


def test_krylov_state_fits():
    """H (m+1 x m, ld = 65) fits before Q for every m <= 64."""
    MMAX = 64
    for m in range(2, MMAX + 1):
        j = m - 1
        assert j * (MMAX + 1) + (j + 1) < (MMAX + 1) * MMAX


def test_normalisation_guards_on_host(tmp_path):
    """The arg-max clamp and the divisor guard of the getevec normalisation
    (stab_amd/csrc/common.h), compiled for the host and run on NaN, Inf, zero
    and ordinary inputs."""
    import shutil
    import subprocess
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(hipcc):
        pytest.skip("hipcc not available")
    src = tmp_path / "guards.cpp"
    src.write_text(r'''
#include "common.h"
#include <cmath>
#include <cstdio>
int main() {
  int bad = 0;
  bad += clamp_index(0x7fffffff, 10) != 0;
  bad += clamp_index(-1, 10) != 0;
  bad += clamp_index(9, 10) != 9;
  bad += clamp_index(10, 10) != 0;
  const double nan = std::nan(""), inf = INFINITY;
  bad += scale_ok(mk(nan, 0.0));
  bad += scale_ok(mk(0.0, inf));
  bad += scale_ok(mk(0.0, 0.0));
  bad += !scale_ok(mk(1e-300, 0.0));
  bad += !scale_ok(mk(-3.0, 4.0));
  bad += !scale_ok(mk(1e308, -1e308));
  std::printf("%d\n", bad);
  return 0;
}
''')
    exe = tmp_path / "guards"
    r = subprocess.run([hipcc, "-x", "hip", "--cuda-host-only", "-O1", "-I",
                        os.path.join(ROOT, "stab_amd", "csrc"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0
    assert out.stdout.strip() == "0"
