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


def _lu_plan_on_host(tmp_path, npads, gw):
    """Run lu_plan (stab_amd/csrc/lu_plan.h) on the host for points of the given
    padded orders, as one segment in largest-first order; returns the steps
    and the task list."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    src = tmp_path / "plan.cpp"
    src.write_text(r'''
#include "lu_plan.h"
#include <cstdio>
#include <cstdlib>
int main(int argc, char **argv) {
  const int gw = std::atoi(argv[1]);
  std::vector<int> npad;
  for (int i = 2; i < argc; ++i) npad.push_back(std::atoi(argv[i]));
  std::vector<int> order(npad.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return npad[a] > npad[b]; });
  std::vector<LuSeg> segs = {{256, 3, 0, (int)npad.size(), {}}};
  std::vector<int> tasks;
  lu_plan(segs, npad, order, gw, tasks);
  for (const LuStep &s : segs[0].steps) std::printf("S %d %d %d %d\n", s.K0, s.npanel, s.ntask, s.toff);
  for (size_t i = 0; i < tasks.size(); i += 2) std::printf("T %d %d\n", tasks[i], tasks[i + 1]);
  return 0;
}
''')
    exe = tmp_path / "plan"
    r = subprocess.run([cxx, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "stab_amd", "csrc"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe), str(gw)] + [str(n) for n in npads], capture_output=True, text=True)
    assert out.returncode == 0
    steps, tasks = [], []
    for line in out.stdout.split("\n"):
        f = line.split()
        if f and f[0] == "S":
            steps.append(tuple(int(x) for x in f[1:]))
        elif f and f[0] == "T":
            tasks.append((int(f[1]), int(f[2])))
    return steps, tasks


@pytest.mark.parametrize("gw", [64, 32])
@pytest.mark.parametrize("npads", [[640, 480, 320, 640, 320, 480, 160, 1280, 32, 96],
                                   [320] * 13, [640], [64, 32]])
def test_lu_plan_covers_every_strip_once(tmp_path, npads, gw):
    """Every launch of the split LU stays inside its buffers: at each block
    step the panel grid is the prefix of points of order above K0, and the
    trailing grid lists every strip [c0, c0 + gw) right of the block of each
    such point exactly once, c0 in [kend, npad), padding entries -1, all
    within tasks[2 toff .. 2 (toff + ntask)); a matrix's strips share one
    queue (task index mod 8)."""
    steps, tasks = _lu_plan_on_host(tmp_path, npads, gw)
    order = sorted(range(len(npads)), key=lambda i: -npads[i])
    assert [s[0] for s in steps] == list(range(0, max(npads), 64))
    end = 0
    for K0, npanel, ntask, toff in steps:
        assert npanel == sum(1 for n in npads if n > K0)
        assert all(npads[order[i]] > K0 for i in range(npanel))
        assert toff == end and ntask % 8 == 0
        end = toff + ntask
        assert end <= len(tasks)
        want = set()
        for p, n in enumerate(npads):
            if n <= K0:
                continue
            kend = K0 + min(64, n - K0)
            want |= {(p, c0) for c0 in range(kend, n, gw)}
        got = [tasks[b] for b in range(toff, toff + ntask)]
        real = [t for t in got if t[0] >= 0]
        assert len(real) == len(set(real)) == len(want) and set(real) == want
        assert all(t == (-1, 0) for t in got if t[0] < 0)
        queue = {}
        for b, (p, c0) in enumerate(got):
            if p >= 0:
                assert queue.setdefault(p, b % 8) == b % 8
    assert end == len(tasks)


def test_lu_split_keeps_runs_and_shapes(tmp_path):
    """lu_split cuts segments into parts that stay contiguous runs of the
    order, cover it once, keep their work-group shape, and halve the work of
    the segment they split."""
    import shutil
    import subprocess
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    src = tmp_path / "split.cpp"
    src.write_text(r'''
#include "lu_plan.h"
#include <cstdio>
int main() {
  std::vector<int> npad;
  for (int i = 0; i < 1339; ++i) npad.push_back(640);
  for (int i = 0; i < 1364; ++i) npad.push_back(480);
  for (int i = 0; i < 1393; ++i) npad.push_back(320);
  std::vector<int> order(npad.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::vector<LuSeg> segs = {{256, 3, 0, 1339, {}}, {256, 2, 1339, 2757, {}}};
  lu_split(segs, npad, order, 4);
  for (const LuSeg &g : segs) std::printf("%d %d %d %d\n", g.nt, g.R, g.start, g.cnt);
  std::vector<LuSeg> one = {{256, 3, 0, 1, {}}};
  lu_split(one, npad, order, 4);
  std::printf("single %zu\n", one.size());
  return 0;
}
''')
    exe = tmp_path / "split"
    r = subprocess.run([cxx, "-O1", "-std=c++17", "-I", os.path.join(ROOT, "stab_amd", "csrc"), str(src), "-o",
                        str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True).stdout.split("\n")
    segs = [tuple(int(x) for x in l.split()) for l in out if l and not l.startswith("single")]
    assert [l for l in out if l.startswith("single")] == ["single 1"]
    assert len(segs) == 4
    pos = 0
    for nt, R, start, cnt in segs:
        assert start == pos and cnt > 0
        pos += cnt
    assert pos == 4096
    npad = [640] * 1339 + [480] * 1364 + [320] * 1393
    assert all(R == (3 if npad[start] == 640 else 2) for nt, R, start, cnt in segs)
    work = [sum(npad[i] ** 3 for i in range(s, s + c)) for _, _, s, c in segs]
    assert max(work) / min(work) < 2.5
