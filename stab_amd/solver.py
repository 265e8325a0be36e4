"""Host-side mirror of stab's temporal/spatial interface over libstab_hip.

The reference drives one stability solve per process through module globals
(stuff.f90:9-55), an ``*.inp`` stream (input.f90:2-165), the profile file
``profile.<ind>`` (getmean.f90:27-32) and the subroutines
``temporal(name, ind)`` / ``spatial(name, ind)`` (temporal.f90:2,
spatial.f90:2), which write ``evec.dat`` (temporal.f90:883-890).  This module
keeps those names, inputs and files, and sends the arithmetic to the GPU
through the C ABI.  It never computes eigenpairs on the CPU.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
import struct
from typing import Optional, Sequence

import numpy as np

from . import _lib
from ._lib import Point, Stats, Z

PT5 = 0.5
GAMMA1 = 0.4


# ----------------------------------------------------------------------------
# input.f90 semantics
# ----------------------------------------------------------------------------
@dataclasses.dataclass
class Inputs:
    """The run parameters of module ``stuff`` that the Chebyshev path reads."""
    mattyp: int = 0
    T0: float = 0.0
    Te: float = 0.0
    Ma: float = 0.0
    Re: float = 0.0
    Pr: float = 0.0
    ny: int = 1
    yi: float = 0.0
    ymax: float = 0.0
    ievec: int = 1
    ider: int = 1
    top: int = 0
    wall: int = 0
    wallt: int = 0
    curve: int = 0
    itype: int = 1
    alpha: complex = 0j
    beta: complex = 0j
    omega: complex = 0j
    ind: int = 0
    x: float = 0.0


def _fields(line: str):
    return [t for t in line.split("!")[0].replace(",", " ").split() if t]


def read_input(text: str) -> Inputs:
    """Read an ``*.inp`` stream in the order of input.f90:14-128 and
    stab.f90:46-55.  Te is set before Ma is read, so Te = T0 when mattyp = 1
    (input.f90:24-25); with mattyp = 0 it stays 0."""
    lines = iter(text.splitlines())

    def nxt():
        for ln in lines:
            f = _fields(ln)
            if f:
                return f
        raise ValueError("input stream ended early")

    p = Inputs()
    p.mattyp = int(nxt()[0])
    if p.mattyp not in (0, 1):
        raise ValueError(f"mattyp {p.mattyp} must be 0 or 1")
    if p.mattyp == 1:
        p.T0 = float(nxt()[0])
        p.Te = p.T0 / (1.0 + PT5 * GAMMA1 * 0.0 ** 2)
    f = nxt(); p.Ma, p.Re, p.Pr = float(f[0]), float(f[1]), float(f[2])
    f = nxt(); p.ny, p.yi, p.ymax = int(f[0]), float(f[1]), float(f[2])
    p.ievec = int(nxt()[0])
    p.ider = int(nxt()[0])
    f = nxt(); p.top, p.wall, p.wallt, p.curve = (int(v) for v in f[:4])
    p.itype = int(nxt()[0])
    if p.itype in (1, 3):
        f = nxt(); p.alpha = complex(float(f[0]), float(f[1]))
        f = nxt(); p.beta = complex(float(f[0]), float(f[1]))
    elif p.itype in (2, 4):
        f = nxt(); p.omega = complex(float(f[0]), float(f[1]))
        f = nxt(); p.beta = complex(float(f[0]), float(f[1]))
    else:
        raise ValueError(f"itype {p.itype}: only the Chebyshev (1, 2) and finite-difference (3, 4) "
                         "temporal and spatial paths are here")
    p.ind = int(nxt()[0])
    if p.itype in (2, 4):
        p.x = float(nxt()[0])
    if p.curve != 0:
        raise ValueError("curve != 0 (wall curvature) is outside this path")
    return p


def profile_name(base: str, iver: int) -> str:
    """makename (mtemporal.f90:53-76): base.<iver>."""
    if iver >= 10000:
        raise ValueError("Error in MakeName:  iver too large")
    return f"{base}.{iver}"


def read_profile(path: str):
    """Rows ``y rho u v w T``; '#' lines skipped (getmean.f90:40-80)."""
    rows = []
    try:
        fh = open(path)
    except OSError as e:
        raise FileNotFoundError(f"Error reading profile... ({path})") from e
    with fh:
        for line in fh:
            if line[:1] == "#":
                continue
            f = _fields(line)
            if f:
                rows.append([float(v.replace("D", "E").replace("d", "e")) for v in f[:6]])
    a = np.array(rows, dtype=np.float64)
    return a[:, 0].copy(), np.ascontiguousarray(a[:, 1:6])


# ----------------------------------------------------------------------------
# GPU context
# ----------------------------------------------------------------------------
class StabError(RuntimeError):
    pass


def make_point(inp: Inputs, shift: complex, nev: int = 8, profile: int = 0) -> Point:
    pt = Point()
    pt.itype, pt.ny, pt.mattyp, pt.wallt, pt.top = inp.itype, inp.ny, inp.mattyp, inp.wallt, inp.top
    pt.profile, pt.nev, pt.ider = profile, nev, inp.ider
    pt.Ma, pt.Re, pt.Pr, pt.Te, pt.yi, pt.ymax = inp.Ma, inp.Re, inp.Pr, inp.Te, inp.yi, inp.ymax
    for nm in ("alpha", "beta", "omega"):
        v = complex(getattr(inp, nm))
        setattr(pt, nm, Z(v.real, v.imag))
    pt.shift = Z(shift.real, shift.imag)
    return pt


def clone_point(pt: Point) -> Point:
    """Field-by-field copy of a Point."""
    q = Point()
    C.pointer(q)[0] = pt
    return q


@dataclasses.dataclass
class SolveResult:
    evals: np.ndarray          # (npts, nev_max) complex, nearest the shift first
    evecs: Optional[list]      # per point: (nvec, nev) complex
    resid: np.ndarray          # (npts, nev_max)
    info: np.ndarray           # (npts,)
    stats: dict


class Solver:
    """One HIP context: profile slots plus batched solves."""

    def __init__(self, device: int = 0):
        self.lib = _lib.load()
        self.ctx = C.c_void_p()
        rc = self.lib.stab_ctx_create(C.byref(self.ctx), device)
        if rc != 0:
            raise StabError(f"stab_ctx_create failed ({rc})")

    def close(self):
        if self.ctx:
            self.lib.stab_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise StabError(f"{self.lib.stab_last_error(self.ctx).decode()} (status {rc})")

    def set_profile(self, slot: int, y, vt):
        y = np.ascontiguousarray(y, dtype=np.float64)
        vt = np.asfortranarray(vt, dtype=np.float64)   # nym x 5, column-major
        dp = C.POINTER(C.c_double)
        self._check(self.lib.stab_profile_set(self.ctx, slot, y.size, y.ctypes.data_as(dp),
                                              vt.ctypes.data_as(dp)))

    def grid(self, pt: Point):
        ny = pt.ny
        outs = [np.zeros(ny) for _ in range(4)] + [np.zeros((ny, 5), order="F"),
                                                    np.zeros((ny, ny), order="F"), np.zeros((ny, ny), order="F")]
        dp = C.POINTER(C.c_double)
        self._check(self.lib.stab_grid(self.ctx, C.byref(pt), *[o.ctypes.data_as(dp) for o in outs]))
        return dict(zip(("y", "eta", "deta", "d2eta", "vm", "d1", "d2"), outs))

    def pencil(self, pt: Point, which: int) -> np.ndarray:
        n = 5 * pt.ny
        out = np.zeros((n, n), dtype=np.complex128, order="F")
        self._check(self.lib.stab_pencil(self.ctx, C.byref(pt), which, out.ctypes.data_as(C.POINTER(Z))))
        return out

    def shifted_solve(self, pt: Point, b, mode: int = _lib.MODE_PLAIN_F64, ir_steps: int = 0):
        """Factor the shifted matrix of `pt` and solve M x = b."""
        b = np.ascontiguousarray(b, dtype=np.complex128)
        x = np.zeros_like(b)
        info = C.c_int32()
        zp = C.POINTER(Z)
        self._check(self.lib.stab_shifted_solve(self.ctx, C.byref(pt), mode, ir_steps, b.ctypes.data_as(zp),
                                                x.ctypes.data_as(zp), C.byref(info)))
        return x, info.value

    def lu_factors(self, pt: Point, mode: int = _lib.MODE_PLAIN_F64):
        """Packed LU factors, row permutation and equilibration scales of the
        shifted matrix (see stab_lu_factors)."""
        npad = -(-5 * pt.ny // 32) * 32
        lu = np.zeros((npad, npad), dtype=np.complex128, order="F")
        perm = np.zeros(npad, dtype=np.int32)
        scales = np.zeros(2 * npad)
        info = C.c_int32()
        self._check(self.lib.stab_lu_factors(self.ctx, C.byref(pt), mode, lu.ctypes.data_as(C.POINTER(Z)),
                                             perm.ctypes.data_as(C.POINTER(C.c_int32)),
                                             scales.ctypes.data_as(C.POINTER(C.c_double)), C.byref(info)))
        return lu, perm, scales[:npad], scales[npad:], info.value

    def solve_leading(self, points: Sequence[Point], mode: int = _lib.MODE_PLAIN_F64,
                      method: int = _lib.LEAD_KRYLOV, maxit: int = 30, fallback_shifts=None,
                      want_vectors: bool = False, shifts=None, nev=None):
        """Leading (least stable discrete) eigenmode per point (stab_solve_leading).
        `shifts` (complex, one per point) replaces the points' shifts, `nev` their
        pair counts."""
        npts = len(points)
        arr = (Point * npts)(*points)
        if nev is not None:
            np.ctypeslib.as_array(arr)["nev"] = int(nev)
        if shifts is not None:
            sh = np.asarray(shifts, dtype=np.complex128)
            assert sh.shape == (npts,)
            view = np.ctypeslib.as_array(arr)
            view["shift"]["re"] = sh.real
            view["shift"]["im"] = sh.imag
        ev = np.zeros(npts, dtype=np.complex128)
        rs = np.zeros(npts)
        ff = np.zeros(npts)
        it = np.zeros(npts, dtype=np.int32)
        info = np.zeros(npts, dtype=np.int32)
        st = Stats()
        zp = C.POINTER(Z)
        fb = None
        if fallback_shifts is not None:
            fb = np.ascontiguousarray(np.asarray(fallback_shifts, dtype=np.complex128))
            assert fb.shape == (npts,)
        nv = [(10 if p.itype in (2, 4) else 5) * p.ny for p in points]
        vec = np.zeros(sum(nv), dtype=np.complex128) if want_vectors else None
        ip = C.POINTER(C.c_int32)
        self._check(self.lib.stab_solve_leading(
            self.ctx, npts, arr, mode, method, maxit, fb.ctypes.data_as(zp) if fb is not None else None,
            ev.ctypes.data_as(zp), vec.ctypes.data_as(zp) if vec is not None else None,
            rs.ctypes.data_as(C.POINTER(C.c_double)), ff.ctypes.data_as(C.POINTER(C.c_double)),
            it.ctypes.data_as(ip), info.ctypes.data_as(ip), C.byref(st)))
        vecs = None
        if vec is not None:
            offs = np.cumsum([0] + nv)
            vecs = [vec[offs[i]:offs[i + 1]] for i in range(npts)]
        return dict(evals=ev, resid=rs, farfield=ff, iters=it, info=info, evecs=vecs, stats={f: getattr(st, f) for f, _ in Stats._fields_})

    def copy_probe(self, nbytes: int, reps: int = 10) -> float:
        """Mean ms of one device-to-device copy of nbytes (stab_copy_probe)."""
        ms = C.c_double()
        self._check(self.lib.stab_copy_probe(self.ctx, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def solve(self, points: Sequence[Point], mode: int = _lib.MODE_PLAIN_F64, want_vectors: bool = True):
        npts = len(points)
        arr = (Point * npts)(*points)
        nev = _lib.STAB_MAX_NEV
        ev = np.zeros(npts * nev, dtype=np.complex128)
        rs = np.zeros(npts * nev)
        info = np.zeros(npts, dtype=np.int32)
        st = Stats()
        evbuf = None
        if want_vectors:
            nbytes = self.lib.stab_evec_bytes(npts, arr)
            evbuf = np.zeros(nbytes // 16, dtype=np.complex128)
        zp = C.POINTER(Z)
        self._check(self.lib.stab_solve(self.ctx, npts, arr, mode, ev.ctypes.data_as(zp),
                                        evbuf.ctypes.data_as(zp) if evbuf is not None else None,
                                        rs.ctypes.data_as(C.POINTER(C.c_double)),
                                        info.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st)))
        vecs = None
        if evbuf is not None:
            vecs, off = [], 0
            for p in points:
                nv = (10 if p.itype in (2, 4) else 5) * p.ny
                blk = evbuf[off:off + nv * nev].reshape(nev, nv).T
                vecs.append(blk[:, :p.nev].copy())
                off += nv * nev
        stats = {f: getattr(st, f) for f, _ in Stats._fields_}
        return SolveResult(ev.reshape(npts, nev), vecs, rs.reshape(npts, nev), info, stats)


# ----------------------------------------------------------------------------
# evec.dat (temporal.f90:883-890; spatial.f90:1120-1126)
# ----------------------------------------------------------------------------
def _rec(payload: bytes) -> bytes:
    n = struct.pack("<i", len(payload))
    return n + payload + n


def write_evec_dat(path: str, inp: Inputs, grid: dict, evals: np.ndarray, evecs: Optional[np.ndarray]):
    """gfortran sequential unformatted, 4-byte record markers.  Record 4 holds the
    computed eigenvalues (the nev nearest the shift), record 5 the matching
    eigenvectors column by column."""
    ndof = 5
    itype = {3: 1, 4: 2}.get(inp.itype, inp.itype)     # stab.f90:59,66: fd_* run with itype 1 / 2
    rec1 = struct.pack("<10i", inp.ind, inp.ny, ndof, itype, inp.ievec, inp.curve, inp.top, inp.wall,
                       inp.wallt, 1 if inp.ider else 0)
    rec2 = struct.pack("<6d", inp.omega.real, inp.omega.imag, inp.alpha.real, inp.alpha.imag, inp.beta.real,
                       inp.beta.imag) + struct.pack("<3d", inp.Re, inp.Ma, inp.Pr)
    rec3 = np.concatenate([[inp.x], grid["y"], grid["eta"], grid["deta"], grid["d2eta"],
                           [inp.yi, inp.ymax]]).astype("<f8").tobytes()
    rec4 = np.asarray(evals, dtype="<c16").tobytes()
    with open(path, "wb") as f:
        for r in (rec1, rec2, rec3, rec4):
            f.write(_rec(r))
        if evecs is not None and inp.ievec == 1:
            f.write(_rec(np.asfortranarray(evecs, dtype="<c16").tobytes(order="F")))


def read_evec_dat(path: str) -> dict:
    with open(path, "rb") as f:
        data = f.read()
    recs, pos = [], 0
    while pos < len(data):
        (ln,) = struct.unpack_from("<i", data, pos)
        recs.append(data[pos + 4:pos + 4 + ln])
        pos += 8 + ln
    ind, ny, ndof, itype, ievec, curve, top, wall, wallt, ider = struct.unpack("<10i", recs[0])
    om, al, be = (complex(*struct.unpack_from("<2d", recs[1], 16 * k)) for k in range(3))
    Re, Ma, Pr = struct.unpack_from("<3d", recs[1], 48)
    r3 = np.frombuffer(recs[2], dtype="<f8")
    ev = np.frombuffer(recs[3], dtype="<c16").copy()
    out = dict(ind=ind, ny=ny, itype=itype, ievec=ievec, omega=om, alpha=al, beta=be, Re=Re, Ma=Ma, Pr=Pr,
               x=r3[0], y=r3[1:1 + ny].copy(), yi=r3[1 + 4 * ny], ymax=r3[2 + 4 * ny], evals=ev, evec=None)
    if len(recs) > 4:
        out["evec"] = np.frombuffer(recs[4], dtype="<c16").reshape(-1, ev.size, order="F").copy()
    return out


# ----------------------------------------------------------------------------
# getevec (getevec.f90:160-227)
# ----------------------------------------------------------------------------
def select_nearest(evals, guess: complex) -> int:
    """First index minimising |guess - eval| (getevec.f90:160-173)."""
    return int(np.argmin(np.abs(guess - np.asarray(evals))))


def normalize(vec, nrows):
    """Divide by the first entry of largest modulus among the first 5 ny
    components (getevec.f90:179-189)."""
    k = int(np.argmax(np.abs(vec[:nrows])))
    s = vec[k]
    return vec / s if s != 0 else vec.copy()


def _e21(v: float) -> str:
    s = "%.13E" % v
    m, e = s.split("E")
    e = int(e)
    return ("%s%s%03d" % (m, "E+" if e >= 0 else "E-", abs(e))).rjust(21)


def _e13(v: float) -> str:
    s = "%.6E" % v
    m, e = s.split("E")
    return ("%sE%+03d" % (m, int(e))).rjust(13)


def eigenfunction_text(d: dict, j: int) -> str:
    """time.N / space.N text for eigenpair j of an evec.dat record set
    (getevec.f90:193-227, format 1p,11(1pe21.13E3,1x))."""
    ny = d["ny"]
    vec = normalize(d["evec"][:, j], 5 * ny)
    lines = ["# Re = %s, Ma = %s, Pr = %s" % (_e13(d["Re"]), _e13(d["Ma"]), _e13(d["Pr"]))]
    ev = d["evals"][j]
    om, al = (ev, d["alpha"]) if d["itype"] in (1, 3) else (d["omega"], ev)
    be = d["beta"]
    lines.append("# Omega = (%s,%s)" % (_e21(om.real), _e21(om.imag)))
    lines.append("# Alpha = (%s,%s)" % (_e21(al.real), _e21(al.imag)))
    lines.append("# Beta  = (%s,%s)" % (_e21(be.real), _e21(be.imag)))
    for i in range(ny - 1, -1, -1):
        vals = [d["y"][i]]
        for k in range(5):
            z = vec[i * 5 + k]
            vals += [z.real, z.imag]
        lines.append(" ".join(_e21(v) for v in vals) + " ")
    return "\n".join(lines) + "\n"


# ----------------------------------------------------------------------------
# temporal / spatial drop-ins
# ----------------------------------------------------------------------------
def run(name: str, inp: Inputs, shift: complex, nev: int = 8, profile_dir: str = ".",
        mode: int = _lib.MODE_PLAIN_F64, solver: Optional[Solver] = None):
    """temporal(name, ind) / spatial(name, ind) with the arithmetic on the GPU:
    reads profile.<ind>, solves for the nev eigenpairs nearest ``shift`` and
    writes ``name`` in the evec.dat layout.  Returns (evals, evecs, resid, info)."""
    own = solver is None
    s = solver or Solver()
    try:
        y, vt = read_profile(os.path.join(profile_dir, profile_name("profile", inp.ind)))
        s.set_profile(0, y, vt)
        pt = make_point(inp, shift, nev, 0)
        res = s.solve([pt], mode)
        g = s.grid(pt)
        evals = res.evals[0, :nev]
        evecs = res.evecs[0]
        if name:
            write_evec_dat(name, inp, g, evals, evecs)
        return evals, evecs, res.resid[0, :nev], int(res.info[0])
    finally:
        if own:
            s.close()
