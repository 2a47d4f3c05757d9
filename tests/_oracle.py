"""ctypes access to the CPU oracle (oracle/liboracle.so) -- test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
# OBL_ORACLE_SO names another build of the same sources (a planted-mistake build, to check that the pins fail)
ORACLE_SO = os.environ.get("OBL_ORACLE_SO") or os.path.join(ORACLE_DIR, "liboracle.so")

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-j4"], cwd=ORACLE_DIR, check=True)
        _L = C.CDLL(ORACLE_SO)
        _L.orc_pcg.restype = C.c_uint32
        _L.orc_pcg.argtypes = [C.c_uint32]
        _L.orc_hash3i.restype = C.c_uint32
        _L.orc_hash3i.argtypes = [C.c_int32] * 3
        _L.orc_set_threads.argtypes = [C.c_int]
    return _L


def P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dt(sfx):
    return np.float32 if sfx == "f32" else np.float64


class Oracle:
    """One precision (``f32`` or ``f64``) of the kernel restatement, on numpy arrays."""

    def __init__(self, U, BS, sfx="f32", threads=None):
        self.U = np.ascontiguousarray(U, dtype=np.float32)
        self.BS = np.ascontiguousarray(BS, dtype=np.float32)
        self.sfx = sfx
        self.dt = _dt(sfx)
        self.nx, self.ny, self.nz = int(self.U[0]), int(self.U[1]), int(self.U[2])
        self.pad, self.nxp, self.nyp = int(self.U[36]), int(self.U[37]), int(self.U[38])
        lib().orc_set_threads(threads or min(8, os.cpu_count() or 1))

    def _f(self, name):
        f = getattr(lib(), f"orc_{name}_{self.sfx}")
        f.restype = None
        return f

    def a(self, x):
        return np.ascontiguousarray(x, dtype=self.dt)

    def vshape(self):
        return (self.nz + 1, self.ny + 1, self.nx + 1, 4)

    def cshape(self, c=4):
        return (self.nz, self.ny, self.nx, c) if c > 1 else (self.nz, self.ny, self.nx)

    def lsf(self, BD):
        A = np.zeros((self.ny, self.nx, 4), self.dt)
        B = np.zeros_like(A)
        ph = np.zeros(self.cshape(1), self.dt)
        self._f("lsf")(P(self.U), P(self.BS), P(np.ascontiguousarray(BD, np.float32)), P(A), P(B), P(ph))
        return A, B, ph

    def ci(self, CI):
        m = np.zeros((self.ny, self.nx, 4), self.dt)
        ci = np.ascontiguousarray(CI if CI is not None and len(CI) else np.zeros(12), np.float32)
        self._f("ci")(P(self.U), P(self.BS), P(ci), P(m))
        return m

    def init(self, lsfA, lsfB, seed=0):
        V = np.zeros(self.vshape(), self.dt)
        Q1 = np.zeros(self.cshape(), self.dt)
        Q2 = np.zeros(self.cshape(), self.dt)
        f = self._f("init")
        f.argtypes = [C.c_void_p] * 7 + [C.c_int]
        f(P(self.U), P(self.BS), P(self.a(lsfA)), P(self.a(lsfB)), P(V), P(Q1), P(Q2), seed)
        return V, Q1, Q2

    def bubble(self, V, x, y, z, rh, rz, dth):
        V = self.a(V).copy()
        f = self._f("bubble")
        f.argtypes = [C.c_void_p, C.c_void_p] + [C.c_double] * 6
        f(P(self.U), P(V), x, y, z, rh, rz, dth)
        return V

    def nut(self, V, Q1):
        aux = np.zeros(self.cshape(2), self.dt)
        self._f("nut")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(aux))
        return aux

    def advv1(self, V):
        Vt = np.zeros(self.vshape(), self.dt)
        self._f("advv1")(P(self.U), P(self.BS), P(self.a(V)), P(Vt))
        return Vt

    def advv2(self, V, Vt):
        Vo = np.zeros(self.vshape(), self.dt)
        self._f("advv2")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Vt)), P(Vo))
        return Vo

    def advq1(self, V, Q1, Q2):
        a = np.zeros(self.cshape(), self.dt)
        b = np.zeros(self.cshape(), self.dt)
        self._f("advq1")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(self.a(Q2)), P(a), P(b))
        return a, b

    def advq2(self, V, Q1, Q2, Q1t, Q2t):
        a = np.zeros(self.cshape(), self.dt)
        b = np.zeros(self.cshape(), self.dt)
        self._f("advq2")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(self.a(Q2)), P(self.a(Q1t)),
                         P(self.a(Q2t)), P(a), P(b))
        return a, b

    def phys(self, Vin, Q1in, Q2in, aux, lsfA, lsfB, phT, ciMap):
        Vo = np.zeros(self.vshape(), self.dt)
        Q1 = np.zeros(self.cshape(), self.dt)
        Q2 = np.zeros(self.cshape(), self.dt)
        self._f("phys")(P(self.U), P(self.BS), P(self.a(Vin)), P(self.a(Q1in)), P(self.a(Q2in)), P(self.a(aux)),
                        P(self.a(lsfA)), P(self.a(lsfB)), P(self.a(phT)), P(self.a(ciMap)), P(Vo), P(Q1), P(Q2))
        return Vo, Q1, Q2

    def micro_point(self, k, st):
        st = self.a(st).copy()
        f = self._f("micro_point")
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        f(P(self.U), P(self.BS), k, P(st))
        return st

    def div(self, V):
        rhs = np.zeros((self.nz, self.nyp, self.nxp), self.dt)
        self._f("div")(P(self.U), P(self.BS), P(self.a(V)), P(rhs))
        return rhs

    def proj(self, V, phi):
        Vo = np.zeros(self.vshape(), self.dt)
        self._f("proj")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(phi)), P(Vo))
        return Vo

    # ---- hurricane mode ----
    @staticmethod
    def f32(x):
        return None if x is None else np.ascontiguousarray(x, np.float32)

    def init2(self, lsfA, lsfB, VX, seed=0):
        V = np.zeros(self.vshape(), self.dt)
        Q1 = np.zeros(self.cshape(), self.dt)
        Q2 = np.zeros(self.cshape(), self.dt)
        f = self._f("init2")
        f.argtypes = [C.c_void_p] * 8 + [C.c_int]
        f(P(self.U), P(self.BS), P(self.a(lsfA)), P(self.a(lsfB)), P(self.f32(VX)), P(V), P(Q1), P(Q2), seed)
        return V, Q1, Q2

    def nut2(self, V, Q1, sfc):
        aux = np.zeros(self.cshape(2), self.dt)
        self._f("nut2")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(self.f32(sfc)), P(aux))
        return aux

    def advv1_2(self, V, Vval):
        Vt = np.zeros(self.vshape(), self.dt)
        self._f("advv1_2")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Vval)), P(Vt))
        return Vt

    def advv2_2(self, V, Vval, Vt):
        Vo = np.zeros(self.vshape(), self.dt)
        self._f("advv2_2")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Vval)), P(self.a(Vt)), P(Vo))
        return Vo

    def phys2(self, Vin, Q1in, Q2in, aux, lsfA, lsfB, phT, ciMap, sfc, phi):
        Vo = np.zeros(self.vshape(), self.dt)
        Q1 = np.zeros(self.cshape(), self.dt)
        Q2 = np.zeros(self.cshape(), self.dt)
        self._f("phys2")(P(self.U), P(self.BS), P(self.a(Vin)), P(self.a(Q1in)), P(self.a(Q2in)), P(self.a(aux)),
                         P(self.a(lsfA)), P(self.a(lsfB)), P(self.a(phT)), P(self.a(ciMap)), P(self.f32(sfc)),
                         P(None if phi is None else self.a(phi)), P(Vo), P(Q1), P(Q2))
        return Vo, Q1, Q2

    def maps2(self, V, Q1, Q2, phi, sfc):
        M = np.zeros((4, self.ny, self.nx, 4), self.dt)
        self._f("maps2")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(self.a(Q2)), P(self.a(phi)),
                         P(self.f32(sfc)), P(M))
        return M

    def prep(self, V, phi):
        Vp = np.zeros(self.vshape(), self.dt)
        self._f("prep")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(phi)), P(Vp))
        return Vp

    def filt(self, V, Q1, Q2):
        Vo = np.zeros(self.vshape(), self.dt)
        Q1o = np.zeros(self.cshape(), self.dt)
        Q2o = np.zeros(self.cshape(), self.dt)
        self._f("filt")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(self.a(Q2)), P(Vo), P(Q1o), P(Q2o))
        return Vo, Q1o, Q2o

    def maps(self, V, Q1, Q2, phi):
        """The four 2-D maps (model.js:1028-1100), shape (4, ny, nx, 4): m0..m3."""
        M = np.zeros((4, self.ny, self.nx, 4), self.dt)
        self._f("maps")(P(self.U), P(self.BS), P(self.a(V)), P(self.a(Q1)), P(self.a(Q2)), P(self.a(phi)), P(M))
        return M

    def densities(self):
        rc = self.BS[2:8 * self.nz:8].astype(np.float64)[: self.nz]
        rw = self.BS[3::8].astype(np.float64)[: self.nz + 1]
        return rc, rw

    def solve(self, rhs, mode="f64"):
        return solve(self.nxp, self.nyp, self.nz, float(self.U[3]), float(self.U[4]), *self.densities(), rhs, mode)


def solve(nx, ny, nz, dx, dz, rhoC, rhoW, rhs, mode="f64", threads=None):
    """FP64 oracle solve (mode f64) or the reference's FP32 arithmetic (mode f32ref)."""
    rc = np.ascontiguousarray(rhoC, np.float64)
    rw = np.ascontiguousarray(rhoW, np.float64)
    nth = threads or min(8, os.cpu_count() or 1)
    if mode == "f64":
        r = np.ascontiguousarray(rhs, np.float64)
        out = np.zeros_like(r)
        f = lib().orc_solve_f64
    else:
        r = np.ascontiguousarray(rhs, np.float32)
        out = np.zeros_like(r)
        f = lib().orc_solve_f32ref
    f.restype = None
    f.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int]
    f(nx, ny, nz, dx, dz, P(rc), P(rw), P(r), P(out), nth)
    return out


def dst_matrix(N):
    q = np.zeros((N, N))
    qi = np.zeros((N, N))
    f = lib().orc_dst_matrix_f64
    f.restype = None
    f.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    f(N, P(q), P(qi))
    return q, qi


def dst_eig(N, dx):
    lam = np.zeros(N)
    f = lib().orc_dst_eig_f64
    f.restype = None
    f.argtypes = [C.c_int, C.c_double, C.c_void_p]
    f(N, dx, P(lam))
    return lam


def step_f32(U, BS, BD, CI, run_lsf, parity, V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi,
             threads=None):
    """The whole FP32 step; every array is modified in place (float32, contiguous)."""
    f = lib().orc_step_f32
    f.restype = None
    f.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int] + [C.c_void_p] * 13 + [C.c_int]
    arrs = [V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi]
    for a in arrs:
        assert a.dtype == np.float32 and a.flags.c_contiguous
    ci = np.ascontiguousarray(CI if CI is not None and len(CI) else np.zeros(12), np.float32)
    f(P(np.ascontiguousarray(U, np.float32)), P(np.ascontiguousarray(BS, np.float32)),
      P(np.ascontiguousarray(BD, np.float32)), P(ci), int(run_lsf), int(parity), *[P(a) for a in arrs],
      threads or min(8, os.cpu_count() or 1))


def step_f64(U, BS, BD, CI, run_lsf, parity, V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi,
             threads=None):
    """The whole step in FP64 (FP64 kernels and solve); every array is float64, modified in place."""
    f = lib().orc_step_f64
    f.restype = None
    f.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int] + [C.c_void_p] * 13 + [C.c_int]
    arrs = [V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi]
    for a in arrs:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    ci = np.ascontiguousarray(CI if CI is not None and len(CI) else np.zeros(12), np.float32)
    f(P(np.ascontiguousarray(U, np.float32)), P(np.ascontiguousarray(BS, np.float32)),
      P(np.ascontiguousarray(BD, np.float32)), P(ci), int(run_lsf), int(parity), *[P(a) for a in arrs],
      threads or min(8, os.cpu_count() or 1))


def step_f64_split(U, BS, BD, CI, run_lsf, parity, V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi,
                   part, threads=None):
    """The FP64 step in two parts around the solve: part 0 every kernel up to div (writes rhs), part 1 the
    projection from phi; every array is float64, modified in place."""
    f = lib().orc_step_f64_split
    f.restype = None
    f.argtypes = [C.c_void_p] * 4 + [C.c_int, C.c_int] + [C.c_void_p] * 13 + [C.c_int, C.c_int]
    arrs = [V2, Q12, Q22, Vt, Q1t, Q2t, aux, phT, lsfA, lsfB, ciMap, rhs, phi]
    for a in arrs:
        assert a.dtype == np.float64 and a.flags.c_contiguous
    ci = np.ascontiguousarray(CI if CI is not None and len(CI) else np.zeros(12), np.float32)
    f(P(np.ascontiguousarray(U, np.float32)), P(np.ascontiguousarray(BS, np.float32)),
      P(np.ascontiguousarray(BD, np.float32)), P(ci), int(run_lsf), int(parity), *[P(a) for a in arrs],
      threads or min(8, os.cpu_count() or 1), int(part))


def radar_gates(ru, BS, V, Q1, Q2, aux, tor, rg, gg=None):
    """Listed gates' moment values before rounding and decision margins (orc_radar_gates): rg = (n, 2) int
    (radial, gate), gg = None or the GPU's geometry of each (n, 7: x, y, ex, ey, zc, hl, hu); returns [n][26] =
    REF, VEL, SW, ZDR, RHO values (v * scale + offset, NaN where not formed), the margins mgeo (m), mecho (dB),
    mwrap (m/s), mT (K), mzg (relative), md0 (mm), mq (relative), the five values at the GPU's geometry (NaN
    without gg), the oracle's own geometry (7), then the gate's KDP (-1 where no echo) at the oracle's geometry
    and at the GPU's (NaN without gg)."""
    rg = np.ascontiguousarray(rg, np.int32).reshape(-1, 2)
    out = np.zeros((len(rg), 26), np.float32)
    if not len(rg):
        return out
    f = lib().orc_radar_gates
    f.restype = None
    a32 = lambda x: np.ascontiguousarray(x, np.float32)  # noqa: E731
    t = a32(tor if tor is not None and len(tor) else np.zeros(8, np.float32))
    g = None if gg is None else a32(gg).reshape(len(rg), 7)
    f(P(a32(ru)), P(a32(BS)), P(a32(V)), P(a32(Q1)), P(a32(Q2)), P(a32(aux)), P(t), C.c_int(len(rg)), P(rg), P(g),
      P(out))
    return out


def radar_cut(ru, BS, V, Q1, Q2, aux, tor, words, kdp_len, out=None):
    """One radar cut (radar.js:41-268) into a uint32 volume of ``words`` words (updated in place when
    ``out`` is given); returns (out, kdp)."""
    out = np.zeros(words, np.uint32) if out is None else out
    kdp = np.zeros(kdp_len, np.float32)
    f = lib().orc_radar_cut
    f.restype = None
    a32 = lambda x: np.ascontiguousarray(x, np.float32)  # noqa: E731
    t = a32(tor if tor is not None and len(tor) else np.zeros(8, np.float32))
    f(P(a32(ru)), P(a32(BS)), P(a32(V)), P(a32(Q1)), P(a32(Q2)), P(a32(aux)), P(t), P(out), P(kdp))
    return out, kdp


def radar_geom(ru, ri, gi):
    """(beam height above the antenna, ground distance, model x, model y) of gate gi on radial ri."""
    o = np.zeros(4, np.float32)
    f = lib().orc_radar_geom
    f.restype = None
    f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
    f(P(np.ascontiguousarray(ru, np.float32)), ri, gi, P(o))
    return o
