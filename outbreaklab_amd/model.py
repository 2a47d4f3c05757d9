"""Host mirror of the reference's step API: RegionalModel (model.js:1156-1496) and
PoissonDirect (solver.js:105-165), driving the gfx950 library through its C ABI.

Method names, argument shapes, parity convention and error texts follow the
reference. ``enc`` arguments are accepted for signature compatibility and
ignored: work is enqueued on the model's own HIP stream in call order, which
is what the reference's single command queue guarantees.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib
from .grid import GRID_PRESETS, grid_bytes, grid_pad
from .sounding import analyze, base_state, parse_sharppy
from .uniforms import CN_OLD, DEG, OUTBREAK_OVERRIDES, UBlock, default_params, write_uniforms

MAX_BOUNDS, MAX_BPTS, MAX_CI = 8, 32, 256


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


class Field:
    """A device field of the model: download()/upload() copy whole arrays (numpy) in the reference layout."""

    def __init__(self, model: "RegionalModel", name: str, copy: int, shape, dtype=np.float32):
        self.model, self.name, self.copy, self.shape, self.dtype = model, name, copy, tuple(shape), dtype

    @property
    def ptr(self) -> int:
        p, n = C.c_void_p(), C.c_size_t()
        _lib.check(_lib.lib().obl_field_ptr(self.model.h, _lib.FIELD[self.name], self.copy, C.byref(p), C.byref(n)),
                   self.model.h)
        return p.value

    def download(self) -> np.ndarray:
        out = np.empty(self.shape, dtype=self.dtype)
        _lib.check(_lib.lib().obl_download(self.model.h, _lib.FIELD[self.name], self.copy, _ptr(out), out.nbytes, 0),
                   self.model.h)
        return out

    def upload(self, a) -> None:
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size * a.itemsize != int(np.prod(self.shape)) * np.dtype(self.dtype).itemsize:
            raise ValueError(f"{self.name}: expected {self.shape}, got {a.shape}")
        _lib.check(_lib.lib().obl_upload(self.model.h, _lib.FIELD[self.name], self.copy, _ptr(a), a.nbytes, 0),
                   self.model.h)


VX_MAXR = 2048  # radii of the vortex table (model.js:38)


def vortex_table(nz, dz, dx, lat, B, v):
    """setVortex's table (model.js:1318-1365), built by the library's host routine obl_vortex_table (the one the
    Node host calls too): tangential wind v(r, z) = V(r) F(z) with V a Rankine-type profile held inertially stable,
    phi in gradient-wind balance, theta' = theta0 (d phi / dz) / g. Returns (table nz x nr x 4 float32, info)."""
    spec = np.array([v["vmax"], v["rmw"], v["r0"], v["ztop"]], np.float64)
    Bf = np.ascontiguousarray(B, np.float32)
    tab = np.zeros((nz, VX_MAXR, 4), np.float32)
    out = np.zeros(5)
    rc = _lib.lib().obl_vortex_table(int(nz), float(dz), float(dx), float(lat), _ptr(Bf), _ptr(spec),
                                     1 if v.get("shape") == "mid" else 0, _ptr(tab), tab.size, _ptr(out))
    if rc != 0:
        raise ValueError(f"setVortex: bad vortex (code {rc})")
    nr = int(out[1])
    tab = np.ascontiguousarray(tab.reshape(-1)[:nz * nr * 4].reshape(nz, nr, 4))
    mr = v.get("moistR") or 0
    info = {"x": v["x"], "y": v["y"], "dr": float(out[0]), "nr": nr, "moistR": mr or 1,
            "moistRH": v.get("moistRH", 0) if mr else 0, "pc": float(out[2]), "minStab": float(out[3]),
            "warm": float(out[4])}
    return tab, info


class RegionalModel:
    """RegionalModel (model.js:1156-1496) on MI355X."""

    def __init__(self):
        self.U = UBlock()
        self.params = default_params()
        self.params.update(OUTBREAK_OVERRIDES)
        self.parity = 0
        self.time = 0.0
        self.steps = 0
        self.bounds = []
        self.ciEvents = []
        self.ciActive = 0
        self.rhoSfc = None
        self.phiScale = 1
        self.hur = None
        self.prep = False
        self.vortex = None
        self.filtNow = False
        self.cnNow = False
        self.h = None
        self.key = None
        self.base = None
        self.noise_seed = 0
        self.lsfT = 0.0

    # ---------------- allocate / destroy (model.js:1186-1247) ----------------
    def allocate(self, key, spec=None):
        self.destroy()
        pr = spec or GRID_PRESETS[key]
        nx, ny, nz = pr["n"]
        self.key, self.nx, self.ny, self.nz, self.dx = key, nx, ny, nz, float(pr["dx"])
        self.dz = 20000.0 / nz
        self.Lx, self.Ly, self.Lz = nx * self.dx, ny * self.dx, nz * self.dz
        self.params["spongeW"] = min(max(0.08 * min(self.Lx, self.Ly), 5000), 40000)
        self.pad = grid_pad(self.dx)
        self.nxp, self.nyp = nx + 2 * self.pad, ny + 2 * self.pad
        g = _lib.Grid(nx, ny, nz, self.dx, self.dz, self.pad)
        h = C.c_void_p()
        rc = _lib.lib().obl_create(C.byref(g), C.byref(h))
        if rc != 0:
            self.key = None
            msg = _lib.lib().obl_global_error().decode()
            if "memory" in msg and not msg.startswith("not enough GPU memory"):
                raise MemoryError(f"not enough GPU memory for this grid (it needs about {grid_bytes(pr) / 1e9:.1f} GB): "
                                  + msg)
            raise _lib.OblError(msg)
        self.h = h
        self.U.set("pad", self.pad, self.nxp, self.nyp, 0)
        self.U.set("dims", nx, ny, nz, self.dx)
        self.U.set("dzv", self.dz, 1 / self.dx, 1 / self.dz, 0)
        vshape, cshape = (nz + 1, ny + 1, nx + 1, 4), (nz, ny, nx, 4)
        self.V = [Field(self, "V", 0, vshape), Field(self, "V", 1, vshape)]
        self.Q1 = [Field(self, "Q1", 0, cshape), Field(self, "Q1", 1, cshape)]
        self.Q2 = [Field(self, "Q2", 0, cshape), Field(self, "Q2", 1, cshape)]
        self.Vt = Field(self, "Vt", 0, vshape)
        self.Q1t = Field(self, "Q1t", 0, cshape)
        self.Q2t = Field(self, "Q2t", 0, cshape)
        self.aux = Field(self, "aux", 0, (nz, ny, nx, 2))
        self.phT = Field(self, "phT", 0, (nz, ny, nx))
        self.lsfA = Field(self, "lsfA", 0, (ny, nx, 4))
        self.lsfB = Field(self, "lsfB", 0, (ny, nx, 4))
        self.ciMap = Field(self, "ciMap", 0, (ny, nx, 4))
        self.phi = Field(self, "phi", 0, (nz, self.nyp, self.nxp))
        self.rhs = Field(self, "rhs", 0, (nz, self.nyp, self.nxp))
        self.bs = Field(self, "bs", 0, ((nz + 1) * 8,))
        return self

    def destroy(self):
        if self.h is not None:
            _lib.lib().obl_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # ---------------- inputs ----------------
    def setSounding(self, text: str):
        """setSounding (model.js:1276-1288): parse, analyse, base state -> bs + solver density."""
        snd = parse_sharppy(text)
        A = analyze(snd)
        self.analysis, self.snd = A, snd
        bsx = base_state(snd, self.nz, self.dz, (0.0, 0.0), analysis=A)
        self.base = bsx
        self.rhoSfc = float(bsx["B"][2])
        self.set_base_table(bsx["B"])
        return A

    def set_base_table(self, B: np.ndarray):
        B = np.ascontiguousarray(B, dtype=np.float32)
        _lib.check(_lib.lib().obl_set_base_state(self.h, _ptr(B), self.nz), self.h)
        self.rhoSfc = float(B[2])
        self._B = B

    def setBoundaries(self, bnds):
        """setBoundaries (model.js:1371-1381): 148 floats per boundary."""
        self.bounds = list(bnds)[:MAX_BOUNDS]
        f = np.zeros(MAX_BOUNDS * 148, dtype=np.float32)
        for i, b in enumerate(self.bounds):
            o = i * 148
            pts = b["pts"][:MAX_BPTS]
            f[o:o + 20] = [len(pts), b["motion"][0], b["motion"][1], b["halfW"], b["dT"], b["Hmax"], b["slope"],
                           b["Hmin"], b["dq"], b["qRatio"], b["du"], b["dv"], b["wRatio"], 1 if b.get("nudge") else 0,
                           b["dFree"], b["side"], b.get("vary") or 0, b.get("seed") or 0, b.get("varyL") or 60000, 0]
            for k, p in enumerate(pts):
                v = (b.get("vel") and b["vel"][k]) or b["motion"]
                f[o + 20 + 4 * k:o + 24 + 4 * k] = [p[0], p[1], v[0], v[1]]
        self._bd = f
        _lib.check(_lib.lib().obl_set_boundaries(self.h, _ptr(f), len(self.bounds)), self.h)

    def setCI(self, events):
        """setCI (model.js:1383-1389): active towers, 12 floats each, at most 256."""
        act = [e for e in events if e["t1"] > self.time][-MAX_CI:]
        self.ciActive = len(act)
        f = np.zeros(max(1, len(act)) * 12, dtype=np.float32)
        for i, e in enumerate(act):
            f[i * 12:(i + 1) * 12] = [e["x"], e["y"], e["zc"], e["rh"], e["rz"], e["wt"], e["t0"], e["t1"],
                                      e.get("vx") or 0, e.get("vy") or 0, e.get("seed") or i * 1.37,
                                      e.get("ramp") or 150]
        self._ci = f
        _lib.check(_lib.lib().obl_set_ci(self.h, _ptr(f), len(act)), self.h)

    def setPrepField(self, on):
        """setPrepField (model.js:1290-1296): the extra wind field hurricane runs carry."""
        _lib.check(_lib.lib().obl_set_prep_field(self.h, 1 if on else 0), self.h)
        self.prep = bool(on)
        if on:
            self.Vp = Field(self, "Vp", 0, self.V[0].shape)

    def setSurface(self, surf=None, hur=None):
        """setSurface (model.js:1298-1303). surf: nx*ny*4 float32 (land fraction, SST K, 0, 0) per column, or
        None (no sea: outbreak mode); hur: {ck, landMoist, sfcP Pa, tauRad s, radCap K/s, blDepth m, u10f, gust m/s,
        ...} or None."""
        a = None if surf is None else np.ascontiguousarray(surf, np.float32).reshape(-1)
        if a is not None and a.size != self.nx * self.ny * 4:
            raise ValueError(f"setSurface: {self.nx * self.ny * 4} floats expected, got {a.size}")
        if a is None and hur is not None:
            a = np.zeros(self.nx * self.ny * 4, np.float32)
        _lib.check(_lib.lib().obl_set_surface(self.h, None if a is None else _ptr(a), 0 if a is None else a.nbytes), self.h)
        self.sfc = None if a is None else Field(self, "sfc", 0, (self.ny, self.nx, 4))
        self.hur = dict(hur) if hur else None
        self.setPrepField(bool(hur))

    def setVortex(self, v=None):
        """setVortex (model.js:1304-1365): the initial warm-core vortex, v = {x, y (model m), vmax m/s, rmw m, r0 m,
        ztop m, moistR m (0 = no moist core), moistRH, shape ('mid' for a mid-level vortex)} or None. Builds the
        (tangential wind, theta') table on the host and hands it to the library for reset; returns the vortex info
        (table spacing, size, central pressure deficit hPa, least core stability K/km, warm-core peak K)."""
        if not v:
            self.vortex = None
            _lib.check(_lib.lib().obl_set_vortex(self.h, None, 0, 0), self.h)
            return None
        tab, info = vortex_table(self.nz, self.dz, self.dx, self.params["lat"], self.base["B"], v)
        _lib.check(_lib.lib().obl_set_vortex(self.h, _ptr(tab), info["nr"], tab.nbytes), self.h)
        self.vortex = info
        self._vx = tab
        return info

    def baseWind(self, z):
        k = min(max(int(math.floor(z / self.dz)), 0), self.nz - 1)
        B = self.base["B"]
        return [float(B[k * 8 + 6]), float(B[k * 8 + 7])]

    # ---------------- uniforms, reset, step ----------------
    def writeUniforms(self, dt):
        return write_uniforms(self.U, self, dt)

    def reset(self):
        """reset (model.js:1427-1440)."""
        self.time, self.steps, self.parity, self.phiScale = 0.0, 0, 0, 1
        u = self.writeUniforms(1)
        self._u_reset = u
        _lib.check(_lib.lib().obl_reset(self.h, _ptr(u), int(self.noise_seed)), self.h)
        self.lsfT = 0.0

    def add_bubble(self, x, y, z, rh, rz, dth):
        _lib.check(_lib.lib().obl_add_bubble(self.h, x, y, z, rh, rz, dth), self.h)

    def lsf_due(self) -> bool:
        """The reference's LSF schedule: at t = 0 and every >= 60 s of sim time (model.js:1461)."""
        return self.time - self.lsfT >= 60 or self.time == 0

    def encStep(self, enc, dt, judge=False):
        """encStep (model.js:1442-1476): the state ends in V[parity], Q1[0], Q2[0]. judge (with a
        surrogate attached): also run the plain microphysics from the same inputs for the judge."""
        u = self.writeUniforms(dt)
        flags = _lib.STEP_JUDGE if judge else 0
        if self.lsf_due():
            flags |= _lib.STEP_LSF
            self.lsfT = self.time
        if getattr(self, "hur", None) and self.filtNow:
            flags |= _lib.STEP_FILT
        self._u_last = u
        _lib.check(_lib.lib().obl_step(self.h, _ptr(u), self.parity, flags), self.h)
        self.parity = 1 - self.parity
        self.time += dt
        self.steps += 1
        cn_old = CN_OLD if getattr(self, "cnOld", None) is None else self.cnOld
        self.phiScale = 1 / (1 - cn_old) if self.cnNow else 1

    def step(self, dt, judge=False):
        self.encStep(None, dt, judge)

    def sync(self):
        _lib.check(_lib.lib().obl_sync(self.h), self.h)

    def set_solver_mode(self, gemm="bf16x6", thomas="mixed", fp64_cond=64.0):
        _lib.check(_lib.lib().obl_set_solver_mode(self.h, _lib.GEMM[gemm], _lib.THOMAS[thomas], float(fp64_cond)),
                   self.h)

    def profile(self, on=True):
        _lib.check(_lib.lib().obl_profile(self.h, 1 if on else 0), self.h)

    def profile_read(self) -> dict:
        names = C.create_string_buffer(4096)
        ms = np.zeros(64)
        cnt = np.zeros(64, dtype=np.int32)
        n = _lib.lib().obl_profile_read(self.h, names, 4096, _ptr(ms), _ptr(cnt), 64)
        if n < 0:
            _lib.check(n, self.h)
        nm = names.value.decode().split("\n")[:n]
        return {nm[i]: (float(ms[i]), int(cnt[i])) for i in range(n)}

    # ---------------- surrogate teacher sampling (no reference counterpart) ----------------
    def sample_micro(self, capacity, p_active=1.0, p_clear=0.0):
        _lib.check(_lib.lib().obl_dump_enable(self.h, int(capacity), float(p_active), float(p_clear)), self.h)

    def micro_samples(self, copy=True):
        rows = C.c_size_t()
        dev = C.c_void_p()
        _lib.check(_lib.lib().obl_dump_read(self.h, None, 0, C.byref(rows), C.byref(dev)), self.h)
        n = rows.value
        if not copy:
            return n, dev.value
        out = np.empty((n, 24), np.float32)
        if n:
            _lib.check(_lib.lib().obl_dump_read(self.h, _ptr(out), n, C.byref(rows), None), self.h)
        return out

    # ---------------- microphysics surrogate in the step (no reference counterpart) ----------------
    def set_surrogate(self, net=None, capacity=0):
        """Attach a trained Surrogate (or detach with None); see obl_set_surrogate."""
        self._net = net
        _lib.check(_lib.lib().obl_set_surrogate(self.h, None if net is None else net.h, int(capacity)), self.h)

    def surrogate_stats(self, reset=False) -> dict:
        """Surrogate usage and, over judged steps, its error against the plain block from the same
        inputs: relative L2 per channel, column condensate, composite reflectivity, precipitation."""
        st = _lib.SurrStats()
        _lib.check(_lib.lib().obl_surrogate_stats(self.h, C.byref(st), 1 if reset else 0), self.h)
        names = ["qv", "qc", "qi", "qr", "qs", "qg", "Nr", "theta"]
        rel = lambda d, r: float(np.sqrt(d / r)) if r > 0 else (0.0 if d == 0 else float("inf"))  # noqa: E731
        out = {"steps": st.steps, "points": st.points, "fallbacks": st.fallbacks, "overflow": st.overflow,
               "judged": st.judged,
               "rel_l2": {names[i]: rel(st.d2[i], st.r2[i]) for i in range(8)},
               "column_condensate_rel_l2": rel(st.cc_d2, st.cc_r2),
               "column_condensate_max_abs_diff": st.cc_maxdiff, "column_condensate_max": st.cc_max,
               "dbz_rms": float(np.sqrt(st.dbz_d2 / st.dbz_cols)) if st.dbz_cols > 0 else 0.0,
               "dbz_max_abs_diff": st.dbz_maxdiff, "dbz_columns": st.dbz_cols,
               "precip_rel_l2": rel(st.pr_d2, st.pr_r2),
               "precip_total_surrogate": st.pr_sum_s, "precip_total_plain": st.pr_sum_p}
        return out

    def requestMaps(self, enc=None):
        """requestMaps (model.js:1478-1494): enqueue the 2-D maps (model.js:1028-1100) of the current state.
        Returns a callable that waits and returns {"m0".."m3": float32 arrays of nx*ny*4, "t": model time
        at the request}, or None while both readback slots are still pending."""
        u = getattr(self, "_u_last", None)
        if u is None:
            u = self.writeUniforms(1.0)
        slot = C.c_int(-1)
        _lib.check(_lib.lib().obl_maps_request(self.h, _ptr(u), self.parity, C.byref(slot)), self.h)
        if slot.value < 0:
            return None
        t, n, q = self.time, self.nx * self.ny, slot.value

        def collect():
            a = np.empty(16 * n, np.float32)
            _lib.check(_lib.lib().obl_maps_collect(self.h, q, _ptr(a), a.size), self.h)
            return {"m0": a[:4 * n], "m1": a[4 * n:8 * n], "m2": a[8 * n:12 * n], "m3": a[12 * n:], "t": t}

        return collect

    @property
    def Vcur(self):
        return self.V[self.parity]

    def state(self):
        """Current prognostic state as numpy arrays: V[parity], Q1[0], Q2[0]."""
        return self.V[self.parity].download(), self.Q1[0].download(), self.Q2[0].download()


class PoissonDirect:
    """PoissonDirect (solver.js:105-165): exact separable solve on MI355X."""

    def __init__(self, nx, ny, nz, dx, dz):
        self.nx, self.ny, self.nz, self.dx, self.dz = nx, ny, nz, dx, dz
        h = C.c_void_p()
        _lib.check(_lib.lib().obl_poisson_create(nx, ny, nz, float(dx), float(dz), C.byref(h)))
        self.h = h
        self.setDensity(np.ones(nz), np.ones(nz + 1))

    def setDensity(self, rhoC, rhoW):
        rc = np.ascontiguousarray(rhoC, dtype=np.float64)
        rw = np.ascontiguousarray(rhoW, dtype=np.float64)
        if rc.size != self.nz or rw.size != self.nz + 1:
            raise ValueError("setDensity: rhoC needs nz values and rhoW nz + 1")
        _lib.check(_lib.lib().obl_poisson_set_density(self.h, _ptr(rc), _ptr(rw)), self.h, poisson=True)

    def solve(self, rhs, gemm="bf16x6", thomas="mixed", fp64_cond=64.0) -> np.ndarray:
        """Host arrays in/out (x fastest, shape (nz, ny, nx))."""
        r = np.ascontiguousarray(rhs, dtype=np.float32).reshape(self.nz, self.ny, self.nx)
        phi = np.empty_like(r)
        _lib.check(_lib.lib().obl_poisson_solve_host(self.h, _ptr(r), _ptr(phi), _lib.GEMM[gemm], _lib.THOMAS[thomas],
                                                     float(fp64_cond)), self.h, poisson=True)
        return phi

    def enc_device(self, rhs_ptr: int, phi_ptr: int, gemm="bf16x6", thomas="mixed", fp64_cond=64.0, stream=None):
        _lib.check(_lib.lib().obl_poisson_solve(self.h, C.c_void_p(rhs_ptr), C.c_void_p(phi_ptr), _lib.GEMM[gemm],
                                                _lib.THOMAS[thomas], float(fp64_cond), stream), self.h, poisson=True)

    def profile(self, on=True):
        _lib.check(_lib.lib().obl_poisson_profile(self.h, 1 if on else 0), self.h, poisson=True)

    def profile_read(self):
        ms = np.zeros(5)
        _lib.check(_lib.lib().obl_poisson_profile_read(self.h, _ptr(ms)), self.h, poisson=True)
        return dict(zip(["ps.xf", "ps.yf", "ps.tri", "ps.yi", "ps.xi"], ms.tolist()))

    def destroy(self):
        if getattr(self, "h", None) is not None:
            _lib.lib().obl_poisson_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass
