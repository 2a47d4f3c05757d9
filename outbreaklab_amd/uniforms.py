"""The step's uniform block: MODEL_FIELDS (model.js:23-37) packed like UBlock (util.js:83-98).

19 vec4f members, 76 float32 values, in the reference's order. ``write_uniforms``
restates RegionalModel.writeUniforms (model.js:1391-1425) for outbreak mode.
"""
from __future__ import annotations

import math

import numpy as np

MODEL_FIELDS = ["dims", "dzv", "step", "sponge", "turb", "micro", "misc", "cor", "ci", "pad", "micro2",
                "hur", "hur2", "hur3", "hur4", "hur5", "vtx", "vtx2", "cn"]
DEG = math.pi / 180.0


class UBlock:
    """UBlock: every member is a vec4f; ``set`` writes four values (missing ones as 0)."""

    def __init__(self, fields=MODEL_FIELDS):
        self.off = {n: 4 * i for i, n in enumerate(fields)}
        self.f = np.zeros(4 * len(fields), dtype=np.float32)
        self.size = self.f.nbytes

    def set(self, name, a=0.0, b=0.0, c=0.0, d=0.0):
        if name not in self.off:
            raise KeyError("UBlock: no field " + name)
        o = self.off[name]
        self.f[o:o + 4] = [0.0 if v is None else v for v in (a, b, c, d)]
        return self


def default_params() -> dict:
    """RegionalModel.params defaults (model.js:1176-1180)."""
    return {
        "cs": 0.18, "nu0": 1.0, "drag": 1, "z0": 0.1, "spongeW": 36000, "spongeTau": 100, "dampZ": 15000,
        "dampTau": 150, "auto": 1e-3, "qc0": 2e-3, "freeze": 1.0, "noise": 0.25, "mc": 1.0, "vcap": 110,
        "evap": 0.22, "sortQ": 5e-4, "ciRate": 0.5, "nudgeTau": 5400, "lat": 35, "moistMix": 0, "mcFloor": 0.25,
    }


# outbreak-mode overrides of startRun (app.js:1613)
OUTBREAK_OVERRIDES = {"qc0": 2e-3, "dampZ": 15000, "dampTau": 150, "mcFloor": 0.25, "evap": 0.22}


CN_OLD = 0.4  # hurricane mode: the share of the last step's pressure force applied before advection (model.js:39)


def write_uniforms(U: UBlock, m, dt: float) -> np.ndarray:
    """writeUniforms (model.js:1391-1425).

    ``m`` needs: params, time, rhoSfc, dz, dx, bounds, ciActive, phiScale; for hurricane mode also hur (the
    setSurface settings or None), steps, Lx, Ly, prep (the prep field exists), vortex (setVortex's result or
    None), and optionally cnOld / traj2. Sets m.filtNow and m.cnNow as the reference does.
    """
    P = m.params
    U.set("step", dt, m.time, m.rhoSfc or 1.15, 9.81)
    U.set("sponge", P["spongeW"], P["spongeTau"], P["dampZ"], P["dampTau"])
    U.set("turb", P["cs"], P["nu0"], P["evap"], 1 if P["drag"] else 0)
    cd = (0.41 / math.log(0.5 * m.dz / P["z0"] + 1)) ** 2
    U.set("micro", P["auto"], P["qc0"], P["freeze"], cd)
    U.set("misc", P["vcap"], P["mc"], P["noise"], P["ciRate"])
    U.set("cor", 2 * 7.292e-5 * math.sin(P["lat"] * DEG), P["nudgeTau"], len(m.bounds), P.get("moistMix", 0) or 0)
    U.set("ci", m.ciActive or 0, 0, 0, 0)
    U.set("micro2", 0, P.get("sortQ", 5e-4), 1 if P.get("scheme") == "thompson" else 0, P.get("evapT", 0.3))
    H = getattr(m, "hur", None)
    if H:
        U.set("hur", 1, H["ck"], H["landMoist"], H["sfcP"])
        U.set("hur2", H["tauRad"], H["radCap"], H["blDepth"], H["u10f"])
        U.set("hur3", (0.41 / math.log(0.5 * m.dz / 0.3 + 1)) ** 2, H["gust"],
              (H.get("thermals") or 0) * (1 if H.get("thermalScale") is None else H["thermalScale"]),
              1 if H.get("maritime") else 0)
        filt4 = 0.003 if H.get("filter4") is None else H["filter4"]
        m.filtNow = m.steps % 4 == 3 and filt4 > 0
        U.set("hur5", 2.4e-3 if H.get("cdCap") is None else H["cdCap"], 1 if H.get("dissHeat") else 0, 4 * filt4, 0)
        C = H.get("centre") or [m.Lx / 2, m.Ly / 2]
        U.set("hur4", C[0], C[1], H.get("thermalR") or 200000, H.get("envR") or 400000)
    else:
        m.filtNow = False
        U.set("hur", 0, 0, 0, 0)
        U.set("hur2", 1, 0, 1, 0.8)
        U.set("hur3", 0, 1, 0, 0)
        U.set("hur4", 0, 0, 1, 1e9)
        U.set("hur5", 2.4e-3, 0, 0, 0)
    cn_old = CN_OLD if getattr(m, "cnOld", None) is None else m.cnOld
    cn = bool(H) and bool(getattr(m, "prep", False)) and getattr(m, "steps", 0) > 0 and cn_old > 0
    m.cnNow = cn
    traj2 = True if getattr(m, "traj2", None) is None else m.traj2
    U.set("cn", cn_old if cn else 0, m.phiScale or 1, 1 / (1 - cn_old) if cn else 1, 1 if H and traj2 else 0)
    X = getattr(m, "vortex", None)
    if X:
        U.set("vtx", X["x"], X["y"], X["dr"], X["nr"])
        U.set("vtx2", X["moistR"], X["moistRH"], 1, 0)
    else:
        U.set("vtx", 0, 0, 1, 0)
        U.set("vtx2", 1, 0, 0, 0)
    U.set("dzv", m.dz, 1 / m.dx, 1 / m.dz, P.get("mcFloor", 0.25))
    return U.f.copy()
