"""A small synthetic radar volume in this build's layout, for the Level II round-trip fixture.

Three sweeps of a VCP 35 radar on a 24 km domain, laid out by the radar host (`plan_radar`): the 0.5 deg
surveillance cut (REF ZDR PHI RHO, 720 radials), the 0.5 deg Doppler cut (REF VEL SW, 720 radials) and the
1.8 deg batch cut (all six moments, 360 radials), 100 gates each, repacked into one buffer. The codes are
smooth deterministic patterns with empty gates (0) and range-folded gates (1), so the fixture compresses
small. The header is the one `RadarSim.collect` returns.
"""
import hashlib
import math
from types import SimpleNamespace

import numpy as np

from outbreaklab_amd.radar import Geo, plan_radar

CUTS = [0, 1, 6]
EPOCH0 = 1714500000000  # 2024-04-30 18:00:00 UTC


def synthetic_volume():
    m = SimpleNamespace(nx=48, ny=48, nz=40, dx=500.0, dz=500.0, Lx=24000.0, Ly=24000.0, rhoSfc=1.1,
                        params={"scheme": "thompson"})
    geo = Geo(35.33, -97.28, m.Lx, m.Ly, 370.0)
    R = plan_radar(m, geo, {"id": "KTLX", "lat": 35.333, "lon": -97.278, "h": 370.0}, {"vcp": 35, "seed": 4})
    off, sweeps, parts = 0, [], []
    az0 = 0.3125
    for k, ci in enumerate(CUTS):
        c = R["cuts"][ci]
        nrad = c["nrad"]
        mom = []
        ri = np.arange(nrad)[:, None].astype(np.float64)
        gi = np.arange(R["ngRef"])[None, :].astype(np.float64)
        for o in c["mom"]:
            ng = o["ngates"]
            g = gi[:, :ng]
            phase = {"REF": 0.0, "VEL": 1.0, "SW": 2.0, "ZDR": 3.0, "PHI": 4.0, "RHO": 5.0}[o["name"]]
            v = 128 + 100 * np.sin(2 * math.pi * ri / nrad * (2 + k) + phase) * np.cos(g / 17.0 + phase)
            code = np.clip(np.rint(v), 2, 255).astype(np.uint8)
            code[(ri.astype(int) % 90 < 20) & (g.astype(int) > ng // 2)] = 0  # no echo
            if o["name"] in ("VEL", "SW"):
                code[(ri.astype(int) % 120 == 7) & (g.astype(int) % 9 == 0)] = 1  # range folded
            parts.append(code.ravel())
            mom.append({"name": o["name"], "ngates": ng, "off": off})
            off += nrad * ng
        t0 = EPOCH0 + round(c["t0"] * 1000)
        sweeps.append({"angle": c["angle"], "superRes": c["superRes"], "kind": c["kind"], "az0": az0, "nrad": nrad,
                       "nyquist": c["nyquist"], "unamb_km": 466 if c["kind"] == "surv" else 300, "start_ms": t0,
                       "dur_ms": round(c["dur"] * 1000), "moments": mom})
    codes = np.concatenate(parts)
    header = {"run": "golden", "icao": "KTLX", "lat": 35.333, "lon": -97.278, "elev": 390.0, "vcp": 35,
              "epoch_ms": EPOCH0, "sweeps": sweeps}
    return header, codes


def digest(codes):
    return hashlib.sha256(np.ascontiguousarray(codes, np.uint8).tobytes()).hexdigest()
