"""The workloads the step is measured and checked on, as concrete synthetic inputs.

The reference ships no such cases (it is a browser app driven by user
settings); these builders define them with the reference's own API:
``allocate('custom', spec)``, ``setSounding(text)``, params, ``reset()``, then
warm bubbles (a build-defined write, see ``RegionalModel.add_bubble``).

  plain_128      128 x 128 x 40 at 2 km (dz 500 m, pad 8), Lin, one 2 K bubble, dt 6 s
  solve_1024     the pressure solve alone, 1024 x 1024 x 80 (dx 500 m, dz 250 m)
  plain_1024     1024 x 1024 x 80 at 500 m (pad 24), Thompson, 8 bubbles, dt 2 s
  combined_1024  1024 x 1024 x 100 at 250 m (pad 40), one cold-pool boundary, dt 1 s
"""
from __future__ import annotations

import functools
import math

import numpy as np

from .sounding import analyze, parse_sharppy, wk_sounding_text


@functools.lru_cache(maxsize=16)
def wk_text_for_cape(target_cape: float, shear6: float, th0: float = 300.0, thtr: float = 343.0) -> str:
    """WK sounding whose surface-based CAPE (the reference's liftParcel) is ``target_cape``,
    found by bisection on the boundary-layer mixing ratio qv0."""
    lo, hi = 0.008, 0.024
    for _ in range(22):
        mid = 0.5 * (lo + hi)
        cape = analyze(parse_sharppy(wk_sounding_text(th0=th0, thtr=thtr, qv0=mid, shear6=shear6)))["P"]["sbcape"]
        if cape < target_cape:
            lo = mid
        else:
            hi = mid
    return wk_sounding_text(th0=th0, thtr=thtr, qv0=0.5 * (lo + hi), shear6=shear6)


def bubble_positions(n: int, Lx: float, Ly: float, margin: float, min_sep: float, seed: int):
    rng = np.random.default_rng(seed)
    pts = []
    tries = 0
    while len(pts) < n:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place the bubbles")
        x, y = rng.uniform(margin, Lx - margin), rng.uniform(margin, Ly - margin)
        if all(math.hypot(x - a, y - b) >= min_sep for a, b in pts):
            pts.append((x, y))
    return pts


def setup_plain_128(model, seed: int = 0, scheme: str = "lin"):
    """128 x 128 x 40, 2 km: WK (300 K, 343 K at 12 km, 14 g/kg), quarter-circle hodograph to 30 m/s,
    one 2 K bubble (10 km x 1.5 km radius, centre 1.5 km AGL), +-0.1 K noise, dt 6 s."""
    model.allocate("custom", {"key": "custom:2000:128x128x40", "dx": 2000, "n": [128, 128, 40]})
    model.params["scheme"] = scheme
    model.params["noise"] = 0.1
    model.params["lat"] = 35.9
    model.noise_seed = seed
    model.setSounding(wk_sounding_text(th0=300.0, thtr=343.0, qv0=0.014, shear6=30.0, hodo_r=5.0))
    model.setBoundaries([])
    model.setCI([])
    model.setSurface(None, None)
    model.reset()
    model.add_bubble(0.5 * model.Lx, 0.5 * model.Ly, 1500.0, 10000.0, 1500.0, 2.0)
    return 6.0


def setup_plain_1024(model, seed: int = 1, scheme: str = "thompson", n=(1024, 1024, 80), cape: float = 3000.0):
    """1024 x 1024 x 80 at 500 m: WK at 3000 J/kg SBCAPE with 25 m/s 0-6 km shear, Thompson,
    8 seeded 2 K bubbles >= 35 km apart and >= spongeW + 10 km from the edges, dt 2 s."""
    nx, ny, nz = n
    model.allocate("custom", {"key": f"custom:500:{nx}x{ny}x{nz}", "dx": 500, "n": [nx, ny, nz]})
    model.params["scheme"] = scheme
    model.params["noise"] = 0.1
    model.params["lat"] = 35.9
    model.noise_seed = seed
    model.setSounding(wk_text_for_cape(cape, 25.0))
    model.setBoundaries([])
    model.setCI([])
    model.setSurface(None, None)
    model.reset()
    margin = model.params["spongeW"] + 10000.0
    nb = 8 if min(model.Lx, model.Ly) > 2 * margin + 70000 else 1
    for x, y in bubble_positions(nb, model.Lx, model.Ly, margin, 35000.0, seed):
        model.add_bubble(x, y, 1500.0, 10000.0, 1500.0, 2.0)
    return 2.0


def setup_outbreak_preset(model, key: str, seed: int = 1, scheme: str = "thompson", cape: float = 3000.0):
    """The 1024 x 1024 x 80 workload's storm environment on one of the reference's outbreak grids (model.js:1143-1149:
    quick, standard, large, fine, m500, m250, m100): WK at `cape` J/kg SBCAPE, 25 m/s 0-6 km shear, 2 K bubbles
    (8 where the domain has room, else one at the centre), dt 6 dx_km s as app.js:1666 at time step x1."""
    model.allocate(key)
    model.params["scheme"] = scheme
    model.params["noise"] = 0.1
    model.params["lat"] = 35.9
    model.noise_seed = seed
    model.setSounding(wk_text_for_cape(cape, 25.0))
    model.setBoundaries([])
    model.setCI([])
    model.setSurface(None, None)
    model.reset()
    margin = model.params["spongeW"] + 10000.0
    if min(model.Lx, model.Ly) > 2 * margin + 70000:
        pts = bubble_positions(8, model.Lx, model.Ly, margin, 35000.0, seed)
    else:
        pts = [(0.5 * model.Lx, 0.5 * model.Ly)]
    for x, y in pts:
        model.add_bubble(x, y, 1500.0, 10000.0, 1500.0, 2.0)
    return 6.0 * model.dx / 1000.0


def setup_combined_1024(model, seed: int = 1, n=(1024, 1024, 100), cape: float = 4000.0):
    """1024 x 1024 x 100 at 250 m (dz 200 m, pad 40): WK at 4000 J/kg SBCAPE with 25 m/s 0-6 km shear,
    Thompson, one seeded outflow-type boundary (dT -6 K, 1.5 km deep, along-front variation 0.8 with
    50 km hot spots) on a north-south line at 30 % of the domain moving east at 6 m/s, dt 1 s.
    The boundary record follows bndModel (app.js:527-541) for an outflow boundary moving east."""
    nx, ny, nz = n
    model.allocate("custom", {"key": f"custom:250:{nx}x{ny}x{nz}", "dx": 250, "n": [nx, ny, nz]})
    model.params["scheme"] = "thompson"
    model.params["noise"] = 0.1
    model.params["lat"] = 35.9
    model.noise_seed = seed
    model.setSounding(wk_text_for_cape(cape, 25.0))
    x0 = 0.3 * model.Lx
    pts = [[x0, model.Ly * (0.08 + 0.84 * i / 7.0)] for i in range(8)]
    ws = model.baseWind(150.0)
    # cold-pool wind 8 m/s from the west (windDir = dir + 180 for an outflow moving toward dir = 90)
    bnd = {"pts": pts, "motion": [6.0, 0.0], "halfW": 4000.0, "dT": -6.0, "Hmax": 1500.0, "slope": 0.05,
           "Hmin": 300.0, "dq": 0.1, "qRatio": 1.0, "du": 8.0 - ws[0], "dv": 0.0 - ws[1], "wRatio": 1.0,
           "nudge": False, "dFree": 25000.0, "side": 1, "vary": 0.8, "seed": seed * 0.731 + 11.3, "varyL": 50000.0}
    model.setBoundaries([bnd])
    model.setCI([])
    model.setSurface(None, None)
    model.reset()
    return 1.0


def solve_rhs_1024(nx=1024, ny=1024, nz=80, seed=1, rms=1e-5):
    """Seeded zero-mean Gaussian divergence RHS with a k^-5/3 horizontal spectrum:
    in DST space g(l, m, k) kappa^(-4/3), kappa = sqrt((l+1)^2 + (m+1)^2), g ~ N(0, 1);
    inverse DST in FP64, per-level mean removed, scaled to the given RMS (s^-2)."""
    from scipy.fft import idstn

    rng = np.random.default_rng(seed)
    l = np.arange(nx) + 1.0
    m = np.arange(ny) + 1.0
    kap = np.sqrt(l[None, :] ** 2 + m[:, None] ** 2)
    spec = rng.standard_normal((nz, ny, nx)) * kap[None] ** (-4.0 / 3.0)
    f = idstn(spec, type=2, norm="ortho", axes=(1, 2))
    f -= f.mean(axis=(1, 2), keepdims=True)
    f *= rms / np.sqrt(np.mean(f * f))
    return f


# the hurricane-mode settings app.js passes to setSurface (app.js:1631), for a storm at the domain centre
HURRICANE_SETTINGS = {"ck": 1.3e-3, "cdCap": 2.0e-3, "dissHeat": True, "thermalFade": True, "landMoist": 0.3,
                      "tauRad": 43200, "radCap": 2 / 86400, "blDepth": 1000, "u10f": 0.8, "gust": 1.5,
                      "thermals": 0.4 / 300, "maritime": True, "thermalR": 200000}
# hurricane-mode parameter overrides of startRun (app.js:1613)
HURRICANE_OVERRIDES = {"qc0": 1e-3, "dampZ": 16500, "dampTau": 300, "mcFloor": 1.0, "evap": 0.10}


def setup_hurricane_128(model, seed: int = 0, scheme: str = "lin", land_from_km: float = 200.0, text=None):
    """128 x 128 x 40, 2 km, hurricane mode: a 302 K sea (land east of ``land_from_km``), the app's hurricane
    settings (environment hold from 100 km), a 30 m/s warm-core vortex (rmw 30 km, r0 120 km, moist core
    60 km) at the centre, one storm-start tower, lat 20, dt 6 s. ``text``: another SHARPpy sounding instead of
    the moist, weakly sheared WK one."""
    model.allocate("custom", {"key": "custom:2000:128x128x40", "dx": 2000, "n": [128, 128, 40]})
    model.params.update(HURRICANE_OVERRIDES)
    model.params["scheme"] = scheme
    model.params["noise"] = 0.1
    model.params["lat"] = 20.0
    model.noise_seed = seed
    model.setSounding(text or wk_sounding_text(th0=300.0, thtr=343.0, qv0=0.016, shear6=5.0, hodo_r=2.0))
    model.setBoundaries([])
    c = [0.5 * model.Lx, 0.5 * model.Ly]
    model.setCI([{"x": c[0] + 40000, "y": c[1], "zc": 2000, "rh": 10000, "rz": 1500, "wt": 8, "t0": 0, "t1": 1800,
                  "ramp": 60}])
    surf = np.zeros((model.ny, model.nx, 4), np.float32)
    surf[..., 1] = 302.0
    xs = (np.arange(model.nx) + 0.5) * model.dx
    surf[:, xs > land_from_km * 1000, 0] = 1.0
    hur = dict(HURRICANE_SETTINGS, sfcP=101000.0, centre=c, envR=100000)
    model.setSurface(surf, hur)
    model.setVortex({"x": c[0], "y": c[1], "vmax": 30.0, "rmw": 30000.0, "r0": 120000.0, "ztop": 16000.0,
                     "moistR": 60000.0, "moistRH": 0.9})
    model.reset()
    return 6.0


def setup_hurricane_preset(model, key: str = "h2", seed: int = 0, sst: float = 302.0):
    """A hurricane on one of the reference's hurricane grids (h2: 640 x 640 x 56 at 2 km; model.js:1150-1153):
    an open 302 K sea, the app's hurricane settings with the environment held beyond 300 km, a 35 m/s balanced
    warm-core vortex (rmw 40 km, r0 250 km, moist core 100 km) at the centre on a moist, weakly sheared WK
    sounding, lat 20, dt 6 dx_km s (app.js:1666)."""
    model.allocate(key)
    model.params.update(HURRICANE_OVERRIDES)
    model.params["scheme"] = "lin"
    model.params["noise"] = 0.1
    model.params["lat"] = 20.0
    model.noise_seed = seed
    model.setSounding(wk_sounding_text(th0=300.0, thtr=343.0, qv0=0.016, shear6=5.0, hodo_r=2.0))
    model.setBoundaries([])
    model.setCI([])
    c = [0.5 * model.Lx, 0.5 * model.Ly]
    surf = np.zeros((model.ny, model.nx, 4), np.float32)
    surf[..., 1] = sst
    model.setSurface(surf, dict(HURRICANE_SETTINGS, sfcP=101000.0, centre=c, envR=300000))
    model.setVortex({"x": c[0], "y": c[1], "vmax": 35.0, "rmw": 40000.0, "r0": 250000.0, "ztop": 16000.0,
                     "moistR": 100000.0, "moistRH": 0.9})
    model.reset()
    return 6.0 * model.dx / 1000.0
