"""Grid helpers of the reference's RegionalModel (model.js:1102-1154).

``levels_for``, ``custom_grid``, ``grid_pad``, ``grid_bytes`` and
``GRID_PRESETS`` follow the reference's functions of the same names; only
the WebGPU 2 GiB-per-field limit (gridFieldBytes/gridFieldLimit) is not a
constraint on MI355X and is kept only as a reported figure.
"""
from __future__ import annotations

import math

# levels by grid spacing (20 km deep), model.js:1104
LEVELS_BY_DX = [(100, 160), (250, 100), (500, 80), (1000, 64), (1500, 60), (2000, 56), (3000, 52), (4000, 48)]


def _clamp(v, lo, hi):
    return min(max(v, lo), hi)


def js_round(x: float) -> int:
    """JavaScript Math.round: halves round toward +infinity."""
    return int(math.floor(x + 0.5))


def levels_for(dx: float) -> int:
    """levelsFor (model.js:1105-1112): log-linear in dx between table rows, rounded to a multiple of 4."""
    T = LEVELS_BY_DX
    l = math.log(_clamp(dx, T[0][0], T[-1][0]))
    for i in range(len(T) - 1):
        a, b = math.log(T[i][0]), math.log(T[i + 1][0])
        if l <= b:
            t = (l - a) / (b - a)
            return js_round((T[i][1] + (T[i + 1][1] - T[i][1]) * t) / 4) * 4
    return T[-1][1]


def custom_grid(c: dict) -> dict:
    """customGrid (model.js:1114-1120): spacing and domain size -> cells per side (multiple of 8, 64..2040)."""
    dx = _clamp(float(c.get("dx") or 500), 100, 4000)

    def side(km):
        return int(_clamp(js_round((float(km or 256)) * 1000 / dx / 8) * 8, 64, 2040))

    nx, ny, nz = side(c.get("w")), side(c.get("h")), levels_for(dx)
    dxi = int(dx) if float(dx).is_integer() else dx
    return {"key": f"custom:{dxi}:{nx}x{ny}x{nz}", "dx": dx, "n": [nx, ny, nz]}


def grid_pad(dx: float) -> int:
    """gridPad (model.js:1122): ~10 km of solve padding per side, a multiple of 8, 8..64 cells."""
    return int(_clamp(js_round(10000 / dx / 8) * 8, 8, 64))


def grid_bytes(g: dict) -> float:
    """gridBytes (model.js:1125-1128): the reference's memory estimate of a grid."""
    nx, ny, nz = g["n"]
    p = grid_pad(g["dx"])
    nxp, nyp = nx + 2 * p, ny + 2 * p
    return 1.05 * ((nx + 1) * (ny + 1) * (nz + 1) * 48 + nx * ny * nz * 108 + nxp * nyp * nz * 20 + nx * ny * 192
                   + 8 * (nxp * nxp + nyp * nyp))


def grid_field_bytes(g: dict) -> int:
    """gridFieldBytes (model.js:1133): the largest field's size as WebGPU sets it up."""
    nx, ny, nz = g["n"]
    return 16 * math.ceil((nx + 1) / 16) * 16 * (ny + 1) * (nz + 1)


GRID_PRESETS = {
    "quick": {"dx": 2000, "n": [320, 320, 56]},
    "standard": {"dx": 1500, "n": [448, 448, 60]},
    "large": {"dx": 1500, "n": [640, 640, 60]},
    "fine": {"dx": 1000, "n": [640, 640, 64]},
    "m500": {"dx": 500, "n": [512, 512, 80]},
    "m250": {"dx": 250, "n": [512, 512, 100]},
    "m100": {"dx": 100, "n": [512, 512, 160]},
    "h2": {"dx": 2000, "n": [640, 640, 56]},
    "h3": {"dx": 3000, "n": [640, 640, 52]},
    "h15": {"dx": 1500, "n": [960, 960, 60]},
}
