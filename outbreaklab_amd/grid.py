"""Grid helpers of the reference's RegionalModel (model.js:1102-1154): the level count of a grid spacing, custom
grids, the solve's padding, and the device memory a grid takes.

The level table and the presets are the reference's (its grids and keys are part of the API). The memory figure
is this library's own layout (``obl_grid_bytes``), not WebGPU's; WebGPU's 2 GiB-per-field limit
(gridFieldBytes / gridFieldLimit) does not constrain MI355X and is kept only as a reported figure.
"""
from __future__ import annotations

import ctypes as C
import math

# (grid spacing m, levels over 20 km), model.js:1104
LEVELS_BY_DX = [(100, 160), (250, 100), (500, 80), (1000, 64), (1500, 60), (2000, 56), (3000, 52), (4000, 48)]
_LOG_DX = [math.log(d) for d, _ in LEVELS_BY_DX]


def js_round(x: float) -> int:
    """JavaScript Math.round: halves round toward +infinity."""
    return int(math.floor(x + 0.5))


def _nearest_multiple(x: float, m: int) -> int:
    return js_round(x / m) * m


def levels_for(dx: float) -> int:
    """levelsFor (model.js:1105-1112): the level count interpolated linearly in log(dx) between the table's rows,
    to the nearest multiple of 4; spacings beyond the table take its end rows."""
    l = math.log(min(max(dx, LEVELS_BY_DX[0][0]), LEVELS_BY_DX[-1][0]))
    i = next((j for j in range(1, len(_LOG_DX)) if l <= _LOG_DX[j]), None)
    if i is None:
        return LEVELS_BY_DX[-1][1]
    (_, n0), (_, n1) = LEVELS_BY_DX[i - 1], LEVELS_BY_DX[i]
    t = (l - _LOG_DX[i - 1]) / (_LOG_DX[i] - _LOG_DX[i - 1])
    return _nearest_multiple(n0 + (n1 - n0) * t, 4)


def custom_grid(c: dict) -> dict:
    """customGrid (model.js:1114-1120): a spacing (100-4000 m, default 500) and a domain size in km (default 256)
    to cells per side, the nearest multiple of 8 within 64..2040, and levels_for's level count."""
    dx = min(max(float(c.get("dx") or 500), 100.0), 4000.0)
    nx, ny = (min(max(_nearest_multiple(float(c.get(k) or 256) * 1000 / dx, 8), 64), 2040) for k in ("w", "h"))
    nz = levels_for(dx)
    label = int(dx) if dx.is_integer() else dx
    return {"key": f"custom:{label}:{nx}x{ny}x{nz}", "dx": dx, "n": [nx, ny, nz]}


def grid_pad(dx: float) -> int:
    """gridPad (model.js:1122): the solve's padding per side, about 10 km of cells to the nearest multiple of 8,
    at least 8 and at most 64."""
    return min(max(_nearest_multiple(10000 / dx, 8), 8), 64)


def grid_bytes(g: dict) -> float:
    """The device bytes this library allocates for the grid g = {dx, n: [nx, ny, nz]} (the counterpart of
    gridBytes, model.js:1125-1128, for this layout; computed by the library)."""
    from . import _lib

    nx, ny, nz = g["n"]
    gr = _lib.Grid(nx, ny, nz, float(g["dx"]), 20000.0 / nz, grid_pad(g["dx"]))
    return float(_lib.lib().obl_grid_bytes(C.byref(gr)))


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
