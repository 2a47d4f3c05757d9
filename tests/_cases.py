"""CPU-side construction of the workloads' uniform blocks and base states (no GPU needed)."""
import types

import numpy as np

from outbreaklab_amd.grid import grid_pad
from outbreaklab_amd.sounding import analyze, base_state, parse_sharppy, wk_sounding_text
from outbreaklab_amd.uniforms import OUTBREAK_OVERRIDES, UBlock, default_params, write_uniforms


def host_model(nx, ny, nz, dx, scheme="lin", noise=0.1, text=None):
    m = types.SimpleNamespace()
    m.nx, m.ny, m.nz, m.dx = nx, ny, nz, float(dx)
    m.dz = 20000.0 / nz
    m.Lx, m.Ly = nx * m.dx, ny * m.dx
    m.params = default_params()
    m.params.update(OUTBREAK_OVERRIDES)
    m.params["spongeW"] = min(max(0.08 * min(m.Lx, m.Ly), 5000), 40000)
    m.params["scheme"] = scheme
    m.params["noise"] = noise
    m.params["lat"] = 35.9
    m.pad = grid_pad(m.dx)
    m.nxp, m.nyp = nx + 2 * m.pad, ny + 2 * m.pad
    m.time, m.bounds, m.ciActive, m.phiScale = 0.0, [], 0, 1
    snd = parse_sharppy(text or wk_sounding_text())
    m.B = base_state(snd, nz, m.dz, analysis=analyze(snd, parcels=False))["B"]
    m.rhoSfc = float(m.B[2])
    m.U = UBlock()
    m.U.set("pad", m.pad, m.nxp, m.nyp, 0)
    m.U.set("dims", nx, ny, nz, m.dx)
    return m


def uniforms(m, dt):
    return write_uniforms(m.U, m, dt)
