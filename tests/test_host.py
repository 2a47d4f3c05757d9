"""Host-side logic: grid helpers, uniform packing, the sounding / base state against the
reference's own sounding.js outputs, and the C ABI's exported symbols (no GPU calls)."""
import ctypes
import json
import math
import os
import re
import subprocess

import numpy as np
import pytest

from outbreaklab_amd import grid, sounding, uniforms
from _cases import host_model, uniforms as pack

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden", "sounding_golden.json")


def test_grid_helpers_closed_forms():
    # gridPad (model.js:1122): round(10000/dx/8)*8 clamped to 8..64
    assert [grid.grid_pad(d) for d in (2000, 1000, 500, 250, 100)] == [8, 8, 24, 40, 64]
    # levelsFor (model.js:1105-1112): the table rows themselves, and the rounding to multiples of 4
    assert [grid.levels_for(d) for d in (100, 250, 500, 1000, 1500, 2000, 3000, 4000)] == [160, 100, 80, 64, 60, 56, 52, 48]
    assert grid.levels_for(750) % 4 == 0 and 64 <= grid.levels_for(750) <= 80
    g = grid.custom_grid({"dx": 500, "w": 512, "h": 512})
    assert g["n"] == [1024, 1024, 80] and g["key"] == "custom:500:1024x1024x80"
    assert grid.custom_grid({"dx": 2000, "w": 5000, "h": 10})["n"][:2] == [2040, 64]
    # the library's allocation at 1024^2 x 80: 3 wind copies, 6 moisture copies, aux, phT and the padded solve
    assert 14e9 < grid.grid_bytes({"dx": 500, "n": [1024, 1024, 80]}) < 18e9


def test_ublock_layout():
    U = uniforms.UBlock()
    assert U.f.size == 76 and U.size == 304
    assert U.off["dims"] == 0 and U.off["pad"] == 36 and U.off["micro2"] == 40 and U.off["cn"] == 72
    U.set("turb", 1, 2)
    assert list(U.f[16:20]) == [1, 2, 0, 0]
    with pytest.raises(KeyError):
        U.set("nope", 1)


def test_write_uniforms_values():
    m = host_model(128, 128, 40, 2000, scheme="thompson")
    m.time = 42.0
    u = pack(m, 6.0)
    assert u[0:4].tolist() == [128, 128, 40, 2000]
    assert u[8] == np.float32(6.0) and u[9] == np.float32(42.0) and u[11] == np.float32(9.81)
    cd = (0.41 / math.log(0.5 * 500.0 / 0.1 + 1)) ** 2
    assert u[23] == np.float32(cd)
    assert u[28] == np.float32(2 * 7.292e-5 * math.sin(35.9 * math.pi / 180))
    assert u[42] == 1.0 and u[41] == np.float32(5e-4) and u[43] == np.float32(0.3)
    assert u[12] == np.float32(min(max(0.08 * 256000, 5000), 40000))  # spongeW
    assert u[36:39].tolist() == [8, 144, 144]
    assert u[48:52].tolist() == [1, 0, 1, np.float32(0.8)] and u[59] == np.float32(1e9)
    assert u[4] == 500.0 and u[7] == np.float32(0.25)


@pytest.mark.parametrize("idx", range(6))
def test_base_state_matches_reference_js(idx):
    """sounding.js evaluated by Node on its samples (tests/golden/make_sounding_golden.js)."""
    case = json.load(open(GOLD))[idx]
    snd = sounding.parse_sharppy(case["text"])
    A = sounding.analyze(snd)
    for k, v in case["P"].items():
        assert abs(A["P"][k] - v) <= 1e-8 * max(1.0, abs(v)), k
    for z, w in zip([0, 0.5, 1, 3, 6, 10], case["wind"]):
        assert np.allclose(A["windAt"](z), w, rtol=0, atol=1e-12)
    for key, ref in case["base"].items():
        nz, dz = key.split("x")
        B = sounding.base_state(snd, int(nz), float(dz), analysis=A)["B"]
        assert np.array_equal(B, np.array(ref, np.float32)), key


def test_base_state_hydrostatic():
    """pi at cell centres follows d(pi)/dz = -g/(cp theta_v) between faces, and rho = p/(Rd T (1+0.61 qv))."""
    snd = sounding.parse_sharppy(sounding.wk_sounding_text())
    nz, dz = 80, 250.0
    B = sounding.base_state(snd, nz, dz)["B"].astype(np.float64).reshape(nz + 1, 8)
    th, qv, rho, pi, p = B[:nz, 0], B[:nz, 1], B[:nz, 2], B[:nz, 4], B[:nz, 5]
    assert np.allclose(p, 1e5 * pi ** (1005.7 / 287.04), rtol=1e-6)
    assert np.allclose(rho, p / (287.04 * th * pi * (1 + 0.61 * qv)), rtol=2e-6)
    # the centre-to-centre pi difference is the mean of the two half-layer drops
    dpi = pi[1:] - pi[:-1]
    expect = -0.5 * 9.80665 * dz * (1 / (1005.7 * th[1:] * (1 + 0.61 * qv[1:])) + 1 / (1005.7 * th[:-1] * (1 + 0.61 * qv[:-1])))
    assert np.allclose(dpi, expect, rtol=5e-3)
    assert B[nz, 3] == pytest.approx(0.94 * B[nz - 1, 2], rel=1e-6)


def test_parse_errors():
    with pytest.raises(sounding.SoundingError, match="at least 4 levels"):
        sounding.parse_sharppy("%RAW%\n1000, 0, 20, 10, 0, 0\n%END%")


def test_wk_sounding_closed_form():
    z = np.array([0.0, 6000.0, 12000.0, 15000.0])
    th, qv, p, T = sounding.wk_profile(z)
    assert th[0] == pytest.approx(300.0) and th[2] == pytest.approx(343.0, rel=1e-9)
    assert th[1] == pytest.approx(300 + 43 * 0.5 ** 1.25, rel=1e-9)
    assert th[3] == pytest.approx(343 * math.exp(9.81 * 3000 / (1004.5 * 213)), rel=1e-6)
    assert qv[0] == pytest.approx(0.014, rel=1e-9)
    u, v = sounding.hodograph(np.array([0.0, 6000.0]), shear6=30.0)
    assert math.hypot(u[1] - u[0], v[1] - v[0]) == pytest.approx(30.0, rel=1e-9)


def _header_symbols():
    txt = open(os.path.join(ROOT, "include", "outbreak_hip.h")).read()
    return sorted(set(re.findall(r"\b(obl_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from outbreaklab_amd import _lib

    so = _lib.LIB_PATH
    if not os.path.exists(so):
        pytest.skip("liboutbreak_hip.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", so], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (obl_\w+)", out))
    declared = _header_symbols()
    missing = [s for s in declared if s not in exported]
    assert not missing, missing
    # loading it (no device calls) and the ctypes signature table cover every declaration
    ctypes.CDLL(so)
    assert set(_lib.symbols()) == set(declared)


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N > 1 exits with status 2 and says so, before touching a device."""
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
