"""The Node facade on the GPU: 20 steps of the 128 x 128 x 40, 2 km Thompson workload driven through
outbreak.js and the N-API addon equal, bit for bit, the same steps driven through the Python host
(same library, same uniform blocks, deterministic kernels), and so do the requestMaps maps and a seeded
radar volume with its header."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "outbreaklab_amd", "js", "obl_napi.node")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                 reason="node or the built addon is missing")]


GOLD_WK = os.path.join(ROOT, "tests", "golden", "wk_sounding.txt")


def _node_inputs(tmp_path):
    """The WK base state (40 levels of 500 m) and the radar site, 40 km west / 30 km south of the centre of the
    128 x 128 x 40 grid at 2 km, in the Python host's projection: what the application hands the Node facade."""
    from outbreaklab_amd.radar import Geo
    from outbreaklab_amd.sounding import base_state, parse_sharppy

    B = base_state(parse_sharppy(open(GOLD_WK).read()), 40, 500.0)["B"]
    geo = Geo(35.9, -97.0, 256000.0, 256000.0, 0.0)
    lat, lon = geo.fromModel(0.5 * 256000 - 40e3, 0.5 * 256000 - 30e3)
    sx, sy = geo.toModel(lat, lon)
    p = tmp_path / "node_in.json"
    p.write_text(json.dumps({"B": [float(x) for x in B], "site": {"lat": lat, "lon": lon, "sx": sx, "sy": sy}}))
    return str(p)


def test_node_facade_steps_equal_python_host(tmp_path):
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_128

    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js_gpu_run.js"), str(tmp_path), _node_inputs(tmp_path)],
                       capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    raw = np.fromfile(info["out"], np.float32)
    nv, nq = info["nV"], info["nQ"]
    Vj, Q1j, Q2j = raw[:nv], raw[nv:nv + nq], raw[nv + nq:]
    m = RegionalModel()
    dt = setup_plain_128(m, scheme="thompson")
    for _ in range(20):
        m.step(dt)
    V, Q1, Q2 = m.state()
    maps = m.requestMaps()()
    from outbreaklab_amd.radar import Geo, RadarSim

    geo = Geo(35.9, -97.0, m.Lx, m.Ly, 0.0)
    sim = RadarSim(m, geo)
    sim.setTornadoes([{"x": 0.5 * m.Lx, "y": 0.5 * m.Ly, "R": 150.0, "v": 60.0, "debrisTop": 1200.0,
                       "debrisDbz": 55.0, "debris": 1.0}])
    lat, lon = geo.fromModel(0.5 * m.Lx - 40e3, 0.5 * m.Ly - 30e3)
    R = sim.addRadar({"id": "KTST", "lat": lat, "lon": lon, "h": 0.0}, {"vcp": 35, "seed": 11})
    done = sim.encDue(None, R["phase"] + R["vcp"]["dur"] - 1)
    vol = sim.collect(done[0], "js", 0.0)
    sim.destroy()
    m.destroy()
    assert info["parity"] == m.parity and info["time"] == m.time
    assert np.array_equal(Vj, V.ravel()) and np.array_equal(Q1j, Q1.ravel()) and np.array_equal(Q2j, Q2.ravel())
    # requestMaps through the Node facade equals the Python host's
    mj = np.fromfile(info["mout"], np.float32)
    assert info["mapsT"] == maps["t"]
    assert np.array_equal(mj, np.concatenate([maps[f"m{q}"] for q in range(4)]))
    # a seeded radar volume (rng.js draws, RADAR_FIELDS blocks, kernels) through either host
    assert info["volumes"] == len(done) == 1
    assert np.array_equal(np.fromfile(info["rout"], np.uint8), vol["bytes"])
    assert info["header"] == vol["header"]


def test_node_hurricane_run_equals_python_host(tmp_path):
    """Hurricane mode through the Node host (setSurface, setVortex and its table, the filter schedule, PREP and
    the phi scale) equals the Python host's run of the same workload."""
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_hurricane_128
    from _util import rel_l2

    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js_gpu_run.js"), str(tmp_path), _node_inputs(tmp_path)],
                       capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])["hinfo"]
    m = RegionalModel()
    dt = setup_hurricane_128(m, text=open(os.path.join(ROOT, "tests", "golden", "wk_sounding.txt")).read())
    for _ in range(9):
        m.step(dt)
    V, Q1, Q2 = m.state()
    m.destroy()
    assert (info["hparity"], info["hsteps"]) == (m.parity, m.steps) and abs(info["hphiScale"] - m.phiScale) < 1e-12
    raw = np.fromfile(info["hout"], np.float32)
    nv, nq = V.size, Q1.size
    same = np.array_equal(raw[:nv], V.ravel()) and np.array_equal(raw[nv:nv + nq], Q1.ravel())
    print("bit-identical" if same else "not bit-identical")
    assert rel_l2(raw[:nv], V.ravel()) <= 1e-6 and rel_l2(raw[nv:nv + nq], Q1.ravel()) <= 1e-5
    assert rel_l2(raw[nv + nq:], Q2.ravel()) <= 1e-5 or not np.any(Q2)
