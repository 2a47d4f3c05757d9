"""Virtual WSR-88D volumes (radar.js:41-418) on the GPU against the FP32 oracle.

The 128 x 128 x 40 Thompson workload after 300 steps (30 min), one radar 40 km west / 30 km south of
the domain centre, VCP 212 (17 cuts, all six moments), one sub-grid tornado with debris at the storm's
strongest low-level rotation. The oracle is fed the same downloaded state (V of the current parity,
Q1[0], Q2[0], aux) and the same RADAR_FIELDS blocks. The 8-bit Archive II codes must match: the GPU's
sin/cos/asin/atan2/pow differ from the C library's by an ulp or two, which moves a code by one where
its value sits on a rounding boundary, and moves the FP32 gate geometry by up to a metre (a code then
differs where reflectivity or the radial wind changes fast). Bar, per moment of each cut, over its echo
gates: <= 1 % off by one (VEL: <= 4 %) and <= 0.1 % (at least 2 gates) off by more; and every differing
code explained, by its value's rounding or by the oracle evaluated at the GPU's own geometry of the gate."""
import numpy as np
import pytest

from _oracle import radar_cut

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def storm():
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_128

    m = RegionalModel()
    dt = setup_plain_128(m, scheme="thompson")
    for _ in range(300):
        m.step(dt)
    yield m
    m.destroy()


def tornado_at_rotation(m):
    got = m.requestMaps()()
    zeta = got["m0"].reshape(m.ny, m.nx, 4)[..., 3]
    j, i = np.unravel_index(np.argmax(zeta), zeta.shape)
    return {"x": (i + 0.5) * m.dx, "y": (j + 0.5) * m.dx, "R": 150.0, "v": 70.0, "debrisTop": 1500.0,
            "debrisDbz": 60.0, "debris": 1.0}


@pytest.fixture(scope="module")
def volume(storm):
    from outbreaklab_amd.radar import Geo, RadarSim

    m = storm
    geo = Geo(35.9, -97.0, m.Lx, m.Ly, 0.0)
    sim = RadarSim(m, geo)
    tor = tornado_at_rotation(m)
    sim.setTornadoes([tor])
    lat, lon = geo.fromModel(m.Lx / 2 - 40e3, m.Ly / 2 - 30e3)
    R = sim.addRadar({"id": "KTST", "lat": lat, "lon": lon, "h": 0.0}, {"vcp": 212, "seed": 3})
    R["phase"] = 0.0
    done = sim.encDue(None, 269.0)
    assert len(done) == 1 and R["next"] == len(R["cuts"])
    vol = sim.collect(done[0], "test", 0.0)
    yield m, geo, sim, R, tor, vol
    sim.destroy()


def oracle_volume(m, geo, R, tor, simT):
    from outbreaklab_amd.radar import cut_block

    V, Q1, Q2 = m.state()
    aux = m.aux.download()
    t = np.array([tor["x"], tor["y"], tor["R"], tor["v"], tor["debrisTop"], tor["debrisDbz"], tor["debris"], 0],
                 np.float32)
    out = None
    for c in R["cuts"]:
        ru = cut_block(m, geo, R, c, simT, 1)
        out, _ = radar_cut(ru, m._B, V, Q1, Q2, aux, t, R["bytes"] // 4, 720 * R["ngDop"], out)
    return out.view(np.uint8)


def test_volume_matches_oracle(volume):
    m, geo, sim, R, tor, vol = volume
    got = vol["bytes"]
    ref = oracle_volume(m, geo, R, tor, 269.0)
    assert got.shape == ref.shape
    hdr = vol["header"]
    assert hdr["vcp"] == 212 and len(hdr["sweeps"]) == 17
    n_echo = 0
    for c, sw in zip(R["cuts"], hdr["sweeps"]):
        assert sw["nrad"] == c["nrad"] and [x["name"] for x in sw["moments"]] == [o["name"] for o in c["mom"]]
        for o in c["mom"]:
            a = got[o["off"]:o["off"] + c["nrad"] * o["ngates"]].astype(np.int32)
            b = ref[o["off"]:o["off"] + c["nrad"] * o["ngates"]].astype(np.int32)
            d = np.abs(a - b)
            off1, big = int((d == 1).sum()), int((d > 1).sum())
            print(f"cut {c['index']:2d} {c['angle']:4.1f} {c['kind']:4s} {o['name']}: {int((a > 0).sum()):7d} echo gates, "
                  f"{off1} off by 1, {big} off by more")
            if o["name"] == "REF":
                n_echo += int((b >= 66 + 2 * 35).sum())
            echo = max(int((b > 0).sum()), 1)
            # VEL: the radial direction is the difference of two FP32 gate positions 200 m apart
            # (radar.js:152), so an ulp of sin/cos moves V . e by up to ~0.1 m/s
            lim1 = 0.04 if o["name"] == "VEL" else 0.01
            assert off1 <= lim1 * echo and big <= max(2, 1e-3 * echo), (c["index"], o["name"], off1, big, echo)
    assert n_echo > 1000, "the storm shows no echo above 35 dBZ"


def test_tornado_signature(volume):
    """The vortex's velocity couplet: radials either side of the tornado read opposite signs."""
    m, geo, sim, R, tor, vol = volume
    import math

    c = R["cuts"][1]  # 0.5 deg Doppler
    o = next(x for x in c["mom"] if x["name"] == "VEL")
    vel = vol["bytes"][o["off"]:o["off"] + c["nrad"] * o["ngates"]].reshape(c["nrad"], o["ngates"]).astype(np.int32)
    dx, dy = tor["x"] - R["sx"], tor["y"] - R["sy"]
    rng, az = math.hypot(dx, dy), (math.degrees(math.atan2(dx, dy)) - R["az0"]) % 360
    ri, gi = int(round(az / 0.5)) % 720, int(round((rng - 2125) / 250))
    left, right = vel[(ri - 1) % 720, gi - 2:gi + 3], vel[(ri + 1) % 720, gi - 2:gi + 3]
    print("left of the vortex", left, "right", right)
    assert right.max() - left.min() > 40


def test_readback_slots_and_schedule(storm):
    from outbreaklab_amd.radar import Geo, RadarSim

    m = storm
    geo = Geo(35.9, -97.0, m.Lx, m.Ly, 0.0)
    sim = RadarSim(m, geo)
    R = sim.addRadar({"id": "KTWO", "lat": 35.9, "lon": -97.0, "h": 300.0}, {"vcp": 35, "seed": 5})
    R["phase"] = 10.0
    assert sim.encDue(None, 5.0) == []  # before this site's first volume
    d1 = sim.encDue(None, 10.0 + 419.0)
    assert len(d1) == 1 and d1[0]["slot"] == 0
    d2 = sim.encDue(None, 10.0 + 420.0 + 419.0)
    assert len(d2) == 1 and d2[0]["slot"] == 1
    assert sim.encDue(None, 10.0 + 3 * 420.0) == []  # both readback slots in flight: the next volume waits
    v1 = sim.collect(d1[0])
    d3 = sim.encDue(None, 10.0 + 3 * 420.0)
    assert len(d3) == 1 and d3[0]["slot"] == 0 and d3[0]["vol"] == 2
    v2, v3 = sim.collect(d2[0]), sim.collect(d3[0])
    # each volume fills the whole layout; its header dates it from the volume start
    assert v1["bytes"].size == v2["bytes"].size == v3["bytes"].size == R["bytes"]
    assert v1["header"]["epoch_ms"] == 10000 and v2["header"]["epoch_ms"] == 430000
    sim.destroy()


def test_upload_body_of_a_gpu_volume(volume):
    """The volume the GPU made, as the upload body (app.js:1743-1750) parsed the way the server parses it
    (server.py:250-266): the header and every moment's codes come back unchanged."""
    import json
    import struct
    import zlib

    from outbreaklab_amd.radar import volume_body

    m, geo, sim, R, tor, vol = volume
    body = zlib.decompress(volume_body(vol["header"], vol["bytes"]), 16 + zlib.MAX_WBITS)
    (hl,) = struct.unpack("<I", body[:4])
    H = json.loads(body[4:4 + hl])
    assert H == json.loads(json.dumps(vol["header"]))
    payload = np.frombuffer(body, np.uint8, offset=4 + hl)
    assert np.array_equal(payload, vol["bytes"])


# Which gates may differ, and by how much. The GPU and the oracle compute every gate in FP32 with different
# transcendentals (hardware exp2 / log2 for powers, the device library's sin / cos / asin / atan2 against the C
# library's) and different FMA contraction, so a gate's values differ by a few ulp, and by more in the radial
# direction e (the difference of two gate positions 200 m apart, radar.js:152: about 5e-5 relative). A code can then
# differ only where the gate sits near one of its discrete boundaries; each bound below is that error budget with a
# margin, in the boundary's own units. The gate geometry differs by more: the great-circle formulas
# (radar.js:57-69) cancel in FP32 (the longitude difference, the y-difference of products), so a transcendental an
# ulp apart moves a gate by up to a metre and turns the radial direction e (the difference of two positions 200 m
# apart) by up to a few 1e-3 rad; near an echo edge a metre moves reflectivity by tenths of a dB, and the turn
# moves VEL and SW by up to a tenth of a code at storm winds. Those gates are checked by evaluating the oracle at
# the GPU's own geometry of the gate (obl_radar_geometry, the cut kernel's geometry code).
ROUND_CODES = {"REF": 1e-3, "VEL": 0.02, "SW": 0.02, "ZDR": 2e-3, "RHO": 1e-3}  # from a rounding boundary k + 1/2
MOM = {"REF": 0, "VEL": 1, "SW": 2, "ZDR": 3, "RHO": 4}
NEAR = {"geometry": (5, 0.05), "echo threshold": (6, 1e-3), "Nyquist wrap": (7, 0.01), "melting level": (8, 2e-3),
        "graupel switch": (9, 1e-5), "drop-size switch": (10, 1e-5), "Thompson size switch": (11, 1e-5)}


def phidp_values(kdp, ng, dr, offset):
    """The oracle's PHI values before rounding (radar.js:243-267, the same FP32 running sum), per gate, with
    each gate's distance from a 360-degree wrap and the running sum itself."""
    acc = np.float32(offset)
    vals, wrap, accs = np.full(ng, np.nan, np.float32), np.full(ng, np.inf), np.zeros(ng)
    for g in range(ng):
        kd = kdp[g]
        if kd >= 0:
            acc = np.float32(acc + np.float32(np.float32(2.0) * kd * np.float32(dr)) * np.float32(0.001))
            phi = np.float32(acc - np.float32(360.0) * np.floor(acc / np.float32(360.0)))
            vals[g] = phi * np.float32(253.0) / np.float32(360.0) + np.float32(2.0)
            wrap[g] = abs(acc / 360.0 - round(acc / 360.0)) * 360.0
        accs[g] = acc
    return vals, wrap, accs


def phi_why(kdp, gi, gpu, ru):
    """Why gate gi's PHI code may be the GPU's ``gpu`` given a radial's KDP (None: no reason): the running sum of
    ulp-different KDP puts the value within its accumulated bound of a rounding boundary or of a 360-degree wrap,
    or the sum gives the GPU's code."""
    ng = len(kdp)
    pv, pw, acc = phidp_values(kdp, ng, float(ru[4 * 3 + 3]), float(ru[4 * 4 + 3]))
    bound = 1e-5 * abs(acc[gi]) + 1e-4  # degrees: the running sum of ulp-different KDP
    x = float(pv[gi]) - 0.5
    if np.isfinite(pv[gi]) and code_of(pv[gi]) == gpu:
        return "geometry"
    if np.isfinite(pv[gi]) and abs(x - round(x)) <= bound * 253.0 / 360.0:
        return "rounding"
    if pw[gi] <= bound:
        return "Nyquist wrap"  # the 360-degree wrap of PHI
    return None


def code_of(v):
    return int(np.clip(np.rint(v), 2, 255)) if np.isfinite(v) else 0


def classify(name, gpu, g):
    """The boundary a gate's differing code of moment ``name`` sits near (None: none within its bound); gpu = the
    GPU's code, g = the gate's row of radar_gates (with the GPU's geometry)."""
    k = MOM.get(name, 0)  # a PHI code exists exactly where REF does
    if name in MOM and np.isfinite(g[k]):
        x = float(g[k]) - 0.5
        if abs(x - round(x)) <= ROUND_CODES[name]:
            return "rounding"
    vg = float(g[12 + k])
    if name in MOM:
        if gpu == code_of(vg):
            return "geometry"  # the oracle at the GPU's own geometry gives the GPU's code
        if np.isfinite(vg) and abs(vg - 0.5 - round(vg - 0.5)) <= ROUND_CODES[name]:
            return "geometry"
    elif (gpu == 0) == (not np.isfinite(vg)) and np.isfinite(g[0]) != np.isfinite(vg):
        return "geometry"
    for c, (i, lim) in NEAR.items():
        if g[i] <= lim:
            return c
    return None


def test_volume_codes_differ_only_near_a_boundary(volume):
    """Every code the GPU's volume and the oracle's differ in is explained: the gate's moment value lies within the
    rounding bound of a code boundary; or the oracle, evaluated at the GPU's own FP32 geometry of that gate
    (obl_radar_geometry), gives the GPU's code (or a value within the rounding bound of its boundary); or the gate
    lies within the stated bound of the echo threshold, a Nyquist wrap, the beam leaving the model, the melting
    level or the graupel / drop-size / Thompson size switches (or, for PHI, a rounding or 360-degree wrap of the
    running sum within its accumulated bound, of the oracle's KDP or of the oracle's KDP at the GPU's own geometry
    of every gate of the radial, the latter also where it gives the GPU's code). Any other difference fails. On a
    sample of gates the GPU's geometry
    lies within a few metres and a few 1e-3 rad of the oracle's (what the FP32 formulas allow)."""
    from collections import Counter

    from _oracle import radar_gates
    from outbreaklab_amd.radar import cut_block, gate_geometry

    m, geo, sim, R, tor, vol = volume
    got = vol["bytes"]
    V, Q1, Q2 = m.state()
    aux = m.aux.download()
    t = np.array([tor["x"], tor["y"], tor["R"], tor["v"], tor["debrisTop"], tor["debrisDbz"], tor["debris"], 0],
                 np.float32)
    rng = np.random.default_rng(7)
    classes, ndiff, ncodes = Counter(), 0, 0
    dgeo, same, nsamp = np.zeros(4), 0, 0
    out = None
    for c in R["cuts"]:
        ru = cut_block(m, geo, R, c, 269.0, 1)
        out, kdp = radar_cut(ru, m._B, V, Q1, Q2, aux, t, R["bytes"] // 4, 720 * R["ngDop"], out)
        ref = out.view(np.uint8)
        diffs = []
        for o in c["mom"]:
            ng, name = o["ngates"], o["name"]
            a = got[o["off"]:o["off"] + c["nrad"] * ng].reshape(c["nrad"], ng)
            b = ref[o["off"]:o["off"] + c["nrad"] * ng].reshape(c["nrad"], ng)
            ncodes += a.size
            diffs += [(name, int(ri), int(gi), int(a[ri, gi])) for ri, gi in zip(*np.nonzero(a != b))]
        sample = np.stack([rng.integers(0, c["nrad"], 2000), rng.integers(0, R["ngRef"], 2000)], 1)
        rg = np.concatenate([np.array([d[1:3] for d in diffs], np.int32).reshape(-1, 2), sample]).astype(np.int32)
        gg = gate_geometry(m, ru, rg)
        rows = radar_gates(ru, m._B, V, Q1, Q2, aux, t, rg, gg)
        d = np.abs(gg - rows[:, 17:24])
        turn = np.abs(gg[:, 2] * rows[:, 20] - gg[:, 3] * rows[:, 19])
        dgeo = np.maximum(dgeo, [d[:, 0].max(), d[:, 1].max(), d[:, 4:7].max(), turn.max()])
        same += int(np.count_nonzero(d[len(diffs):].max(axis=1) == 0))
        nsamp += len(sample)
        for (name, ri, gi, gpu), g in zip(diffs, rows):
            ndiff += 1
            why = classify(name, gpu, g)
            if why is None and name == "PHI":
                why = phi_why(kdp[ri * R["ngDop"]:(ri + 1) * R["ngDop"]], gi, gpu, ru)
                if why is None:
                    # the running sum of the oracle's KDP at the GPU's own geometry of every gate of the radial
                    rr = np.stack([np.full(R["ngDop"], ri), np.arange(R["ngDop"])], 1).astype(np.int32)
                    kg = radar_gates(ru, m._B, V, Q1, Q2, aux, t, rr, gate_geometry(m, ru, rr))[:, 25]
                    if phi_why(kg, gi, gpu, ru) is not None:
                        why = "geometry" 
            classes[why or "unexplained"] += 1
    print(f"{ndiff} of {ncodes} codes differ:", dict(classes), "| GPU vs oracle geometry, largest |dx| |dy| |dz| (m),"
          f" turn (rad): {dgeo}; sampled gates with identical geometry: {same} of {nsamp}")
    assert classes["unexplained"] == 0, dict(classes)
    assert dgeo[0] < 5.0 and dgeo[1] < 5.0 and dgeo[2] < 0.1 and dgeo[3] < 0.02, dgeo
