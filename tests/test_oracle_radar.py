"""Pins of the virtual-radar oracle (radar.js:41-268) and of the radar host's layout and projection
(radar.js:271-418, geo.js:6-29), on the CPU.

Each check is fixed by something other than the restated formula:
- the beam height and ground distance against the geometry of a straight ray in an earth of radius
  ke*a (the 4/3-earth model);
- the gate position against a 3-D unit-vector construction of the great-circle destination and the
  azimuthal-equidistant projection;
- echo-free air gives all-zero moments and no KDP;
- uniform rain gives its Z from the Smith (1975) Z-q relation (Lin) or the exponential-DSD moment
  identity Z = 720 rho Nr / lambda^6 (Thompson), Marshall-Palmer ZDR of D0 = 3.67 / lambda, rho_hv of
  pure rain and the 1 m/s spectrum-width floor;
- a uniform wind gives V . e cos(el) - vt sin(el), aliased into [-Nyquist, Nyquist); the radials of an
  azimuthal-equidistant projection about the site run straight at their azimuth;
- PHIDP is the system offset plus twice the running sum of KDP times the gate spacing, mod 360;
- a cyclonic sub-grid vortex north of the radar reads outbound on its east side, inbound on its west;
- the gate detail for the GPU comparison rounds to the cut's codes and follows a geometry handed to it.
"""
import math
from types import SimpleNamespace

import numpy as np
import pytest

from _oracle import radar_cut, radar_gates, radar_geom
from outbreaklab_amd.radar import (GATE0, GATEDR, KE, RE_M, VCPS, Geo, cut_block, plan_radar)

LAT0, LON0 = 35.2, -97.4


def mk_model(n=96, nz=20, dx=1000.0, scheme="lin"):
    return SimpleNamespace(nx=n, ny=n, nz=nz, dx=dx, dz=20000.0 / nz, Lx=n * dx, Ly=n * dx, rhoSfc=1.0,
                           params={"scheme": scheme})


def mk_state(m, qr=0.0, Nr=0.0, u=0.0, v=0.0, th=300.0, rho=1.0):
    BS = np.zeros((m.nz, 8), np.float32)
    BS[:, 2] = rho
    BS[:, 4] = 1.0  # Exner: T = theta
    V = np.zeros((m.nz + 1, m.ny + 1, m.nx + 1, 4), np.float32)
    V[..., 0], V[..., 1], V[..., 3] = u, v, th
    Q1 = np.zeros((m.nz, m.ny, m.nx, 4), np.float32)
    Q1[..., 3] = qr
    Q2 = np.zeros((m.nz, m.ny, m.nx, 4), np.float32)
    Q2[..., 2] = Nr
    aux = np.zeros((m.nz, m.ny, m.nx, 2), np.float32)
    return BS.ravel(), V, Q1, Q2, aux


def mk_radar(m, dx_site=0.0, dy_site=0.0, vcp=212):
    geo = Geo(LAT0, LON0, m.Lx, m.Ly, 0.0)
    lat, lon = geo.fromModel(m.Lx / 2 + dx_site, m.Ly / 2 + dy_site)
    R = plan_radar(m, geo, {"id": "KTST", "lat": lat, "lon": lon, "h": 0.0}, {"vcp": vcp})
    R["vol"], R["az0"] = 0, 0.0
    return geo, R


def run_cut(m, geo, R, ci, st, tor=None, out=None):
    c = R["cuts"][ci]
    ru = cut_block(m, geo, R, c, 0.0, 0 if tor is None else len(tor) // 8)
    out, kdp = radar_cut(ru, *st, tor, R["bytes"] // 4, 720 * R["ngDop"], out)
    return ru, out, kdp


def moment(R, ci, name, out):
    """(nrad, ngates) uint8 codes of one moment of one cut."""
    c = R["cuts"][ci]
    o = next(x for x in c["mom"] if x["name"] == name)
    b = out.view(np.uint8)
    return b[o["off"]:o["off"] + c["nrad"] * o["ngates"]].reshape(c["nrad"], o["ngates"])


# ---- independent geometry (FP64) ----
def unit(lat, lon):
    return np.array([math.cos(lat) * math.cos(lon), math.cos(lat) * math.sin(lon), math.sin(lat)])


def east_north(lat, lon):
    return (np.array([-math.sin(lon), math.cos(lon), 0.0]),
            np.array([-math.sin(lat) * math.cos(lon), -math.sin(lat) * math.sin(lon), math.cos(lat)]))


def destination(lat, lon, az, s):
    """Great-circle point at bearing az (rad), ground distance s (m), as a unit vector."""
    E, N = east_north(lat, lon)
    d = s / RE_M
    return math.cos(d) * unit(lat, lon) + math.sin(d) * (math.cos(az) * N + math.sin(az) * E)


def aeqd(p, lat0, lon0):
    """Azimuthal-equidistant (x east, y north) of unit vector p about (lat0, lon0)."""
    u0 = unit(lat0, lon0)
    E, N = east_north(lat0, lon0)
    c = math.atan2(np.linalg.norm(np.cross(u0, p)), float(u0 @ p))
    t = p - (u0 @ p) * u0
    tn = np.linalg.norm(t)
    if tn < 1e-15:
        return 0.0, 0.0
    return RE_M * c * float(t @ E) / tn, RE_M * c * float(t @ N) / tn


def ray(r, el, a):
    """Straight ray from an antenna at radius a: (height above the antenna's sphere, arc distance)."""
    P = np.array([r * math.cos(el), a + r * math.sin(el)])
    return float(np.hypot(*P) - a), float(a * math.atan2(P[0], P[1]))


def test_layout_matches_reference_shape():
    m = mk_model()
    _, R = mk_radar(m)
    # reach: the far corner plus 10 km, in whole groups of 4 gates from 2125 m every 250 m
    far = math.hypot(m.Lx / 2, m.Ly / 2) + 10000
    assert R["ngRef"] == max(64, math.ceil(math.ceil((far - GATE0) / GATEDR) / 4) * 4)
    assert R["ngDop"] == min(1192, R["ngRef"])
    cuts = R["cuts"]
    assert [c["kind"] for c in cuts[:6]] == ["surv", "dop"] * 3 and all(c["kind"] == "both" for c in cuts[6:])
    assert len(cuts) == 17 and cuts[0]["nrad"] == 720 and cuts[6]["nrad"] == 360
    # moments tile the volume without gaps, radial-major
    off = 0
    for c in cuts:
        for o in c["mom"]:
            assert o["off"] == off
            off += c["nrad"] * o["ngates"]
    assert off == R["bytes"] and R["bytes"] % 4 == 0
    # the scan schedule covers the VCP's duration, split cuts taking twice as long
    assert cuts[0]["t0"] == 0 and abs(cuts[-1]["t0"] + cuts[-1]["dur"] - VCPS[212]["dur"]) < 1e-9
    assert abs(cuts[0]["dur"] - 2 * cuts[6]["dur"]) < 1e-9


def test_geo_projection_against_vectors():
    g = Geo(LAT0, LON0, 200e3, 150e3)
    D = math.pi / 180
    for lat, lon in [(LAT0, LON0), (36.1, -96.2), (34.0, -98.9), (35.2, -95.0), (37.5, -97.4)]:
        x, y = g.fwd(lat, lon)
        xe, ye = aeqd(unit(lat * D, lon * D), LAT0 * D, LON0 * D)
        assert abs(x - xe) < 1e-4 and abs(y - ye) < 1e-4
        la, lo = g.inv(x, y)
        assert abs(la - lat) < 1e-9 and abs(lo - lon) < 1e-9
        mx, my = g.toModel(lat, lon)
        assert abs(mx - (x + 100e3)) < 1e-9 and abs(my - (y + 75e3)) < 1e-9


@pytest.mark.parametrize("ci", [0, 13])
def test_beam_geometry(ci):
    m = mk_model()
    geo, R = mk_radar(m)
    ru = cut_block(m, geo, R, R["cuts"][ci], 0.0)
    el, a = float(ru[8]), float(ru[3])
    assert abs(a - KE * RE_M) < 1
    for gi in [0, 7, 150, R["ngRef"] - 1]:
        h, s, _, _ = radar_geom(ru, 5, gi)
        he, se = ray(GATE0 + GATEDR * gi, el, a)
        assert abs(h - he) <= 1e-5 * abs(he) + 0.02, (gi, h, he)
        assert abs(s - se) <= 1e-5 * se + 0.5, (gi, s, se)


@pytest.mark.parametrize("dxy", [(0.0, 0.0), (-30e3, 22e3)])
def test_gate_position(dxy):
    m = mk_model()
    geo, R = mk_radar(m, *dxy)
    ru = cut_block(m, geo, R, R["cuts"][1], 0.0)
    D = math.pi / 180
    slat, slon = R["site"]["lat"] * D, R["site"]["lon"] * D
    worst = 0.0
    for ri in [0, 1, 97, 250, 360, 533, 719]:
        az = (R["az0"] + 0.5 * ri) * D
        for gi in [0, 40, 200, R["ngRef"] - 1]:
            _, s, x, y = radar_geom(ru, ri, gi)
            xe, ye = aeqd(destination(slat, slon, az, float(s)), LAT0 * D, LON0 * D)
            worst = max(worst, math.hypot(x - (xe + m.Lx / 2), y - (ye + m.Ly / 2)))
    assert worst < 5.0, worst


def test_gate_values_and_geometry_override():
    """orc_radar_gates (the GPU comparison's gate detail): its values before rounding round to the cut's own codes;
    its geometry is radar_geom's; handed back its own geometry it reproduces its values bit for bit, and a geometry
    moved out of the grid, or turned, changes them as it must (no echo; the radial wind read along the new
    direction)."""
    m = mk_model()
    geo, R = mk_radar(m)
    U, Vw = 25.0, -9.0
    st = mk_state(m, qr=2e-3, u=U, v=Vw)
    ru, out, kdp = run_cut(m, geo, R, 1, st)
    c = R["cuts"][1]
    rng = np.random.default_rng(3)
    rg = np.stack([rng.integers(0, c["nrad"], 3000), rng.integers(0, R["ngDop"], 3000)], 1).astype(np.int32)
    rows = radar_gates(ru, *st, None, rg)
    assert np.all(np.isnan(rows[:, 12:17]))
    for name, k in [("REF", 0), ("VEL", 1), ("SW", 2)]:
        code = moment(R, 1, name, out)[rg[:, 0], rg[:, 1]]
        v = rows[:, k]
        ok = np.isfinite(v)
        assert ok.sum() > 300
        assert np.array_equal(np.clip(np.rint(v[ok]), 2, 255).astype(np.uint8), code[ok]), name
        assert np.all(code[~ok] == 0), name
    for (ri, gi), g in list(zip(rg, rows))[:50]:
        h, _, x, y = radar_geom(ru, int(ri), int(gi))
        assert g[17] == x and g[18] == y and g[21] == np.float32(np.float32(ru[2] - ru[33]) + h)
        assert abs(math.hypot(g[19], g[20]) - 1.0) < 1e-6
    gg = rows[:, 17:24].copy()
    same = radar_gates(ru, *st, None, rg, gg)
    assert np.array_equal(same[:, 12:17], rows[:, 0:5], equal_nan=True)
    # the KDP columns: the cut's own KDP buffer at the oracle's geometry, and the same again at that geometry
    # handed back (rain everywhere: every gate inside the domain has one)
    assert np.array_equal(rows[:, 24], kdp[rg[:, 0] * R["ngDop"] + rg[:, 1]]) and np.count_nonzero(rows[:, 24] > 0) > 300
    assert np.array_equal(same[:, 25], rows[:, 24]) and np.all(np.isnan(rows[:, 25]))
    out_ = gg.copy()
    out_[:, 0] = -1000.0
    assert np.all(np.isnan(radar_gates(ru, *st, None, rg, out_)[:, 12:17]))
    turn = gg.copy()
    turn[:, 2:4] = [1.0, 0.0]  # east: the radial wind is U cos(el) - vt sin(el)
    t = radar_gates(ru, *st, None, rg, turn)
    ok = np.isfinite(rows[:, 1])
    el = c["angle"] * math.pi / 180
    # turned east, the reading gains 2 cos(el) (U - V . e) codes, modulo the Nyquist wrap (4 Nyquist codes)
    want = rows[ok, 1] + 2.0 * math.cos(el) * (U - (U * gg[ok, 2] + Vw * gg[ok, 3]))
    d = (t[ok, 13] - want) / (4.0 * c["nyquist"])
    assert ok.sum() > 300 and np.all(np.abs(d - np.round(d)) * 4.0 * c["nyquist"] < 2e-3)


def test_echo_free_is_zero():
    m = mk_model()
    geo, R = mk_radar(m, -10e3, 5e3)
    st = mk_state(m, u=12.0, v=-4.0)
    out = None
    for ci in (0, 1, 7):
        _, out, kdp = run_cut(m, geo, R, ci, st, out=out)
        assert np.all(kdp[:R["cuts"][ci]["nrad"] * R["ngDop"]] == -1.0)
    assert not out.any()


def inside_margin(m, R, ci, margin=1500.0):
    """+1 for gates inside the grid by > margin, -1 outside by > margin, 0 near an edge (FP64 geometry,
    the same unit-vector construction as test_gate_position, vectorised)."""
    c = R["cuts"][ci]
    D = math.pi / 180
    slat, slon = R["site"]["lat"] * D, R["site"]["lon"] * D
    el, a = c["angle"] * D, KE * RE_M
    r = GATE0 + GATEDR * np.arange(R["ngRef"], dtype=np.float64)
    P = np.stack([r * math.cos(el), a + r * math.sin(el)])
    h, s = np.hypot(P[0], P[1]) - a, a * np.arctan2(P[0], P[1])
    az = (R["az0"] + 360.0 / c["nrad"] * np.arange(c["nrad"])) * D
    E, N = east_north(slat, slon)
    d = (s / RE_M)[None, :, None]
    dirs = np.cos(az)[:, None, None] * N + np.sin(az)[:, None, None] * E
    p = np.cos(d) * unit(slat, slon) + np.sin(d) * dirs
    u0 = unit(LAT0 * D, LON0 * D)
    E0, N0 = east_north(LAT0 * D, LON0 * D)
    dot = p @ u0
    cang = np.arctan2(np.linalg.norm(np.cross(u0, p), axis=-1), dot)
    t = p - dot[..., None] * u0
    tn = np.maximum(np.linalg.norm(t, axis=-1), 1e-300)
    x = RE_M * cang * (t @ E0) / tn + m.Lx / 2
    y = RE_M * cang * (t @ N0) / tn + m.Ly / 2
    top = m.nz * m.dz - 200.0 - 20.0 - h[None, :]
    dist_in = np.minimum.reduce([x, m.Lx - x, y, m.Ly - y, np.broadcast_to(top, x.shape)])
    return np.where(dist_in > margin, 1, np.where(dist_in < -margin, -1, 0)).astype(np.int8)


@pytest.fixture(scope="module")
def lab_c1():
    m = mk_model()
    geo, R = mk_radar(m)
    return inside_margin(m, R, 1)


def test_uniform_rain_lin(lab_c1):
    m = mk_model()
    geo, R = mk_radar(m)
    qr = 2e-3
    st = mk_state(m, qr=qr)
    _, out, _ = run_cut(m, geo, R, 0, st)  # surveillance: REF ZDR PHI RHO
    _, out, _ = run_cut(m, geo, R, 1, st, out=out)  # Doppler: REF VEL SW
    Z = 3.63e9 * qr ** 1.75
    ref_f = 2 * 10 * math.log10(Z) + 66
    lam = (math.pi * 1000 * 8e6 / qr) ** 0.25  # Marshall-Palmer slope (m^-1)
    d0 = 3.67 / lam * 1000  # median volume diameter (mm)
    zdr_f = 16 * min(max(0.2 + 1.15 * (d0 - 0.5), 0.1), 4.8) + 128
    rho_f = 300 * (0.986 if d0 > 2.5 else 0.995) - 60.5
    inn, out_ = lab_c1 == 1, lab_c1 == -1
    assert inn.sum() > 10000 and out_.sum() > 10000
    for ci, name, f in [(0, "REF", ref_f), (1, "REF", ref_f), (0, "ZDR", zdr_f), (0, "RHO", rho_f)]:
        code = moment(R, ci, name, out).astype(np.int32)
        assert abs(round(f) - f) < 0.45
        assert np.all(code[inn] == round(f)), (ci, name, np.unique(code[inn]), f)
        assert np.all(code[out_] == 0), (ci, name)
    sw = moment(R, 1, "SW", out).astype(np.int32)
    assert np.all(sw[inn] == 131) and np.all(sw[out_] == 0)


def test_uniform_rain_thompson(lab_c1):
    m = mk_model(scheme="thompson")
    geo, R = mk_radar(m)
    qr, Nr, rho = 2e-3, 1e4, 0.9
    m.rhoSfc = rho
    st = mk_state(m, qr=qr, Nr=Nr, rho=rho)
    _, out, _ = run_cut(m, geo, R, 1, st)
    lam = (math.pi * 1000 * Nr / qr) ** (1 / 3)  # exponential DSD: rho qr / (rho Nr) = pi rho_w / lambda^3
    Z = 720e18 * rho * Nr / lam ** 6  # 6th moment: N0 Gamma(7) / lambda^7 = 720 (rho Nr) / lambda^6
    f = 2 * 10 * math.log10(Z) + 66
    assert abs(round(f) - f) < 0.45
    code = moment(R, 1, "REF", out).astype(np.int32)
    assert np.all(code[lab_c1 == 1] == round(f)), (np.unique(code[lab_c1 == 1]), f)


def fold(v, n):
    return v - 2 * n * math.floor((v + n) / (2 * n))


def test_uniform_wind_radial_velocity(lab_c1):
    m = mk_model()
    geo, R = mk_radar(m)
    U, Vw = 35.0, -12.0
    st = mk_state(m, qr=0.1, u=U, v=Vw)  # heavy rain: fall speed at its 10 m/s cap (rho = rho_sfc)
    ru, out, _ = run_cut(m, geo, R, 1, st)
    c = R["cuts"][1]
    el, nyq = c["angle"] * math.pi / 180, c["nyquist"]
    vel = moment(R, 1, "VEL", out).astype(np.int32)
    checked = folded = hits = total = 0
    for ri in range(c["nrad"]):
        az = (R["az0"] + 0.5 * ri) * math.pi / 180
        vr = (U * math.sin(az) + Vw * math.cos(az)) * math.cos(el) - 10.0 * math.sin(el)
        f = 2 * fold(vr, nyq) + 129
        if abs(abs(vr) - nyq) < 0.5:
            continue
        row = vel[ri][lab_c1[ri] == 1]
        # the radial's direction comes from two FP32 gate positions 200 m apart (radar.js:152): up to
        # ~0.2 m/s (0.4 code) of rounding noise in V . e
        assert np.all(np.abs(row - f) < 0.9), (ri, np.unique(row), f)
        hits += np.count_nonzero(row == round(f))
        total += row.size
        checked += 1
        folded += abs(vr) > nyq
    assert checked > 600 and folded > 50 and hits > 0.9 * total, (checked, folded, hits / total)


def test_phidp_accumulates_kdp():
    m = mk_model()
    geo, R = mk_radar(m, 20e3, -15e3)
    st = mk_state(m, qr=0.02)
    ru, out, kdp = run_cut(m, geo, R, 0, st)
    c = R["cuts"][0]
    ng = R["ngDop"]
    K = kdp[:c["nrad"] * ng].reshape(c["nrad"], ng).astype(np.float64)
    phi = moment(R, 0, "PHI", out).astype(np.int32)
    assert (K > 0).sum() > 10000
    wrapped = 0
    for ri in range(c["nrad"]):
        acc = 25.0
        for gi in range(ng):
            if K[ri, gi] < 0:
                assert phi[ri, gi] == 0
                continue
            acc += 2 * K[ri, gi] * GATEDR * 0.001
            wrapped |= acc > 360
            f = min(max((acc % 360) * 253 / 360 + 2, 2), 255)
            if abs(f - round(f) - 0.5) > 0.02 and abs(f - round(f) + 0.5) > 0.02:
                assert phi[ri, gi] == round(f), (ri, gi, phi[ri, gi], f)
    assert wrapped


def test_tornado_couplet():
    m = mk_model()
    geo, R = mk_radar(m)
    st = mk_state(m, qr=2e-3)
    tor = np.array([m.Lx / 2, m.Ly / 2 + 40e3, 150.0, 60.0, 0.0, 0.0, 0.0, 0.0], np.float32)
    _, out, _ = run_cut(m, geo, R, 1, st, tor=tor)
    vel = moment(R, 1, "VEL", out).astype(np.int32)
    gi = round((40e3 - GATE0) / GATEDR)
    east, west = vel[1, gi - 1:gi + 2], vel[719, gi - 1:gi + 2]
    assert east.max() > 129 + 30 and west.min() < 129 - 30, (east, west)
    # far from the vortex (beyond 8 effective radii) the still air reads 0 m/s less the rain's fall speed
    assert np.all(vel[180, 20:170] == 129) and np.all(vel[540, 20:170] == 129)
    _, out0, _ = run_cut(m, geo, R, 1, st)
    v0 = moment(R, 1, "VEL", out0)
    assert np.array_equal(v0[180:541], moment(R, 1, "VEL", out)[180:541])
