"""Soundings: SHARPpy parser, parcel/kinematic analysis and the hydrostatic base state.

Restates the reference's Sounding module (sounding.js:10-197) on the host, and
adds the analytic Weisman-Klemp sounding the workloads use (no reference
counterpart: the reference reads SHARPpy text, so ``wk_sounding_text`` writes
one).

``base_state(snd, nz, dz)`` returns the (nz+1)*8 float32 table the step reads:
per level k (θ0, qv0, ρ0 centre, ρ0 bottom face, π0, p0 Pa, u0, v0) and ρ at
the lid face in slot 3 of level nz (sounding.js:139-197).
"""
from __future__ import annotations

import math
import re

import numpy as np

KT = 0.514444
RD, CP, LV, EPS, G, RV = 287.04, 1005.7, 2.501e6, 0.622, 9.80665, 461.5


class SoundingError(ValueError):
    pass


def _isfinite(v) -> bool:
    return v is not None and isinstance(v, (int, float)) and math.isfinite(v)


def parse_sharppy(text: str) -> dict:
    """parseSharppy (sounding.js:10-32): %TITLE% / %RAW% blocks of PRES, HGHT, TEMP, DWPT, WDIR, WSPD(kt)."""
    out = {"title": "", "station": "", "time": None, "levels": []}
    mode = ""
    for raw in text.replace("\r", "").split("\n"):
        L = raw.strip()
        if not L:
            continue
        if re.match(r"^%TITLE%", L, re.I):
            mode = "title"
            continue
        if re.match(r"^%RAW%", L, re.I):
            mode = "raw"
            continue
        if re.match(r"^%END%", L, re.I):
            mode = ""
            continue
        if mode == "title":
            if not out["title"]:
                out["title"] = L
                out["station"] = L.split()[0]
            continue
        if mode == "raw" or (not mode and re.match(r"^-?\d", L)):
            toks = [t for t in re.split(r"[,\s]+", L) if t]
            try:
                f = [float(t) for t in toks]
            except ValueError:
                continue
            if len(f) < 6 or any(not math.isfinite(v) for v in f):
                continue
            p, h, t, td, wd, ws = f[:6]
            if p <= 0 or h < -1000 or h > 40000 or p == -9999:
                continue
            nan = float("nan")
            out["levels"].append({"p": p, "h": h, "t": t if t > -9990 else nan, "td": td if td > -9990 else nan,
                                  "wd": wd if wd > -9990 else nan, "ws": ws if ws > -9990 else nan})
    # Array.prototype.sort is stable
    out["levels"].sort(key=lambda l: -l["p"])
    if len(out["levels"]) < 4:
        raise SoundingError("Could not read at least 4 levels. Expected SHARPpy format: %RAW% rows of  PRES, HGHT, "
                            "TEMP, DWPT, WDIR, WSPD(kt)  ... %END%")
    return out


def es(t: float) -> float:
    """Saturation vapour pressure over water (hPa), t in deg C (sounding.js:33)."""
    return 6.112 * math.exp(17.67 * t / (t + 243.5))


def mixr(p: float, t: float) -> float:
    e = es(t)
    return EPS * e / max(p - e, 1e-3)


def _tvK(tK, w):
    return tK * (1 + 0.61 * w)


def _interp_logp(levels, key, p):
    for i in range(len(levels) - 1):
        a, b = levels[i], levels[i + 1]
        if a["p"] >= p >= b["p"]:
            va, vb = a[key], b[key]
            if not (math.isfinite(va) and math.isfinite(vb)):
                return va if math.isfinite(va) else vb
            f = math.log(a["p"] / p) / math.log(a["p"] / b["p"])
            return va + (vb - va) * f
    return float("nan")


def _moist_step(p, tK, dp):
    def f(pp, T):
        rs = mixr(pp, T - 273.15)
        return (1 / pp) * (RD * T + LV * rs) / (CP + LV * LV * rs * EPS / (RD * T * T))

    k1 = f(p, tK)
    k2 = f(p + dp / 2, tK + dp / 2 * k1)
    k3 = f(p + dp / 2, tK + dp / 2 * k2)
    k4 = f(p + dp, tK + dp * k3)
    return tK + dp / 6 * (k1 + 2 * k2 + 2 * k3 + k4)


def _lift_parcel(grid, p0, t0, td0):
    """liftParcel (sounding.js:53-84)."""
    T0, TD0 = t0 + 273.15, td0 + 273.15
    tl = 1 / (1 / (TD0 - 56) + math.log(T0 / TD0) / 800) + 56
    pl = p0 * (tl / T0) ** (1 / 0.2857)
    w0 = mixr(p0, td0)
    theta = T0 * (1000 / p0) ** 0.2857
    cin, lfcH, elH, above, capeTot = 0.0, float("nan"), float("nan"), False, 0.0
    moistP, moistT = pl, tl
    pts = []
    for k, p in enumerate(grid["p"]):
        if p > p0:
            continue
        if p >= pl:
            tp, wp = theta * (p / 1000) ** 0.2857, w0
        else:
            while moistP > p + 1e-6:
                dp = max(-5, p - moistP)
                moistT = _moist_step(moistP, moistT, dp)
                moistP += dp
            tp, wp = moistT, mixr(p, moistT - 273.15)
        pts.append((p, grid["h"][k], _tvK(tp, wp), grid["tv"][k]))
    lclH = grid["hAt"](pl)
    for k in range(1, len(pts)):
        a, b = pts[k - 1], pts[k]
        buoy = ((a[2] - a[3]) / a[3] + (b[2] - b[3]) / b[3]) / 2
        e = G * buoy * (b[1] - a[1])
        if buoy > 0:
            if not above and b[0] < pl + 1:
                above = True
                if not math.isfinite(lfcH):
                    lfcH = a[1]
            if above:
                capeTot += e
                elH = b[1]
        elif not above:
            cin += e
    return {"cape": max(0.0, capeTot), "cin": min(0.0, cin), "lclH": lclH, "lfcH": lfcH, "elH": elH, "pLcl": pl}


def analyze(snd: dict, parcels: bool = True) -> dict:
    """analyze (sounding.js:86-137): winds (windAt) and, with ``parcels``, SB/ML/MU CAPE."""
    lv = [l for l in snd["levels"] if math.isfinite(l["h"])]
    sfc = lv[0]
    sfcH = sfc["h"]
    wl = []
    for l in lv:
        if math.isfinite(l["wd"]) and math.isfinite(l["ws"]):
            d, s = l["wd"] * math.pi / 180, l["ws"] * KT
            wl.append(((l["h"] - sfcH) / 1000, -s * math.sin(d), -s * math.cos(d)))

    def windAt(z):
        if z <= wl[0][0]:
            return wl[0][1], wl[0][2]
        for i in range(len(wl) - 1):
            if wl[i][0] <= z <= wl[i + 1][0]:
                f = (z - wl[i][0]) / max(1e-6, wl[i + 1][0] - wl[i][0])
                return wl[i][1] + (wl[i + 1][1] - wl[i][1]) * f, wl[i][2] + (wl[i + 1][2] - wl[i][2]) * f
        return wl[-1][1], wl[-1][2]

    A = {"snd": snd, "sfcH": sfcH, "windAt": windAt, "P": {}}
    if not parcels:
        return A
    grid = {"p": [], "h": [], "tv": []}
    p = sfc["p"]
    pend = max(100, lv[-1]["p"])
    while p >= pend:
        t, td, h = _interp_logp(lv, "t", p), _interp_logp(lv, "td", p), _interp_logp(lv, "h", p)
        if math.isfinite(t) and math.isfinite(h):
            grid["p"].append(p)
            grid["h"].append(h - sfcH)
            grid["tv"].append(_tvK(t + 273.15, mixr(p, td) if math.isfinite(td) else 0))
        p -= 2

    def hAt(pp):
        for k in range(len(grid["p"]) - 1):
            if grid["p"][k] >= pp >= grid["p"][k + 1]:
                f = math.log(grid["p"][k] / pp) / math.log(grid["p"][k] / grid["p"][k + 1])
                return grid["h"][k] + (grid["h"][k + 1] - grid["h"][k]) * f
        return 0 if pp > grid["p"][0] else grid["h"][-1]

    grid["hAt"] = hAt
    sb = _lift_parcel(grid, sfc["p"], sfc["t"], sfc["td"])
    thS = wS = 0.0
    n = 0
    p = sfc["p"]
    while p >= sfc["p"] - 100:
        t, td = _interp_logp(lv, "t", p), _interp_logp(lv, "td", p)
        if math.isfinite(t) and math.isfinite(td):
            thS += (t + 273.15) * (1000 / p) ** 0.2857
            wS += mixr(p, td)
            n += 1
        p -= 1
    thS /= n
    wS /= n
    mlT = thS * (sfc["p"] / 1000) ** 0.2857 - 273.15
    eML = wS * sfc["p"] / (EPS + wS)
    lnE = math.log(eML / 6.112)
    mlTd = 243.5 * lnE / (17.67 - lnE)
    ml = _lift_parcel(grid, sfc["p"], mlT, min(mlTd, mlT))
    best, bestTe = None, -1.0
    for l in lv:
        if l["p"] < sfc["p"] - 300 or not math.isfinite(l["t"]) or not math.isfinite(l["td"]):
            continue
        w = mixr(l["p"], l["td"])
        te = (l["t"] + 273.15) * (1000 / l["p"]) ** 0.2857 * math.exp(LV * w / (CP * (l["t"] + 273.15)))
        if te > bestTe:
            bestTe, best = te, l
    mu = _lift_parcel(grid, best["p"], best["t"], best["td"]) if best else sb
    A["P"] = {"sbcape": sb["cape"], "sbcin": sb["cin"], "sblcl": sb["lclH"], "mlcape": ml["cape"],
              "mlcin": ml["cin"], "mucape": mu["cape"]}
    return A


def base_state(snd: dict, nz: int, dz: float, motion=(0.0, 0.0), analysis: dict | None = None) -> dict:
    """baseState (sounding.js:143-197) -> {"B": float32[(nz+1)*8], "sfcP", "sfcH"}."""
    lv = sorted([l for l in snd["levels"] if math.isfinite(l["h"]) and math.isfinite(l["t"])], key=lambda l: l["h"])
    sfcH, sfcP = lv[0]["h"], lv[0]["p"]

    def byH(key, zAgl):
        h = zAgl + sfcH
        if h <= lv[0]["h"]:
            return lv[0][key]
        for i in range(len(lv) - 1):
            if h <= lv[i + 1]["h"]:
                a, b = lv[i], lv[i + 1]
                f = (h - a["h"]) / max(b["h"] - a["h"], 1e-3)
                va, vb = a[key], b[key]
                if not (math.isfinite(va) and math.isfinite(vb)):
                    return va if math.isfinite(va) else vb
                if key == "p":
                    return math.exp(math.log(va) + (math.log(vb) - math.log(va)) * f)
                return va + (vb - va) * f
        return float("nan")

    top = lv[-1]
    A = analysis or analyze(snd, parcels=False)
    B = np.zeros((nz + 1) * 8, dtype=np.float32)
    th, qv = [], []
    for k in range(nz):
        z = (k + 0.5) * dz
        T, Td, p = byH("t", z), byH("td", z), byH("p", z)
        if not (math.isfinite(T) and math.isfinite(p)):
            dzt = z + sfcH - top["h"]
            T = top["t"]
            p = top["p"] * math.exp(-G * dzt / (RD * (top["t"] + 273.15)))
            Td = min(top["td"], -80)
        if not math.isfinite(Td):
            Td = T - 30
        Td = min(Td, T - 0.3)
        th.append((T + 273.15) * (1000 / p) ** (RD / CP))
        qv.append(max(1e-7, mixr(p, Td)))
    piF = [0.0] * (nz + 1)
    piF[0] = (sfcP / 1000) ** (RD / CP)
    for k in range(nz):
        piF[k + 1] = piF[k] - G * dz / (CP * th[k] * (1 + 0.61 * qv[k]))
    for k in range(nz):
        pi = 0.5 * (piF[k] + piF[k + 1])
        p = 100000 * pi ** (CP / RD)
        T = th[k] * pi
        esx = 611.2 * math.exp(17.67 * (T - 273.15) / (T - 29.65))
        qvs = EPS * esx / (p - esx)
        qv[k] = min(qv[k], 0.97 * qvs)
        rho = p / (RD * T * (1 + 0.61 * qv[k]))
        w = A["windAt"]((k + 0.5) * dz / 1000)
        B[k * 8:k * 8 + 8] = [th[k], qv[k], rho, 0, pi, p, w[0] - motion[0], w[1] - motion[1]]
    for k in range(nz + 1):
        if k == 0:
            r = float(B[2]) + (float(B[2]) - float(B[10])) * 0.5
        elif k == nz:
            r = float(B[(nz - 1) * 8 + 2]) * 0.94
        else:
            r = 0.5 * (float(B[(k - 1) * 8 + 2]) + float(B[k * 8 + 2]))
        B[k * 8 + 3] = r
    return {"B": B, "sfcP": sfcP, "sfcH": sfcH, "analysis": A}


# ---------------------------------------------------------------------------------------------
# Analytic Weisman-Klemp (1982) sounding, written as SHARPpy text (no reference counterpart).
#   theta(z) = th0 + (thtr - th0) (z/ztr)^(5/4)            z <= ztr
#            = thtr exp(g (z - ztr) / (cp Ttr))             z >  ztr
#   RH(z)    = 1 - 3/4 (z/ztr)^(5/4)  (0.25 above ztr),   qv = min(RH qvs, qv0)
#   p from hydrostatic integration (virtual theta) from psfc.
# Winds: a quarter-circle hodograph of radius ``hodo_r`` over 0-2 km followed by a
# straight segment, so that the 0-6 km shear is ``shear6`` m/s.
# ---------------------------------------------------------------------------------------------
def wk_profile(z, th0=300.0, thtr=343.0, ztr=12000.0, Ttr=213.0, qv0=0.014, psfc=1000.0, dz=50.0):
    """theta, qv, p (hPa), T (K) at heights z (m AGL), integrating the hydrostatic equation upward."""
    zz = np.arange(0.0, float(np.max(z)) + dz, dz)

    def theta(h):
        return np.where(h <= ztr, th0 + (thtr - th0) * (h / ztr) ** 1.25, thtr * np.exp(9.81 * (h - ztr) / (1004.5 * Ttr)))

    def rh(h):
        return np.where(h <= ztr, 1.0 - 0.75 * (h / ztr) ** 1.25, 0.25)

    th = theta(zz)
    pi = np.empty_like(zz)
    qv = np.empty_like(zz)
    pi[0] = (psfc / 1000.0) ** (RD / CP)
    for i in range(len(zz)):
        if i > 0:
            thv = 0.5 * (th[i] * (1 + 0.61 * qv[i - 1]) + th[i - 1] * (1 + 0.61 * qv[i - 1]))
            pi[i] = pi[i - 1] - G * dz / (CP * thv)
        p = 1000.0 * pi[i] ** (CP / RD)
        T = th[i] * pi[i]
        qvs = mixr(p, T - 273.15)
        qv[i] = min(rh(zz[i]) * qvs, qv0)
    p = 1000.0 * pi ** (CP / RD)
    T = th * pi
    return (np.interp(z, zz, th), np.interp(z, zz, qv), np.exp(np.interp(z, zz, np.log(p))), np.interp(z, zz, T))


def dewpoint_from_qv(p_hpa, qv):
    e = qv * p_hpa / (EPS + qv)
    lnE = np.log(np.maximum(e, 1e-12) / 6.112)
    return 243.5 * lnE / (17.67 - lnE)


def hodograph(z, shear6=30.0, hodo_r=5.0, z_arc=2000.0, z_top=6000.0):
    """u, v (m/s): a quarter circle of radius hodo_r over 0..z_arc, then a straight segment
    to |V(z_top) - V(0)| = shear6, constant above z_top. Mean wind removed below 6 km."""
    z = np.asarray(z, dtype=float)
    # quarter circle starting at (-r, 0) turning clockwise to (0, r)
    ang = np.pi * (1 - 0.5 * np.clip(z / z_arc, 0, 1))
    u = hodo_r * np.cos(ang)
    v = hodo_r * np.sin(ang)
    # straight segment eastward from (0, r): |(L, r) - (-r, 0)| = shear6
    L = max(math.sqrt(max(shear6 ** 2 - hodo_r ** 2, 0.0)) - hodo_r, 0.0)
    f = np.clip((z - z_arc) / (z_top - z_arc), 0, 1)
    u = np.where(z > z_arc, L * f, u)
    v = np.where(z > z_arc, hodo_r, v)
    # remove the 0-6 km mean so storms stay in the domain
    zm = np.linspace(0, z_top, 121)
    am = np.pi * (1 - 0.5 * np.clip(zm / z_arc, 0, 1))
    um = np.where(zm > z_arc, L * np.clip((zm - z_arc) / (z_top - z_arc), 0, 1), hodo_r * np.cos(am))
    vm = np.where(zm > z_arc, hodo_r, hodo_r * np.sin(am))
    return u - um.mean(), v - vm.mean()


def wk_sounding_text(th0=300.0, thtr=343.0, ztr=12000.0, Ttr=213.0, qv0=0.014, shear6=30.0, hodo_r=5.0,
                     top=24000.0, dz=250.0, title="WK82 analytic") -> str:
    z = np.arange(0.0, top + dz, dz)
    th, qv, p, T = wk_profile(z, th0, thtr, ztr, Ttr, qv0)
    Td = dewpoint_from_qv(p, qv)
    u, v = hodograph(z, shear6, hodo_r)
    spd = np.hypot(u, v) / KT
    wdir = (np.degrees(np.arctan2(-u, -v)) + 360.0) % 360.0
    rows = ["%TITLE%", f" WKA   {title}", "", "   LEVEL       HGHT       TEMP       DWPT       WDIR       WSPD", "%RAW%"]
    for i in range(len(z)):
        rows.append(f" {p[i]:.3f}, {z[i]:.2f}, {T[i] - 273.15:.3f}, {Td[i]:.3f}, {wdir[i]:.3f}, {spd[i]:.3f}")
    rows.append("%END%")
    return "\n".join(rows) + "\n"
