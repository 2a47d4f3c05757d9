"""Virtual WSR-88D radars on the MI355X model state: the host side of the reference's RadarSim
(radar.js:271-418) over the C ABI's radar calls, and the azimuthal-equidistant projection it uses
(geo.js:6-29).

The VCP tables (radar.js:24-35), the volume layout (cuts x moments, radial-major, 8-bit Archive II
codes; radar.js:289-313), the scan schedule (radar.js:314-316, 344-367) and the RADAR_FIELDS block
(radar.js:37-39, 369-382) follow the reference; the GPU passes are `k_radar_cut` / `k_radar_phi`.
The reference draws each site's volume phase and each volume's first azimuth from Math.random; with
`opts["seed"]` they come from the reference's own seeded generator (rng.js:3-27), so runs reproduce and
the Node host draws the same numbers.
"""
from __future__ import annotations

import ctypes as C
import gzip
import json
import math
import random
import struct

import numpy as np

from . import _lib
from .uniforms import UBlock

RE_M, KE = 6371000.0, 4.0 / 3.0
GATE0, GATEDR, MAX_REF_GATES, MAX_DOP_GATES = 2125, 250, 1832, 1192
L2_SCALE = {"REF": (2, 66), "VEL": (2, 129), "SW": (2, 129), "ZDR": (16, 128), "PHI": (253 / 360, 2), "RHO": (300, -60.5)}
CUT_MOMENTS = {"surv": ["REF", "ZDR", "PHI", "RHO"], "dop": ["REF", "VEL", "SW"],
               "both": ["REF", "VEL", "SW", "ZDR", "PHI", "RHO"]}
MOMENT_BITS = {n: 1 << i for i, n in enumerate(["REF", "VEL", "SW", "ZDR", "PHI", "RHO"])}
RADAR_FIELDS = ["site", "geo", "cut", "gates", "nyq", "offA", "offB", "mdl", "mdl2", "phys", "phys2"]
DEG = math.pi / 180


_M32 = 0xFFFFFFFF


def _imul(a, b):
    return (a * b) & _M32


class Rng:
    """The reference's seeded generator (rng.js:3-27): the xmur3 hash of the seed's decimal text gives
    the state of a mulberry32 stream; ``u()`` is uniform on [0, 1) in steps of 2^-32."""

    def __init__(self, seed):
        text = str(seed)
        h = (1779033703 ^ len(text)) & _M32
        for ch in text:
            h = _imul(h ^ ord(ch), 3432918353)
            h = ((h << 13) | (h >> 19)) & _M32
        h = _imul(h ^ (h >> 16), 2246822507)
        h = _imul(h ^ (h >> 13), 3266489909)
        self.state = (h ^ (h >> 16)) & _M32

    def u(self):
        self.state = (self.state + 0x6D2B79F5) & _M32
        t = self.state
        t = _imul(t ^ (t >> 15), t | 1)
        t ^= (t + _imul(t ^ (t >> 7), t | 61)) & _M32
        return ((t ^ (t >> 14)) & _M32) / 4294967296.0


def _jsround(x):
    """Math.round: halves go up."""
    return int(math.floor(x + 0.5))


def split_vcp(split, rest, nyq_split, nyq_rest):
    """splitVcp (radar.js:24-29): split cuts (surveillance + Doppler) then the batch cuts."""
    cuts = []
    for a in split:
        cuts.append({"angle": a, "kind": "surv", "superRes": True, "nyquist": 8.9})
        cuts.append({"angle": a, "kind": "dop", "superRes": True, "nyquist": nyq_split})
    for a in rest:
        cuts.append({"angle": a, "kind": "both", "superRes": False, "nyquist": nyq_rest if a < 7 else 32.3})
    return cuts


VCPS = {
    212: {"label": "VCP 212", "dur": 270, "cuts": split_vcp([0.5, 0.9, 1.3], [1.8, 2.4, 3.1, 4.0, 5.1, 6.4, 8.0, 10.0, 12.5, 15.6, 19.5], 28.4, 29.4)},
    12: {"label": "VCP 12", "dur": 250, "cuts": split_vcp([0.5, 0.9, 1.3], [1.8, 2.4, 3.1, 4.0, 5.1, 6.4, 8.0, 10.0, 12.5, 15.6, 19.5], 28.4, 29.4)},
    215: {"label": "VCP 215", "dur": 360, "cuts": split_vcp([0.5, 0.9, 1.3], [1.8, 2.4, 3.1, 4.0, 5.1, 6.4, 8.0, 10.0, 12.0, 14.0, 16.7, 19.5], 28.4, 29.4)},
    35: {"label": "VCP 35", "dur": 420, "cuts": split_vcp([0.5, 0.9, 1.3], [1.8, 2.4, 3.1, 4.0, 5.1, 6.4], 26.0, 27.0)},
}


class Geo:
    """Azimuthal-equidistant projection about the domain centre (geo.js:6-29), metres."""

    def __init__(self, lat0, lon0, Lx=0.0, Ly=0.0, groundMSL=0.0):
        self.lat0, self.lon0, self.Lx, self.Ly, self.groundMSL = lat0, lon0, Lx, Ly, groundMSL
        self.p0, self.l0 = lat0 * DEG, lon0 * DEG
        self.sp0, self.cp0 = math.sin(self.p0), math.cos(self.p0)

    def fwd(self, lat, lon):
        p, dl = lat * DEG, lon * DEG - self.l0
        sp, cp, cdl = math.sin(p), math.cos(p), math.cos(dl)
        cc = max(-1.0, min(1.0, self.sp0 * sp + self.cp0 * cp * cdl))
        c = math.acos(cc)
        k = 1.0 if c < 1e-9 else c / math.sin(c)
        return RE_M * k * cp * math.sin(dl), RE_M * k * (self.cp0 * sp - self.sp0 * cp * cdl)

    def inv(self, x, y):
        rho = math.hypot(x, y)
        if rho < 1e-6:
            return self.lat0, self.lon0
        c = rho / RE_M
        sc, ccs = math.sin(c), math.cos(c)
        p = math.asin(ccs * self.sp0 + y * sc * self.cp0 / rho)
        lon = self.l0 + math.atan2(x * sc, rho * self.cp0 * ccs - y * self.sp0 * sc)
        return p / DEG, ((lon / DEG + 540) % 360) - 180

    def toModel(self, lat, lon):
        x, y = self.fwd(lat, lon)
        return x + self.Lx / 2, y + self.Ly / 2

    def fromModel(self, mx, my):
        return self.inv(mx - self.Lx / 2, my - self.Ly / 2)


def plan_radar(m, geo, site, opts=None):
    """addRadar's layout (radar.js:289-316) without the device buffers: gate counts from the reach of
    the model, each cut's moments at byte offsets (radial-major), the scan time of each cut."""
    opts = dict(opts or {})
    vcp = VCPS.get(opts.get("vcp", 212), VCPS[212])
    sx, sy = geo.toModel(site["lat"], site["lon"])
    far = max(math.hypot(cx - sx, cy - sy) for cx, cy in ((0, 0), (m.Lx, 0), (0, m.Ly), (m.Lx, m.Ly)))
    rng = min(460000, far + 10000)
    ng_ref = min(MAX_REF_GATES, math.ceil((rng - GATE0) / GATEDR))
    ng_ref = max(64, math.ceil(ng_ref / 4) * 4)
    ng_dop = min(MAX_DOP_GATES, ng_ref)
    off, cuts = 0, []
    for ci, c in enumerate(vcp["cuts"]):
        nrad = 720 if c["superRes"] else 360
        mom = []
        for name in CUT_MOMENTS[c["kind"]]:
            ng = ng_ref if name == "REF" else ng_dop
            mom.append({"name": name, "ngates": ng, "off": off})
            off += nrad * ng
        cuts.append({**c, "index": ci, "nrad": nrad, "mom": mom})
    wsum = sum(2 if c["superRes"] else 1 for c in cuts)
    acc = 0
    for c in cuts:
        w = 2 if c["superRes"] else 1
        c["t0"] = acc / wsum * vcp["dur"]
        acc += w
        c["dur"] = w / wsum * vcp["dur"]
    rand = Rng(opts["seed"]).u if "seed" in opts else random.random
    return {"site": site, "opts": opts, "vcp": vcp, "vcpId": int(opts.get("vcp", 212)), "cuts": cuts, "bytes": off,
            "ngRef": ng_ref, "ngDop": ng_dop, "sx": sx, "sy": sy, "h": None, "U": UBlock(RADAR_FIELDS), "rand": rand,
            "phase": rand() * vcp["dur"], "vol": -1, "next": 0, "volStart": 0.0, "az0": 0.0, "busy": [False, False]}


def cut_block(m, geo, R, c, simT, n_tor=0):
    """The cut's RADAR_FIELDS block (radar.js:369-382): 44 float32."""
    U = R["U"]
    off = {o["name"]: o["off"] // 4 for o in c["mom"]}
    mask = sum(MOMENT_BITS[n] for n in off)
    z_ant = R["site"]["h"] + 20
    nyq = 63.5 if R["opts"].get("aliasing") is False else c["nyquist"]
    U.set("site", R["sx"], R["sy"], z_ant, KE * RE_M)
    U.set("geo", geo.lat0 * DEG, geo.lon0 * DEG, R["site"]["lat"] * DEG, R["site"]["lon"] * DEG)
    U.set("cut", c["angle"] * DEG, R["az0"], 360 / c["nrad"], c["nrad"])
    U.set("gates", R["ngRef"], R["ngDop"], GATE0, GATEDR)
    U.set("nyq", nyq, simT, (R["vol"] * 97 + c["index"] * 13) % 100000, 25)
    U.set("offA", off.get("REF", 0), off.get("VEL", 0), off.get("SW", 0), off.get("ZDR", 0))
    U.set("offB", off.get("PHI", 0), off.get("RHO", 0), 0, mask)
    U.set("mdl", m.nx, m.ny, m.nz, m.dx)
    U.set("mdl2", m.dz, geo.groundMSL, -m.Lx / 2, -m.Ly / 2)
    U.set("phys", m.rhoSfc or 1.15, R["opts"].get("minDbz", 0), 1 if m.params["scheme"] == "thompson" else 0, n_tor)
    hur = getattr(m, "hur", None)
    U.set("phys2", 1 if hur and hur.get("maritime") else 0, 0, 0, 0)  # radar.js:384
    return U.f.copy()


def gate_geometry(m, ru, rg):
    """The GPU's geometry of listed gates of a cut (obl_radar_geometry, a diagnostic): ru = the cut's block,
    rg = (n, 2) (radial, gate); returns (n, 7) float32 = x, y, ex, ey, beam-centre height, heights below / above."""
    rg = np.ascontiguousarray(rg, np.int32).reshape(-1, 2)
    ru = np.ascontiguousarray(ru, np.float32)
    out = np.zeros((len(rg), 7), np.float32)
    _lib.check(_lib.lib().obl_radar_geometry(m.h, ru.ctypes.data_as(C.c_void_p), len(rg), rg.ctypes.data_as(C.c_void_p),
                                             out.ctypes.data_as(C.c_void_p)), m.h)
    return out


class RadarSim:
    """RadarSim (radar.js:271-418) over `obl_radar_*`."""

    def __init__(self, model, geo):
        self.m, self.geo = model, geo
        self.radars = []
        self.nTor = 0

    def setTornadoes(self, lst):
        act = list(lst)[:64]
        f = np.zeros(max(1, len(act)) * 8, np.float32)
        for i, t in enumerate(act):
            f[i * 8:i * 8 + 8] = [t["x"], t["y"], t["R"], t["v"], t["debrisTop"], t["debrisDbz"], t["debris"], 0]
        _lib.check(_lib.lib().obl_radar_set_tornadoes(self.m.h, f.ctypes.data_as(C.c_void_p), len(act)), self.m.h)
        self.nTor = len(act)

    def addRadar(self, site, opts=None):
        R = plan_radar(self.m, self.geo, site, opts)
        h = C.c_void_p()
        _lib.check(_lib.lib().obl_radar_create(self.m.h, R["bytes"], 720 * R["ngDop"], C.byref(h)), self.m.h)
        R["h"] = h
        self.radars.append(R)
        return R

    def destroy(self):
        # the model frees its radars itself when it is destroyed first
        if self.m.h is not None:
            for R in self.radars:
                _lib.lib().obl_radar_destroy(R["h"])
        self.radars = []

    def encDue(self, enc, simT):
        """encDue (radar.js:344-367): encode every cut whose scan time has come; returns finished volumes."""
        done = []
        for R in self.radars:
            while True:
                if R["vol"] < 0 or R["next"] >= len(R["cuts"]):
                    vstart = R["phase"] if R["vol"] < 0 else R["volStart"] + R["vcp"]["dur"]
                    if simT < vstart or (R["busy"][0] and R["busy"][1]):
                        break
                    R["vol"] += 1
                    R["volStart"], R["next"], R["az0"] = vstart, 0, R["rand"]()
                c = R["cuts"][R["next"]]
                if simT < R["volStart"] + c["t0"]:
                    break
                self.encCut(enc, R, c, simT)
                R["next"] += 1
                if R["next"] >= len(R["cuts"]):
                    slot = C.c_int(-1)
                    _lib.check(_lib.lib().obl_radar_readback(R["h"], C.byref(slot)), self.m.h)
                    if slot.value < 0:
                        raise _lib.OblError(f"radar {R['site']['id']}: both readback slots are pending")
                    R["busy"][slot.value] = True
                    done.append({"R": R, "slot": slot.value, "vol": R["vol"], "volStart": R["volStart"], "az0": R["az0"]})
        return done

    def encCut(self, enc, R, c, simT):
        ru = cut_block(self.m, self.geo, R, c, simT, self.nTor)
        _lib.check(_lib.lib().obl_radar_cut(R["h"], ru.ctypes.data_as(C.c_void_p), self.m.parity), self.m.h)
        c["simT"] = simT

    def collect(self, d, run="run", epoch0=0.0):
        """collect (radar.js:394-416): the finished volume's bytes and the header the server reads."""
        R = d["R"]
        buf = np.empty(R["bytes"], np.uint8)
        _lib.check(_lib.lib().obl_radar_collect(R["h"], d["slot"], buf.ctypes.data_as(C.c_void_p), buf.nbytes), self.m.h)
        R["busy"][d["slot"]] = False
        t0 = epoch0 + d["volStart"] * 1000
        nyq = lambda c: 63.5 if R["opts"].get("aliasing") is False else c["nyquist"]  # noqa: E731
        header = {"run": run, "icao": R["site"]["id"], "lat": R["site"]["lat"], "lon": R["site"]["lon"],
                  "elev": R["site"]["h"] + 20, "vcp": R["vcpId"], "epoch_ms": _jsround(t0),
                  "sweeps": [{"angle": c["angle"], "superRes": c["superRes"], "kind": c["kind"], "az0": d["az0"],
                              "nrad": c["nrad"], "nyquist": nyq(c), "unamb_km": 466 if c["kind"] == "surv" else 300,
                              "start_ms": _jsround(t0 + c["t0"] * 1000), "dur_ms": _jsround(c["dur"] * 1000),
                              "moments": [{"name": o["name"], "ngates": o["ngates"], "off": o["off"]} for o in c["mom"]]}
                             for c in R["cuts"]]}
        return {"R": R, "header": header, "bytes": buf}


# ---------------- the volume upload (app.js:1735-1750) ----------------
def volume_body(header, codes, compress=True):
    """The /api/volume request body for a collected volume (app.js:1743-1750): the header's JSON length as a
    little-endian u32, the JSON (UTF-8), then the volume's codes, gzip-compressed (sent with X-OB-Gzip: 1)."""
    hj = json.dumps(header, separators=(",", ":"), ensure_ascii=False).encode("utf-8")
    raw = struct.pack("<I", len(hj)) + hj + np.ascontiguousarray(codes, np.uint8).tobytes()
    return gzip.compress(raw, compresslevel=6, mtime=0) if compress else raw


def level2_sweeps(header, codes):
    """The arguments the Archive II writer takes for a collected volume, as the server builds them from an
    upload (server.py:256-266): (icao, lat, lon, elev, epoch_ms, sweeps, vcp), each sweep's moments as
    {name: (codes of nrad x ngates, ngates)}."""
    b = memoryview(np.ascontiguousarray(codes, np.uint8).tobytes())
    sweeps = []
    for s in header["sweeps"]:
        mom = {m["name"]: (b[m["off"]:m["off"] + int(s["nrad"]) * int(m["ngates"])], int(m["ngates"]))
               for m in s["moments"]}
        sweeps.append({**s, "moments": mom})
    return (header["icao"].upper(), float(header["lat"]), float(header["lon"]), float(header["elev"]),
            int(header["epoch_ms"]), sweeps, int(header.get("vcp", 212)))
