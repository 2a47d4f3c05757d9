"""Reader of NEXRAD Archive II volumes (message 31 radials in bz2 LDM records) -- test infrastructure.

Written from the published format (NOAA RDA/RPG ICD 2620002 and 2620010), not from any writer:
- volume header: "AR2V00xx.xxx", Julian date (>i), milliseconds of day (>i), ICAO (4 bytes), 24 bytes;
- LDM records: a big-endian i32 byte count (negative on the last record) then a bz2 stream;
- a record holds messages, each a 12-byte channel-terminal-manager prefix and a 16-byte header
  (size in halfwords including the header, channel, type, sequence, Julian date, ms of day, segments,
  segment); messages other than 31 occupy fixed 2432-byte frames;
- message 31: data header (ICAO, ms, date, radial number, azimuth f32, compression, spare, radial length,
  azimuth spacing code, radial status, elevation number, cut sector, elevation f32, spot blanking,
  azimuth indexing, block count, block pointers); blocks "RVOL" (size, version, lat, lon, site height,
  feedhorn height, calibration, H / V transmit power, ZDR and PHIDP calibration, VCP, status), "RELV",
  "RRAD" (size, unambiguous range in 100 m, H / V noise, Nyquist in cm/s, spare, H / V calibration) and
  moment blocks "D<name>" (reserved u32, gate count, first gate m, gate spacing m, threshold, SNR threshold, control
  flags, word size, scale f32, offset f32, then the gate words).
"""
from __future__ import annotations

import bz2
import struct

import numpy as np

FRAME = 2432


def _messages(rec: bytes):
    o = 0
    while o + 28 <= len(rec):
        size_hw, chan, mtype, seq, jd, jms, nseg, seg = struct.unpack_from(">HBBHHIHH", rec, o + 12)
        if mtype == 31:
            n = 12 + 2 * size_hw
            yield mtype, seq, jd, jms, rec[o:o + n]
            o += n
        else:
            yield mtype, seq, jd, jms, rec[o:o + FRAME]
            o += FRAME


def read_volume(data: bytes) -> dict:
    tape = data[:12].decode("ascii")
    jd, jms = struct.unpack_from(">ii", data, 12)
    icao = data[20:24].decode("ascii")
    o, records = 24, []
    while o < len(data):
        (n,) = struct.unpack_from(">i", data, o)
        o += 4
        records.append(bz2.decompress(data[o:o + abs(n)]))
        o += abs(n)
        if n < 0:
            break
    vol = {"tape": tape, "jd": jd, "jms": jms, "icao": icao, "records": len(records), "other": [], "sweeps": {}}
    for rec in records:
        for mtype, seq, mjd, mjms, msg in _messages(rec):
            if mtype != 31:
                vol["other"].append((mtype, seq))
                if mtype == 5:
                    vol["msg5"] = msg
                continue
            d = 28
            r_icao = msg[d:d + 4].decode("ascii")
            (r_ms, r_jd, r_num, az, comp, _sp, r_len, az_sp, status, el_num, cut_sec, el, spot, az_idx,
             nblk) = struct.unpack_from(">IHHfBBHBBBBfBBH", msg, d + 4)
            ptrs = struct.unpack_from(">" + "I" * nblk, msg, d + 32)
            rad = {"icao": r_icao, "ms": r_ms, "jd": r_jd, "num": r_num, "az": az, "status": status, "el": el,
                   "az_spacing": az_sp, "len": r_len, "seq": seq, "hdr_jd": mjd, "hdr_jms": mjms, "moments": {}}
            for p in ptrs:
                b = d + p
                name = msg[b:b + 4].decode("ascii")
                if name == "RVOL":
                    (ln, vmaj, vmin, lat, lon, hgt, fh, calib, tx_h, tx_v, zdr_cal, phi_init, vcp,
                     _s) = struct.unpack_from(">HBBffhHfffffHH", msg, b + 4)
                    rad["rvol"] = {"lat": lat, "lon": lon, "height": hgt, "feedhorn": fh, "vcp": vcp}
                elif name == "RELV":
                    rad["relv"] = struct.unpack_from(">Hhf", msg, b + 4)
                elif name == "RRAD":
                    (ln, unamb, nh, nv, nyq, _s, ch, cv) = struct.unpack_from(">HhffhHff", msg, b + 4)
                    rad["rrad"] = {"unamb_100m": unamb, "nyquist_cms": nyq}
                elif name[0] == "D":
                    (_r, ng, g0, gdr, thr, snr, flags, wsize, scale, offset) = struct.unpack_from(">IHHHHhBBff", msg, b + 4)
                    assert wsize == 8
                    rad["moments"][name[1:].strip()] = {
                        "ngates": ng, "first": g0, "dr": gdr, "scale": scale, "offset": offset,
                        "codes": np.frombuffer(msg, np.uint8, ng, b + 28)}
                else:
                    raise ValueError(f"unknown data block {name!r}")
            vol["sweeps"].setdefault(el_num, []).append(rad)
    return vol


def sweep_codes(sw: list, name: str) -> np.ndarray:
    """(nrad, ngates) codes of one moment of a sweep, radials in radial-number order."""
    rs = sorted(sw, key=lambda r: r["num"])
    return np.stack([r["moments"][name]["codes"] for r in rs])
