"""The Level II wire and disk formats of a radar volume (CPU).

- The upload body (app.js:1743-1750), built by either host, parses back as server.py:249-266 parses it
  (gunzip, u32-LE header length, JSON, per-moment slices) to the volume's header and codes.
- A volume in this build's layout and header, passed through the reference's own Archive II writer
  (l2writer.py:161-182, run once to make tests/golden/l2_small.ar2), reads back -- with a reader written from
  the published message-31 format (tests/_archive2.py) -- to the same codes, azimuths, elevations, times,
  Nyquist velocities, scales and site. The writer's scales and offsets equal the radar host's Archive II
  code table (L2_SCALE), so the codes the kernels quantise are the codes the files declare.
"""
import json
import os
import shutil
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

from _archive2 import read_volume, sweep_codes
from outbreaklab_amd.radar import L2_SCALE, level2_sweeps, volume_body

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLD)
from l2_volume import digest, synthetic_volume  # noqa: E402


@pytest.fixture(scope="module")
def vol():
    header, codes = synthetic_volume()
    meta = json.load(open(os.path.join(GOLD, "l2_small.json")))
    data = open(os.path.join(GOLD, "l2_small.ar2"), "rb").read()
    return header, codes, meta, data


def server_parse(body):
    """server.py:250-266: gunzip, header length, JSON, each moment's nrad x ngates slice of the payload."""
    body = zlib.decompress(body, 16 + zlib.MAX_WBITS)
    (hl,) = struct.unpack("<I", body[:4])
    H = json.loads(body[4:4 + hl])
    payload = memoryview(body)[4 + hl:]
    out = []
    for s in H["sweeps"]:
        out.append({m["name"]: np.frombuffer(payload[m["off"]:m["off"] + int(s["nrad"]) * int(m["ngates"])], np.uint8)
                    .reshape(int(s["nrad"]), int(m["ngates"])) for m in s["moments"]})
    return H, out


def moments_of(header, codes):
    return [{m["name"]: codes[m["off"]:m["off"] + s["nrad"] * m["ngates"]].reshape(s["nrad"], m["ngates"])
             for m in s["moments"]} for s in header["sweeps"]]


def test_fixture_is_this_layout(vol):
    header, codes, meta, data = vol
    assert digest(codes) == meta["codes_sha256"]
    assert header == meta["header"] and len(data) == meta["bytes"]


def test_upload_body_round_trip(vol):
    header, codes, _, _ = vol
    H, moms = server_parse(volume_body(header, codes))
    assert H == header
    for a, b in zip(moms, moments_of(header, codes)):
        assert a.keys() == b.keys() and all(np.array_equal(a[k], b[k]) for k in a)
    # the writer's arguments as the server builds them
    icao, lat, lon, elev, epoch, sweeps, vcp = level2_sweeps(header, codes)
    assert (icao, lat, lon, elev, epoch, vcp) == ("KTLX", 35.333, -97.278, 390.0, header["epoch_ms"], 35)
    for s, b in zip(sweeps, moments_of(header, codes)):
        for k, (mv, ng) in s["moments"].items():
            assert np.array_equal(np.frombuffer(mv, np.uint8).reshape(-1, ng), b[k])


@pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(os.path.join(ROOT, "outbreaklab_amd", "js", "obl_napi.node")),
                    reason="node or the built addon is missing")
def test_node_upload_body_round_trip(vol, tmp_path):
    header, codes, _, _ = vol
    hp, cp, op = tmp_path / "h.json", tmp_path / "c.bin", tmp_path / "body.gz"
    hp.write_text(json.dumps(header))
    cp.write_bytes(codes.tobytes())
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js_body_check.js"), str(hp), str(cp), str(op)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    H, moms = server_parse(op.read_bytes())
    assert H == header
    for a, b in zip(moms, moments_of(header, codes)):
        assert all(np.array_equal(a[k], b[k]) for k in b)


def test_archive2_reads_back(vol):
    header, codes, meta, data = vol
    V = read_volume(data)
    ep = header["epoch_ms"]
    assert V["tape"] == "AR2V0006.001" and V["icao"] == "KTLX"
    assert (V["jd"], V["jms"]) == (ep // 86400000 + 1, ep % 86400000)
    # metadata record: 131 empty frames, then messages 3, 5 and 2; message 5 lists the three tilts
    types = [t for t, _ in V["other"]]
    assert types[:131] == [0] * 131 and types[131:134] == [3, 5, 2]
    ntilt, = struct.unpack_from(">H", V["msg5"], 28 + 6)
    vcp5, = struct.unpack_from(">H", V["msg5"], 28 + 4)
    assert ntilt == 3 and vcp5 == 35
    for k, s in enumerate(header["sweeps"]):
        ang = struct.unpack_from(">H", V["msg5"], 28 + 22 + 46 * k)[0]
        assert ang == round(s["angle"] * 65536 / 360) & 0xFFFF
    assert sorted(V["sweeps"]) == [1, 2, 3]
    nsw = len(header["sweeps"])
    for ei, s in enumerate(header["sweeps"]):
        rads = sorted(V["sweeps"][ei + 1], key=lambda r: r["num"])
        nrad = s["nrad"]
        assert [r["num"] for r in rads] == list(range(1, nrad + 1))
        ri = np.arange(nrad)
        az = np.array([r["az"] for r in rads], np.float32)
        assert np.array_equal(az, ((s["az0"] + ri * (360.0 / nrad)) % 360.0).astype(np.float32))
        assert all(r["el"] == np.float32(s["angle"]) for r in rads)
        assert all(r["az_spacing"] == (1 if s["superRes"] else 2) for r in rads)
        tms = s["start_ms"] + np.rint(s["dur_ms"] * ri / nrad).astype(np.int64)
        assert [r["ms"] for r in rads] == list(tms % 86400000)
        st = [r["status"] for r in rads]
        assert st[0] == (3 if ei == 0 else 0) and st[-1] == (4 if ei == nsw - 1 else 2) and set(st[1:-1]) == {1}
        rr = rads[0]["rrad"]
        assert rr["nyquist_cms"] == round(s["nyquist"] * 100) and rr["unamb_100m"] == round(s["unamb_km"] * 10)
        rv = rads[0]["rvol"]
        assert rv["lat"] == np.float32(header["lat"]) and rv["lon"] == np.float32(header["lon"])
        assert rv["height"] == round(header["elev"]) and rv["vcp"] == 35
        ref = moments_of(header, codes)[ei]
        assert set(rads[0]["moments"]) == set(ref)
        for name, c in ref.items():
            got = sweep_codes(rads, name)
            assert np.array_equal(got, c), (ei, name)
            md = rads[0]["moments"][name]
            assert (md["ngates"], md["first"], md["dr"]) == (c.shape[1], 2125, 250)
            sc, of = L2_SCALE[name]
            assert md["scale"] == np.float32(sc) and md["offset"] == np.float32(of), name
