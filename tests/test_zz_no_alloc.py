"""obl_radar_cut_into, the allocation-free radar cut (its KDP scratch is a fixed per-GPU device array), against
obl_radar_cut (a radar handle's own volume and scratch): bit-identical Archive II codes, on two calls in a row, the
second a smaller cut (fewer radials), on the 128 x 128 x 40 Thompson workload after 15 minutes."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _hip():
    """The HIP runtime the library is linked to (already loaded in this process): the test's own device buffers."""
    return C.CDLL("libamdhip64.so.7")


@pytest.fixture(scope="module")
def radar():
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.configs import setup_plain_128
    from outbreaklab_amd.radar import Geo, RadarSim

    m = RegionalModel()
    dt = setup_plain_128(m, scheme="thompson")
    for _ in range(150):
        m.step(dt)
    m.sync()
    geo = Geo(35.9, -97.0, m.Lx, m.Ly, 0.0)
    sim = RadarSim(m, geo)
    sim.setTornadoes([{"x": 0.5 * m.Lx, "y": 0.5 * m.Ly, "R": 150.0, "v": 60.0, "debrisTop": 1200.0,
                       "debrisDbz": 55.0, "debris": 1.0}])
    lat, lon = geo.fromModel(0.5 * m.Lx - 40e3, 0.5 * m.Ly - 30e3)
    R = sim.addRadar({"id": "KTST", "lat": lat, "lon": lon, "h": 0.0}, {"vcp": 212, "seed": 5})
    R["phase"] = 0.0
    done = sim.encDue(None, 269.0)
    assert len(done) == 1
    vol = sim.collect(done[0], "ref", 0.0)
    yield m, geo, R, vol["bytes"]
    sim.destroy()
    m.destroy()


def _cut_words(c, ref_bytes):
    """(first word, word count) of each moment of cut c in the volume."""
    return [(o["off"] // 4, c["nrad"] * o["ngates"] // 4) for o in c["mom"]]


def test_no_alloc_cut_equals_the_radar_handles_cut(radar):
    from outbreaklab_amd import _lib
    from outbreaklab_amd.radar import cut_block

    m, geo, R, ref = radar
    hip = _hip()
    L = _lib.lib()
    nbytes = R["bytes"]
    dev = C.c_void_p()
    assert hip.hipMalloc(C.byref(dev), C.c_size_t(nbytes)) == 0
    try:
        ref_w = ref.view(np.uint32)
        # a super-resolution cut (720 radials, all six moments), then a smaller one (360 radials)
        cuts = [R["cuts"][0], R["cuts"][1], R["cuts"][6]]
        assert cuts[2]["nrad"] == 360 and cuts[0]["nrad"] == 720
        for c in cuts:
            print(f"cut {c['index']}: {c['nrad']} radials, moments {[o['name'] for o in c['mom']]}", flush=True)
            ru = np.ascontiguousarray(cut_block(m, geo, R, c, c["simT"], 1), np.float32)
            rc = L.obl_radar_cut_into(m.h, ru.ctypes.data_as(C.c_void_p), m.parity, dev, C.c_size_t(nbytes))
            assert rc == 0, L.obl_last_error(m.h).decode()
            m.sync()
            got = np.empty(nbytes // 4, np.uint32)
            assert hip.hipMemcpy(got.ctypes.data_as(C.c_void_p), dev, C.c_size_t(nbytes), 2) == 0
            for w0, n in _cut_words(c, ref):
                assert np.array_equal(got[w0:w0 + n], ref_w[w0:w0 + n]), (c["index"], w0)
            assert np.count_nonzero(got[cuts[0]["mom"][0]["off"] // 4:][:100000]) > 0
        # a cut whose Doppler radials exceed the fixed scratch is refused
        big = np.ascontiguousarray(cut_block(m, geo, R, cuts[1], 0.0, 1), np.float32).copy()
        big[11] = 1200.0  # cut.w: radials
        assert L.obl_radar_cut_into(m.h, big.ctypes.data_as(C.c_void_p), m.parity, dev, C.c_size_t(1 << 30)) != 0
        assert "scratch" in L.obl_last_error(m.h).decode()
    finally:
        hip.hipFree(dev)
