"""The C ABI rejects malformed inputs before any kernel runs (counts that index fixed-size device tables, short
host buffers), with a status and a last-error text, as the reference's allocate/set* throw."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small():
    from outbreaklab_amd import RegionalModel
    from outbreaklab_amd.sounding import wk_sounding_text

    m = RegionalModel()
    m.allocate("custom", {"key": "custom:2000:32x32x10", "dx": 2000, "n": [32, 32, 10]})
    m.setSounding(wk_sounding_text())
    m.setBoundaries([])
    m.setCI([])
    m.setSurface(None, None)
    m.reset()
    yield m
    m.destroy()


def _step_error(m, u):
    from outbreaklab_amd import _lib

    u = np.ascontiguousarray(u, np.float32)
    rc = _lib.lib().obl_step(m.h, u.ctypes.data_as(C.c_void_p), m.parity, 0)
    return rc, _lib.lib().obl_last_error(m.h).decode()


def test_uniform_counts_out_of_range_are_rejected(small):
    m = small
    u = m.writeUniforms(6.0)
    for idx, bad, word in ((30, 9.0, "cor.z"), (30, -1.0, "cor.z"), (32, 257.0, "ci.x"), (32, float("nan"), "ci.x")):
        v = u.copy()
        v[idx] = bad
        rc, msg = _step_error(m, v)
        assert rc != 0 and word in msg, (idx, bad, msg)
    rc, _ = _step_error(m, u)
    assert rc == 0
    m.sync()


def test_boundary_records_and_short_buffers_are_rejected(small):
    from outbreaklab_amd import _lib

    m = small
    L = _lib.lib()
    rec = np.zeros(148, np.float32)
    rec[0] = 33
    assert L.obl_set_boundaries(m.h, rec.ctypes.data_as(C.c_void_p), 1) != 0
    assert "points" in L.obl_last_error(m.h).decode()
    rec[0] = 2
    assert L.obl_set_boundaries(m.h, rec.ctypes.data_as(C.c_void_p), 1) == 0
    assert L.obl_set_boundaries(m.h, None, 0) == 0
    surf = np.zeros(m.nx * m.ny * 4 - 1, np.float32)
    assert L.obl_set_surface(m.h, surf.ctypes.data_as(C.c_void_p), surf.nbytes) != 0
    assert "setSurface" in L.obl_last_error(m.h).decode()
    tab = np.zeros(m.nz * 8 * 4 - 4, np.float32)
    assert L.obl_set_vortex(m.h, tab.ctypes.data_as(C.c_void_p), 8, tab.nbytes) != 0
    assert "setVortex" in L.obl_last_error(m.h).decode()
