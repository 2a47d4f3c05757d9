"""The Node N-API facade (outbreaklab_amd/js): loads under the image's Node, reproduces the
reference base state and packs the same uniform block as the Python host (CPU only)."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from _cases import host_model, uniforms

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "outbreaklab_amd", "js", "obl_napi.node")

pytestmark = pytest.mark.skipif(shutil.which("node") is None or not os.path.exists(ADDON),
                                reason="node or the built addon is missing")


def _run():
    r = subprocess.run(["node", os.path.join(ROOT, "tests", "js_facade_check.js")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_facade_loads_and_matches_golden():
    out = _run()
    for name in ["create", "step", "reset", "download", "upload", "sync", "poissonSolve", "fieldPtr", "setBaseState",
                 "sampleMicro", "microSamples", "setSurrogate", "surrogateStats", "mlpCreate", "mlpTrain", "mlpInfer",
                 "mapsRequest", "mapsCollect"]:
        assert name in out["exports"]
    assert out["mismatches"] == 0
    assert out["levels"] == [56, 80, 100] and out["pads"] == [8, 24, 40]
    assert "multiples of 8" in out["poissonError"]


def test_facade_uniforms_equal_python_host():
    out = _run()
    m = host_model(128, 128, 40, 2000, scheme="thompson", text=open(os.path.join(ROOT, "tests", "golden",
                                                                                  "wk_sounding.txt")).read())
    m.time = 42.0
    u = uniforms(m, 6.0)
    assert np.array_equal(np.array(out["uniforms"], np.float32), u)


def test_facade_surrogate_targets_equal_python_host():
    from outbreaklab_amd.surrogate import targets

    out = _run()
    rows = np.zeros((2, 24), np.float32)
    rows[0, 9], rows[0, 21], rows[1, 9], rows[1, 21] = 0, 3e4, 5e3, -4e3
    assert np.array_equal(np.array(out["targets"], np.float32).reshape(2, 24), targets(rows))

