"""Occupancy of the hot kernels, read from the compiler's resource remarks (no GPU needed).

Each of these kernels was measured at the occupancy below and loses a large share of its speed when one register
too many drops it a wave per SIMD (the hurricane moisture corrector at 257 registers ran 1.6 times slower). The
test compiles the device code as the build does and fails when a kernel's waves per SIMD fall below the floor or
it starts spilling to scratch.
"""
import os
import re
import shutil
import subprocess

import pytest

from outbreaklab_amd import build as B

CSRC = os.path.join(os.path.dirname(B.__file__), "csrc")

# kernel (demangled, without arguments) -> minimum waves per SIMD
FLOORS = {
    "k_adv.hip": {
        "obl::k_adv_t": 2,
    },
    "k_step.hip": {
        "void obl::k_phys<0, false>": 3,
        "void obl::k_phys<2, false>": 3,
        "void obl::k_phys<0, true>": 2,
        "obl::k_advq1_hr": 2,
        "obl::k_advq2_hr": 2,
        "void obl::k_advv2_hr<true>": 3,
        "void obl::k_div_fold<3>": 8,
        "obl::k_proj": 8,
    },
    "k_solve.hip": {
        "void obl::k_thomas<80, true>": 2,
        "void obl::k_dst_mx<0, 3, 3>": 1,
        "void obl::k_dst_mx<3, 3, 3>": 1,
    },
}


def resources(src):
    cmd = [B.HIPCC] + [f for f in B.HIP_FLAGS if f != "-shared"] + B.FILE_FLAGS.get(src, []) + [
        "--offload-device-only", "-c", os.path.join(CSRC, src), "-o", os.devnull, "-Rpass-analysis=kernel-resource-usage"]
    out = subprocess.run(cmd, capture_output=True, text=True, cwd=CSRC)
    assert out.returncode == 0, out.stderr[-2000:]
    res, name = {}, None
    for line in out.stderr.splitlines():
        m = re.search(r"remark:\s+(.*?)\s*\[-Rpass", line)
        if not m:
            continue
        t = m.group(1)
        if t.startswith("Function Name:"):
            mangled = t.split(":", 1)[1].strip()
            name = subprocess.run(["c++filt", mangled], capture_output=True, text=True).stdout.strip().split("(")[0]
            res[name] = {}
        elif ":" in t and name is not None:
            k, v = t.split(":", 1)
            res[name][k.strip()] = v.strip()
    return res


@pytest.mark.skipif(not os.path.exists(B.HIPCC) or shutil.which("c++filt") is None, reason="hipcc / c++filt absent")
@pytest.mark.parametrize("src", sorted(FLOORS))
def test_hot_kernels_keep_their_occupancy(src):
    res = resources(src)
    for kernel, floor in FLOORS[src].items():
        assert kernel in res, f"{kernel} not in {src}"
        r = res[kernel]
        occ = int(r["Occupancy [waves/SIMD]"])
        scratch = int(r["ScratchSize [bytes/lane]"])
        assert occ >= floor, f"{kernel}: {occ} waves per SIMD (VGPRs {r['VGPRs']}, AGPRs {r['AGPRs']}), floor {floor}"
        assert r.get("VGPRs Spill", "0") == "0", f"{kernel} spills {r['VGPRs Spill']} VGPRs (scratch {scratch} B)"
