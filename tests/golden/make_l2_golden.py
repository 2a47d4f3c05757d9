"""Writes tests/golden/l2_small.ar2 and l2_small.json: the synthetic volume of l2_volume.py (this build's
layout and header) passed through the reference's own Archive II writer, l2writer.encode_volume
(l2writer.py:161-182), exactly as server.py:256-268 calls it for an uploaded volume.

Run from the repository root (needs the reference tree; the writer is imported, not copied):
    python tests/golden/make_l2_golden.py /root/reference
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

from l2_volume import digest, synthetic_volume  # noqa: E402
from outbreaklab_amd.radar import level2_sweeps  # noqa: E402

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
sys.path.insert(0, ref)
import l2writer  # noqa: E402

header, codes = synthetic_volume()
args = level2_sweeps(header, codes)
data = l2writer.encode_volume(*args)
name = l2writer.archive_name(args[0], args[4])
with open(os.path.join(HERE, "l2_small.ar2"), "wb") as f:
    f.write(data)
with open(os.path.join(HERE, "l2_small.json"), "w") as f:
    json.dump({"header": header, "codes_sha256": digest(codes), "archive_name": name, "bytes": len(data)}, f, indent=1)
print(name, len(data), "bytes")
