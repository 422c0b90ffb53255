"""Writes dc_qlookup_8bit.json: AV1's normative 8-bit DC quantizer lookup
(the first bit-depth block of the {dc, ac} pairs in goav1's
internal/av1/quantize/tables.go), used by the search harness to price
lambda at any qindex.  Run once in the build container:

    python tests/golden/make_dc_table.py /root/reference
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
txt = open(os.path.join(ref, "internal", "av1", "quantize", "tables.go")).read()
body = txt[txt.index("var dequantLookup"):]
pairs = re.findall(r"\{\s*(\d+)\s*,\s*(\d+)\s*\}", body)
dc = [int(a) for a, _ in pairs[:256]]
ac = [int(b) for _, b in pairs[:256]]
assert len(dc) == 256 and dc[0] == 4 and dc[160] == 223 and ac[120] == 152
json.dump({"source": "internal/av1/quantize/tables.go (dequantLookup, 8-bit)", "dc": dc, "ac": ac},
          open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dc_qlookup_8bit.json"), "w"))
