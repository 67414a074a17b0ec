"""Extract the reference direction tables into velocity_sets.json.

Reads BoltzmannFluids/Source/LBMConstants/VelocitySet.cpp of the reference as
text (nothing is compiled or run) and records, per velocity set, the list of
directions (cx, cy, cz) and the weight as the literal fraction written in the
source.  The JSON is the committed fixture; this script documents how it was
made and is not used at test time.

usage: python tests/golden/extract_velocity_sets.py /root/reference
"""
import json
import os
import re
import sys

ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
src = open(os.path.join(ref, "BoltzmannFluids/Source/LBMConstants/VelocitySet.cpp")).read()
out = {}
for name in ["D2Q9", "D3Q15", "D3Q19", "D3Q27", "D2Q5", "D3Q7"]:
    # the `case NAME:` inside get_velosity_vectors is followed by `return{ ... };`
    for m in re.finditer(r"case\s+%s\s*:\s*return\s*\{(.*?)\};" % name, src, re.S):
        rows = re.findall(r"glm::vec4\(\s*(-?\d+)\s*,\s*(-?\d+)\s*,\s*(-?\d+)\s*,\s*([\d.]+)\s*/\s*(\d+)\s*\)", m.group(1))
        if rows:
            out[name] = [{"c": [int(a), int(b), int(c)], "w": [float(n), int(d)]} for a, b, c, n, d in rows]
json.dump({"source": "BoltzmannFluids/Source/LBMConstants/VelocitySet.cpp:53-141,180-203", "sets": out},
          open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "velocity_sets.json"), "w"), indent=1)
print({k: len(v) for k, v in out.items()})
