"""Extract the D3Q27 tables (weights t, bounce-back bb, velocities e) from the
reference's tgv/src/lbm.cpp:27-190 into tests/golden/lattice_d3q27.json.

Run in a container that has the reference checkout at /root/reference; the
JSON it writes is the committed fixture (index -> value data only).
"""
import json
import os
import re
import sys

SRC = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/tgv/src/lbm.cpp"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lattice_d3q27.json")

text = open(SRC).read()
t = {int(i): (int(a), int(b)) for i, a, b in re.findall(r"\bt\((\d+)\)\s*=\s*(\d+)\.0\s*/\s*(\d+)\.0", text)}
bb = {int(i): int(v) for i, v in re.findall(r"\bbb\((\d+)\)\s*=\s*(\d+)\s*;", text)}
e = {}
for i, d, v in re.findall(r"\be\((\d+),\s*(\d)\)\s*=\s*(-?\d+)\s*;", text):
    e.setdefault(int(i), [0, 0, 0])[int(d)] = int(v)
q = 27
assert sorted(t) == list(range(q)) and sorted(bb) == list(range(q)) and sorted(e) == list(range(q))
fixture = {
    "source": "tgv/src/lbm.cpp:27-190",
    "weights_num_den": [list(t[i]) for i in range(q)],
    "bb": [bb[i] for i in range(q)],
    "e": [e[i] for i in range(q)],
}
with open(OUT, "w") as f:
    json.dump(fixture, f, indent=1)
print(OUT)
