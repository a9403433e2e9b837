"""Collect the known-answer vectors of the reference's projection tests
(tests/cxx/isce3/core/projections/{utm,polar,cea,geocent}.cpp) into
projection_kats.json: one record per test, {"epsg", "llh", "xyz"} with llh in
radians (lon, lat, h) and xyz the projected coordinates.  Reads the test
sources as text from the reference tree; writes data only.

    python tests/golden/make_projection_kats.py [/path/to/reference]
"""
import json
import math
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
SRC = os.path.join(REF, "tests", "cxx", "isce3", "core", "projections")
NUM = r"[-+0-9.eE*_A-Z]+"


def vec(text):
    vals = []
    for tok in text.split(","):
        tok = tok.strip().replace("M_PI", repr(math.pi))
        vals.append(float(eval(tok, {"__builtins__": {}})))  # arithmetic literals only
    assert len(vals) == 3, text
    return vals


def body(name):
    src = open(os.path.join(SRC, name)).read()
    src = re.sub(r"//[^\n]*", "", src)
    return re.sub(r"/\*.*?\*/", "", src, flags=re.S)


records = []
src = body("utm.cpp")
for m in re.finditer(r"utmOrigin(North|South)Test\((\d+)\);", src):
    zone = int(m.group(2))
    code = (32600 if m.group(1) == "North" else 32700) + zone
    lon = (-177.0 + (zone - 1) * 6.0) * math.pi / 180.0
    records.append({"epsg": code, "llh": [lon, 0.0, 0.0],
                    "xyz": [500000.0, 0.0 if m.group(1) == "North" else 10000000.0, 0.0]})
for m in re.finditer(r"utm(North|South)Test\((\d+),\s*\{([^}]*)\},\s*\{([^}]*)\}\s*\)", src):
    code = (32600 if m.group(1) == "North" else 32700) + int(m.group(2))
    records.append({"epsg": code, "llh": vec(m.group(3)), "xyz": vec(m.group(4))})
src = body("polar.cpp")
for m in re.finditer(r"polarTest\((North|South),\s*\w+,\s*\{([^}]*)\},\s*\{([^}]*)\}\s*\)", src):
    records.append({"epsg": 3413 if m.group(1) == "North" else 3031,
                    "llh": vec(m.group(2)), "xyz": vec(m.group(3))})
src = body("cea.cpp")
for m in re.finditer(r"ceaTest\(\w+,\s*\{([^}]*)\},\s*\{([^}]*)\}\s*\)", src):
    records.append({"epsg": 6933, "llh": vec(m.group(1)), "xyz": vec(m.group(2))})
# geocent.cpp: a, b of WGS84 (a = 6378137, b = a sqrt(1 - e^2))
A = 6378137.0
B = A * math.sqrt(1.0 - 0.006694379990141317)
src = body("geocent.cpp").replace("M_PI", repr(math.pi))
src = re.sub(r"\bb\b", repr(B), re.sub(r"\ba\b", repr(A), src))
for m in re.finditer(r"geocentTest\(\w+,\s*\{([^}]*)\},\s*\{([^}]*)\}\s*\)", src):
    records.append({"epsg": 4978, "llh": vec(m.group(1)), "xyz": vec(m.group(2))})
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "projection_kats.json")
with open(out, "w") as f:
    json.dump(records, f, indent=0)
counts = {}
for r in records:
    counts[r["epsg"] // 100] = counts.get(r["epsg"] // 100, 0) + 1
print(f"wrote {len(records)} vectors to {out}: {counts}")
