#!/usr/bin/env python3
"""Write tests/golden/ fixtures from the reference's own test data and from the
reference building blocks compiled in oracle/_ref (run in the container that
has the reference tree; the fixtures travel, the reference does not).

  ellipsoid_kat.json   the 21 LLH<->XYZ known answers of
                       tests/cxx/isce3/core/ellipsoid/ellipsoid.cpp:40-105
                       (numbers parsed from that file)
  ref_blocks.npz       outputs of the reference's KnabKernel/TabulatedKernel
                       tables, TabulatedKernel<float> evaluation, interp1d on a
                       seeded complex line, Linspace::search and Brent roots on
                       the Brent-1973 test functions
                       (tests/cxx/isce3/math/root_find1d.cpp:160-230)
"""
import ctypes as C
import json
import math
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = os.environ.get("ISCE3_REF", "/root/reference")
sys.path.insert(0, os.path.join(ROOT, "oracle"))


def ellipsoid_kat():
    src = open(os.path.join(REF, "tests/cxx/isce3/core/ellipsoid/ellipsoid.cpp")).read()
    a = 6378137.0
    e2 = 0.0066943799901
    b = a * math.sqrt(1.0 - e2)
    env = {"a": a, "b": b, "M_PI": math.pi}
    cases = []
    for m in re.finditer(r"ellipsoidTest\((\w+),\s*\{([^}]*)\},\s*\{([^}]*)\}\)", src):
        name = m.group(1)
        llh = [eval(v.strip(), {}, env) for v in m.group(2).split(",")]
        xyz = [eval(v.strip(), {}, env) for v in m.group(3).split(",")]
        cases.append({"name": name, "llh": llh, "xyz": xyz})
    assert len(cases) == 21, len(cases)
    return {"a": a, "e2": e2, "tol_xyz": 1e-6, "tol_lonlat": 1e-9, "tol_h": 1e-6,
            "source": "tests/cxx/isce3/core/ellipsoid/ellipsoid.cpp:40-105",
            "cases": cases}


def brent_fns():
    def f1(x):
        return math.cos(x)

    def f2(x):
        return -100.0 * x * math.exp(-2.0 * x)

    def f3(x):
        if abs(x) < 3.8e-4:
            return 0.0
        try:
            return x / math.exp(x ** -2)
        except OverflowError:  # C's exp overflows to inf, giving a signed zero
            return x / math.inf

    def f4(x):
        return x ** 9.0

    def f5(x):
        if abs(x) < 1e-3:
            return float("nan")
        return x ** 3.0

    return [(1, f1, 4.0, 5.0), (2, f2, -9.0, 31.0), (3, f3, -1.0, 4.0), (4, f4, -1.0, 1.1),
            (5, f5, -5.0, 5.0)]


def ref_blocks():
    import oracle
    ref = oracle.ref_lib()
    if ref is None:
        raise SystemExit("oracle/_ref not built: run `make -C oracle ref`")
    out = {}
    for w in (8.0, 9.0):
        for bw in (0.8, 0.83):
            n = 2048
            t32 = np.empty(n, np.float32)
            t64 = np.empty(n, np.float64)
            ref.ref_tab_table_f32(w, bw, n, t32.ctypes.data)
            ref.ref_tab_table_f64(w, bw, n, t64.ctypes.data)
            out[f"table_f32_w{w:g}_bw{bw:g}"] = t32
            out[f"table_f64_w{w:g}_bw{bw:g}"] = t64
    rng = np.random.Generator(np.random.PCG64(2024))
    x = np.concatenate([rng.uniform(-4.6, 4.6, 4000), np.arange(-5, 5.01, 0.25)])
    out["tab_eval_x"] = x
    for w in (8.0, 9.0):
        y = np.empty(x.size, np.float32)
        ref.ref_tab_eval_f32(w, 0.8, 2048, x.ctypes.data, x.size, y.ctypes.data)
        out[f"tab_eval_w{w:g}"] = y
    line = (rng.standard_normal(257) + 1j * rng.standard_normal(257)).astype(np.complex64)
    t = np.concatenate([rng.uniform(-6.0, 262.0, 3000), np.arange(-5.0, 262.0, 0.5)])
    out["interp_line"] = line
    out["interp_t"] = t
    for w in (8.0, 9.0):
        y = np.empty(t.size, np.complex64)
        ref.ref_interp1d_c64(w, 0.8, 2048, line.ctypes.data, line.size, t.ctypes.data, t.size,
                             y.ctypes.data)
        out[f"interp_w{w:g}"] = y
    vals = np.concatenate([rng.uniform(-20, 140, 500), [-10.0, 0.0, 10.0, 100.0, 100.5]])
    out["search_vals"] = vals
    out["search_idx"] = np.array([ref.ref_linspace_search(0.0, 10.0, 11, v) for v in vals],
                                 dtype=np.int32)
    roots, status = [], []
    for fid, f, a, b in brent_fns():
        cb = ref.FN(lambda xx, _ctx, f=f: f(xx))
        r = C.c_double(0.0)
        st = ref.ref_find_zero_brent(a, b, cb, None, 1e-12, C.byref(r))
        roots.append(r.value)
        status.append(st)
    out["brent_roots"] = np.array(roots)
    out["brent_status"] = np.array(status, dtype=np.int32)
    out["knab_x"] = np.linspace(-4.5, 4.5, 181)
    out["knab_w8"] = np.array([ref.ref_knab(8.0, 0.8, v) for v in out["knab_x"]])
    return out


if __name__ == "__main__":
    with open(os.path.join(HERE, "ellipsoid_kat.json"), "w") as f:
        json.dump(ellipsoid_kat(), f, indent=1)
    np.savez_compressed(os.path.join(HERE, "ref_blocks.npz"), **ref_blocks())
    print("wrote ellipsoid_kat.json and ref_blocks.npz")
