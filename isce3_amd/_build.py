"""Build the native libraries in-tree.

    python -m isce3_amd._build

Compiles the HIP library for gfx950 (libisce3_amd.so) and the host scene
generator (libisce3_amd_sim.so) into isce3_amd/lib/.  Packed FP32
instructions are disabled for all device code.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ISCE3_AMD_ARCH", "gfx950")

HIP_FLAGS = ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-shared",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in res.stderr.splitlines() if "packed-fp32-ops" not in l)
    if err.strip():
        print(err, file=sys.stderr)
    if res.returncode != 0:
        raise RuntimeError(f"build step failed ({res.returncode}): {' '.join(cmd)}")


def build_hip():
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = [os.path.join(CSRC, f)
            for f in ("bp_kernels.hip", "bp_api.hip", "bp_serve.hip", "rc_kernels.hip",
                      "rc_api.hip", "rslc.hip", "presum.hip")]
    out = os.path.join(LIBDIR, "libisce3_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, out)
    return out


def build_sim():
    os.makedirs(LIBDIR, exist_ok=True)
    out = os.path.join(LIBDIR, "libisce3_amd_sim.so")
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared", "-Wall", "-o", tmp,
          os.path.join(HERE, "sim", "sim.cpp")])
    os.replace(tmp, out)
    return out


def build_all():
    return [build_hip(), build_sim()]


if __name__ == "__main__":
    for p in build_all():
        print("built", p)
