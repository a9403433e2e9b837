#!/usr/bin/env python3
"""Write tests/golden/small_oracle.npz: the CPU oracle (reference mode) on the
small scene (512 x 2048 pulses -> 256 x 256, 3 targets, flat DEM, noise 0.01,
seed 0), every 4th pixel in both directions, with heights and the update count."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle  # noqa: E402
from isce3_amd import sim  # noqa: E402

sc = sim.make_scene("small")
raw = sim.simulate(sc)
out, h, diag, rc = oracle.backproject(sc.out_geometry(), raw, sc.in_geometry(), sc.dem, sc.fc,
                                      sc.ds, sc.kernel, sc.dry_tropo_model, want_height=True,
                                      want_diag=True)
assert rc == 0
L, W = out.shape
rows, cols = np.meshgrid(np.arange(0, L, 4), np.arange(0, W, 4), indexing="ij")
idx = (rows * W + cols).ravel()
updates = sum(d.kstop - d.kstart for d in diag)
np.savez_compressed(os.path.join(HERE, "small_oracle.npz"), index=idx, out=out.ravel()[idx],
                    height=h.ravel()[idx], updates=np.int64(updates), seed=np.int64(sc.seed))
print("wrote small_oracle.npz", idx.size, "pixels, updates", updates)
