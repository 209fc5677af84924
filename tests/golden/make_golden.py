"""Writes tests/golden/c0_oracle.npz: the oracle's results on the 64x64
convection-diffusion workload (16 shifts, tolerance 1e-10) for first seeds
0, 8 and 15 -- returned k, per-shift stop iterations, the seed-switch log,
every 41st entry of each x_j and each x_j's sum -- so later changes to the
oracle or the GPU path are held against fixed numbers.

Run from the repository root: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "hybrid-bicg_amd"), os.path.join(ROOT, "oracle")]

import pyoracle as O  # noqa: E402
from hbicg import problems as P  # noqa: E402

SEEDS = (0, 8, 15)
STRIDE = 41


def main():
    p = P.config0()
    out = {}
    for s in SEEDS:
        r = O.solve(p.A, p.b, p.sigma, s, tol=p.tol)
        out[f"k_{s}"] = np.array(r.k)
        out[f"stop_{s}"] = r.stop_iter.astype(np.int32)
        out[f"switch_{s}"] = np.array(r.switches, dtype=np.int32).reshape(-1, 2)
        out[f"xs_{s}"] = r.x[:, ::STRIDE].copy()
        out[f"xsum_{s}"] = r.x.sum(axis=1)
    np.savez(os.path.join(ROOT, "tests", "golden", "c0_oracle.npz"), **out)


if __name__ == "__main__":
    main()
