"""Writes tests/golden/c1_oracle.npz: the oracle's complete solves of the
128^3 7-point convection-diffusion workload (n = 2,097,152, 64 shifts
log-spaced in [1e-3, 1], b uniform in [-1, 1] from seed 1, tolerance
1e-10), for first seed 0 (the smallest shift: no seed switch) and first
seed 63 (the largest shift: it converges first and the seed switches).

Stored per first seed s:
  k_s          returned k (iterations + 1)
  stop_s       [64] iteration at which each shift stopped
  switch_s     [n_switch, 2] (iteration, new seed)
  xs_s         [64, 128] x_j at rows 0, STRIDE, 2 STRIDE, ...
  xbits_s      [64] the sum modulo 2^64 of every x_j's entries read as 64-bit
               integers: it does not depend on the order of summation, so it
               is the same on any machine, and two vectors that differ in
               one bit differ in it
  xnorm_s      [64] 2-norm of every x_j (for the relative comparisons)
  rs_s, rbits_s, rnorm_s  the returned r (the last seed's residual) at the
               same rows, its bit-pattern sum and its norm
  tres_s       [64] the oracle's true residual ||b - (A + sigma_j I) x_j|| / ||b||
  k_s_rP, stop_s_rP, switch_s_rP, xs_s_rP
               the same solve as the reference computes it on P = 2, 4, ..., 64 MPI
               ranks (the oracle's orc_shifted_lopbicg_ranks: split row sums,
               per-rank dots): the reference's own spread across its
               decompositions, which the default-order tests are held to

The oracle takes one to two minutes per solve on 8 threads, too long for the
CPU test suite, so the GPU tests compare against these numbers instead of
running it.  Run from the repository root:
    python tests/golden/make_c1_fixture.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "hybrid-bicg_amd"), os.path.join(ROOT, "oracle")]

import pyoracle as O  # noqa: E402
from hbicg import problems as P  # noqa: E402

SEEDS = (0, 63)
RANKS = (2, 4, 8, 16, 32, 64)
STRIDE = 16411   # prime: 128 rows per shift, spread over the whole grid


def bit_sums(a):
    """Per row, the sum modulo 2^64 of the entries' IEEE bit patterns."""
    return np.ascontiguousarray(a).view(np.uint64).sum(axis=1, dtype=np.uint64)


def main():
    p = P.config1()
    out = {"stride": np.array(STRIDE)}
    for s in SEEDS:
        t = time.time()
        r = O.solve(p.A, p.b, p.sigma, s, tol=p.tol, max_iter=p.max_iter)
        print(f"seed {s}: k={r.k}, switches {r.switches}, {time.time() - t:.0f} s", flush=True)
        out[f"k_{s}"] = np.array(r.k)
        out[f"stop_{s}"] = r.stop_iter.astype(np.int32)
        out[f"switch_{s}"] = np.array(r.switches, dtype=np.int32).reshape(-1, 2)
        out[f"xs_{s}"] = r.x[:, ::STRIDE].copy()
        out[f"xbits_{s}"] = bit_sums(r.x)
        out[f"xnorm_{s}"] = np.linalg.norm(r.x, axis=1)
        out[f"rs_{s}"] = r.r[::STRIDE].copy()
        out[f"rbits_{s}"] = bit_sums(r.r[None, :])
        out[f"rnorm_{s}"] = np.array(np.linalg.norm(r.r))
        out[f"tres_{s}"] = O.true_residuals(p.A, r.x, p.b, p.sigma)
        del r
        for P_ in RANKS:
            t = time.time()
            r = O.solve(p.A, p.b, p.sigma, s, tol=p.tol, max_iter=p.max_iter, ranks=P_)
            print(f"seed {s}, {P_} ranks: k={r.k}, switches {r.switches}, {time.time() - t:.0f} s", flush=True)
            out[f"k_{s}_r{P_}"] = np.array(r.k)
            out[f"stop_{s}_r{P_}"] = r.stop_iter.astype(np.int32)
            out[f"switch_{s}_r{P_}"] = np.array(r.switches, dtype=np.int32).reshape(-1, 2)
            out[f"xs_{s}_r{P_}"] = r.x[:, ::STRIDE].copy()
            del r
    np.savez_compressed(os.path.join(ROOT, "tests", "golden", "c1_oracle.npz"), **out)


if __name__ == "__main__":
    main()
