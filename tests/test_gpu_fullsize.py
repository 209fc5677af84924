"""Parity at the workloads' full sizes.

128^3 7-point convection-diffusion, 64 shifts (n = 2,097,152), against the
oracle's complete solves committed in tests/golden/c1_oracle.npz (written by
tests/golden/make_c1_fixture.py), for first seed 0 (no switch) and first
seed 63 (the largest shift converges first; the seed switches at iteration
31 to shift 0, src/solver_dynamic.c:309-359):
  * exact order (the reference's summation order) equals the oracle bit for
    bit: k, every stop iteration, the switch log, the sampled rows and a
    bit-pattern checksum of every x_j and of the returned r;
  * the default order -- a device plan, and shifted_lopbicg_static and
    shifted_lopbicg_dynamic on host buffers -- stays within the parity bar:
    k, stop and switch iterations against the reference's own results (see
    below), the same switch targets, x_j within 1e-9 relative of the
    one-rank run on the sampled rows and in norm, and every true residual
    within 10x the stopping rule; static and dynamic agree bit for bit with
    each other.
"The reference's own results" are its runs on 1, 2, 4, ..., 64 MPI ranks
(stored in the fixture): on this operator BiCG amplifies a rounding
difference about 1e5-fold every ten iterations, so the split row sums and
per-rank dot sums of another decomposition already move the reference's k
by several iterations and single shifts' stop iterations by 10 to 20,
while x stays within a few 1e-10.  The tree-summed dots of the default
order are one more such rounding: k must lie within 2 of the range the
reference's decompositions span, and no shift's stop iteration may move
further from the one-rank run than the reference's own decompositions
move one, plus 2.
The GPU true-residual checker is held against the oracle's on the same x.

The irregular power-law matrix (n = 1e6, 31M nonzeros, 128 shifts): the
first iterations run by the oracle on the host and by the GPU with the same
cap agree bit for bit in exact order and within 1e-9 in the default order;
the complete solve meets the stopping rule for every shift with sigma >=
0.05, and every shift it stops has a true residual within 10x the rule.
The smallest shifts, the seed among them, do not reach 1e-10 within the
2000-iteration cap, so the solve returns cap + 1 like the reference.  The
shifts just above them stop near the cap, where summation order decides:
measured on MI355X, sigma = 0.0284 stops at iteration 1903 in the default
order with one SpMV pass, and not at all with column panels or in the
reference's order; stop iterations there move by up to 318 between orders.
"""
import os

import numpy as np
import pytest
import torch

import pyoracle as O
from hbicg import Plan, dropin_solve, make_params
from hbicg import problems as P

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c1_oracle.npz")


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD))


@pytest.fixture(scope="module")
def c1():
    return P.config1()


@pytest.fixture(scope="module")
def c2():
    print("generating the 1e6-row power-law matrix", flush=True)
    return P.config2()


def _bit_sums(a):
    """Per row, the sum modulo 2^64 of the entries' bit patterns (as the
    fixture script computes it): independent of summation order, so equal
    on any host exactly when the rows are -- up to a collision of two
    different vectors, which random data does not produce."""
    return np.ascontiguousarray(a).view(np.uint64).sum(axis=1, dtype=np.uint64)


def _rel_rows(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


RANKS = (2, 4, 8, 16, 32, 64)


def _within_bar(k, stop, sw, x, gold, s, what):
    """The parity bar of the default order against the reference's runs on
    1, 2, 4, ..., 64 ranks: k within 2 of the range they span; the largest
    change of a shift's stop iteration from the one-rank run no larger than
    the largest the reference's own decompositions show, plus 2; the same
    switch targets, at iterations within 2 of the range; x within 1e-9 of
    the one-rank run."""
    S = int(gold["stride"])
    stop1 = gold[f"stop_{s}"].astype(np.int64)
    ks = [int(gold[f"k_{s}"])] + [int(gold[f"k_{s}_r{P_}"]) for P_ in RANKS]
    sws = [gold[f"switch_{s}"]] + [gold[f"switch_{s}_r{P_}"] for P_ in RANKS]
    ref_dev = max(int(np.abs(gold[f"stop_{s}_r{P_}"] - stop1).max()) for P_ in RANKS)
    dev = int(np.abs(stop.astype(np.int64) - stop1).max())
    assert all(np.array_equal(w[:, 1], sws[0][:, 1]) for w in sws)   # the reference agrees on targets
    assert min(ks) - 2 <= k <= max(ks) + 2, (what, k, ks)
    assert dev <= ref_dev + 2, (what, dev, ref_dev, stop - stop1)
    assert [t for _, t in sw] == [int(t) for t in sws[0][:, 1]], (what, sw)
    for i, (kg, _) in enumerate(sw):
        kr = [int(w[i, 0]) for w in sws]
        assert min(kr) - 2 <= kg <= max(kr) + 2, (what, sw, kr)
    rel = _rel_rows(x[:, ::S], gold[f"xs_{s}"])
    assert rel.max() <= 1e-9, (what, rel.max())
    nrel = np.abs(np.linalg.norm(x, axis=1) - gold[f"xnorm_{s}"]) / gold[f"xnorm_{s}"]
    assert nrel.max() <= 1e-9, (what, nrel.max())
    print(f"  {what}: k={k} (reference on 1..64 ranks: {ks}), largest stop change {dev} "
          f"(the reference's own: {ref_dev}), max rel dx on samples {rel.max():.2e}, "
          f"in norm {nrel.max():.2e}", flush=True)


@pytest.mark.parametrize("seed", [0, 63])
def test_c1_exact_order_equals_oracle(c1, gold, seed):
    print(f"case c1 exact order, seed {seed}", flush=True)
    pl = Plan(c1.A, c1.M, make_params(tol=c1.tol, max_iter=c1.max_iter, exact_order=True, graph=1))
    b = torch.as_tensor(c1.b).cuda()
    r = torch.empty_like(b)
    k, x, _ = pl.solve(b, c1.sigma, seed, r_out=r)
    _, stop, sw = pl.report()
    tr = pl.true_residuals(b, c1.sigma, x)
    pl.close()
    S = int(gold["stride"])
    xh = x.cpu().numpy()
    rh = r.cpu().numpy()
    del x
    assert k == int(gold[f"k_{seed}"])
    assert np.array_equal(stop, gold[f"stop_{seed}"])
    assert sw == [tuple(int(v) for v in row) for row in gold[f"switch_{seed}"]]
    if seed == 63:
        assert len(sw) >= 1
    assert np.array_equal(xh[:, ::S], gold[f"xs_{seed}"])
    assert np.array_equal(_bit_sums(xh), gold[f"xbits_{seed}"])
    assert np.array_equal(rh[::S], gold[f"rs_{seed}"])
    assert np.array_equal(_bit_sums(rh[None, :]), gold[f"rbits_{seed}"])
    # the same x: the GPU checker against the oracle's.  Each residual entry
    # is the same double in both; they differ in how the 2M squared entries
    # are summed: the oracle's serial sum of n positive terms carries a
    # relative error of up to about n u = 2.3e-10, the GPU's tree far less
    np.testing.assert_allclose(tr, gold[f"tres_{seed}"], rtol=2.5e-10, atol=0)
    assert tr.max() <= 10 * c1.tol


@pytest.mark.parametrize("seed", [0, 63])
def test_c1_default_order_plan_and_dropins(c1, gold, seed):
    print(f"case c1 default order, seed {seed}", flush=True)
    b = torch.as_tensor(c1.b).cuda()
    pl = Plan(c1.A, c1.M, make_params(tol=c1.tol, max_iter=c1.max_iter, graph=1))
    k, x, rep = pl.solve(b, c1.sigma, seed)
    _, stop, sw = pl.report()
    tr = pl.true_residuals(b, c1.sigma, x)
    assert rep["stop_count"] == c1.M and tr.max() <= 10 * c1.tol, tr.max()
    _within_bar(k, stop, sw, x.cpu().numpy(), gold, seed, "device plan")
    del x
    pl.close()
    xs = {}
    for variant in ("static", "dynamic"):
        got = dropin_solve(c1.A, c1.b, c1.sigma, seed, variant, make_params(tol=c1.tol))
        assert got.report["stop_count"] == c1.M
        _within_bar(got.k, got.stop_iter, got.switches, got.x, gold, seed, f"shifted_lopbicg_{variant}")
        xs[variant] = got
    assert xs["static"].k == xs["dynamic"].k
    assert np.array_equal(xs["static"].x, xs["dynamic"].x)
    assert np.array_equal(xs["static"].r, xs["dynamic"].r)


def test_c1_fused_levels_bitwise(c1):
    """At full size, with the seed switching (first seed 63): the four-launch
    iteration (scalar stage inside the shift kernel, whose tiles walk the
    seed p update and the flush grid-stride) against the eight-launch one,
    bit for bit, as a device plan with hipGraph chunks."""
    print("case c1 fused levels 1 and 0, seed 63", flush=True)
    b = torch.as_tensor(c1.b).cuda()
    out = []
    for level in (1, 0):
        pl = Plan(c1.A, c1.M, make_params(tol=c1.tol, max_iter=c1.max_iter, graph=1, fused=level))
        r = torch.empty_like(b)
        k, x, rep = pl.solve(b, c1.sigma, 63, r_out=r)
        _, stop, sw = pl.report()
        assert not (rep["status"] & 16)
        out.append((k, stop, sw, x, r))
        pl.close()
    (k1, s1, w1, x1, r1), (k0, s0, w0, x0, r0) = out
    assert k1 == k0 and w1 == w0 and len(w1) >= 1
    assert np.array_equal(s1, s0)
    assert torch.equal(x1, x0) and torch.equal(r1, r0)


def test_true_residual_checker_against_oracle(c1):
    """hbicg_plan_true_residuals against the oracle's checker on x far from
    converged (every residual O(1)): the 64x64 workload's 16 shifts and 4
    shifts of the 128^3 workload."""
    rng = np.random.default_rng(41)
    c0 = P.config0()
    cases = [(c0.A, c0.b, c0.sigma), (c1.A, c1.b, c1.sigma[[0, 21, 42, 63]])]
    for A, bh, sigma in cases:
        print(f"case true residual checker n={A.n} M={len(sigma)}", flush=True)
        M = len(sigma)
        xh = rng.uniform(-1, 1, (M, A.n))
        pl = Plan(A, M, make_params(tol=1e-10))
        b = torch.as_tensor(bh).cuda()
        got = pl.true_residuals(b, sigma, torch.as_tensor(xh).cuda())
        pl.close()
        want = O.true_residuals(A, xh, bh, sigma)
        assert want.min() > 0.1
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=0)


def _capped(prob, cap, exact, window=0):
    pl = Plan(prob.A, prob.M, make_params(tol=prob.tol, max_iter=cap, exact_order=exact,
                                          window=window, graph=1))
    b = torch.as_tensor(prob.b).cuda()
    r = torch.empty_like(b)
    k, x, _ = pl.solve(b, prob.sigma, prob.seed, r_out=r)
    out = k, x.cpu().numpy(), r.cpu().numpy()
    pl.close()
    return out


def test_c2_capped_iterations_match_oracle(c2):
    cap = 8
    print(f"case c2 capped at {cap} iterations: oracle", flush=True)
    ref = O.solve(c2.A, c2.b, c2.sigma, c2.seed, tol=c2.tol, max_iter=cap)
    print("case c2: GPU exact order", flush=True)
    k, x, r = _capped(c2, cap, exact=True)
    assert k == ref.k
    assert np.array_equal(r, ref.r)
    assert np.array_equal(x, ref.x)
    print("case c2: GPU default order", flush=True)
    k, x, r = _capped(c2, cap, exact=False)
    assert k == ref.k
    rel = np.linalg.norm(x - ref.x, axis=1) / np.maximum(np.linalg.norm(ref.x, axis=1), 1e-300)
    assert rel.max() <= 1e-9, rel.max()


def test_c2_full_solve_true_residuals(c2):
    print("case c2: full solve, graph and direct launches", flush=True)
    b = torch.as_tensor(c2.b).cuda()
    xs = []
    for graph in (1, 0):
        pl = Plan(c2.A, c2.M, make_params(tol=c2.tol, graph=graph))
        k, x, rep = pl.solve(b, c2.sigma, c2.seed)
        _, stop, sw = pl.report()
        tr = pl.true_residuals(b, c2.sigma, x)
        stopped = stop > 0
        assert rep["stop_count"] == int(stopped.sum())
        assert np.all(stopped[c2.sigma >= 0.05])
        assert k == c2.max_iter + 1 if not np.all(stopped) else k < c2.max_iter + 1
        assert tr[stopped].max() <= 10 * c2.tol, tr[stopped].max()
        xs.append(x)
        pl.close()
    assert torch.equal(xs[0], xs[1])
