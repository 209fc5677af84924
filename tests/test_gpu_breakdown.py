"""Breakdowns and degenerate input through the C-ABI, against the oracle.

The reference traps none of these (src/solver_dynamic.c:167 divides by
(r^, s), :189 by (q, y), :245-253 by 1 - omega~ (sigma_s - sigma_j) and by
pi_j[k]): the NaN or infinity goes on through the iteration, nothing meets
the stopping rule, and the loop runs to its cap.  The GPU solver does the
same arithmetic, ends at the same iteration, and reports what happened in
hbicg_report.status.  In exact order the whole run equals the oracle's:
k, stop iterations and x (NaN where the oracle has NaN).  Every case is a
numerical NaN, not a fault: each kernel still walks the same indices.
"""
import numpy as np
import pytest

import pyoracle as O
from hbicg import dropin_solve, make_params
from hbicg import problems as P

pytestmark = pytest.mark.gpu

RTS, QTY, NO_TARGET, SHIFT = 1, 2, 4, 8


def _same(got, ref):
    assert got.k == ref.k, (got.k, ref.k)
    np.testing.assert_array_equal(got.stop_iter, ref.stop_iter)
    assert got.switches == ref.switches
    assert np.array_equal(got.x, ref.x, equal_nan=True)


@pytest.fixture(scope="module")
def c0():
    return P.config0()


def test_zero_rhs(c0):
    """b = 0: r^ = 0, so (r^, s) = 0 and alpha~ = 0/0 at the first step."""
    b = np.zeros(c0.n)
    cap = 30
    print("case breakdown b = 0", flush=True)
    ref = O.solve(c0.A, b, c0.sigma, 0, tol=c0.tol, max_iter=cap)
    assert ref.k == cap + 1
    got = dropin_solve(c0.A, b, c0.sigma, 0, "static",
                       make_params(tol=c0.tol, max_iter=cap, exact_order=True))
    _same(got, ref)
    assert got.report["status"] & RTS
    for variant in ("normal", "dynamic"):
        got = dropin_solve(c0.A, b, c0.sigma, 0, variant, make_params(tol=c0.tol, max_iter=cap))
        assert got.k == cap + 1 and got.report["status"] & RTS
        assert got.report["stop_count"] == 0


def test_q_zero(c0):
    """A = 2I with sigma_s = 0: s = 2b, alpha~ = (b,b)/(b,2b) = 1/2 exactly
    (serial sums in exact order), so q = r - alpha~ s = 0 and (q, y) = 0."""
    n = 1000
    A = P.CSR(np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.full(n, 2.0), n)
    b = P.uniform(9, 1, n, -1.0, 1.0)
    sigma = np.array([0.0, 0.5, 1.0])
    cap = 20
    print("case breakdown q = 0", flush=True)
    ref = O.solve(A, b, sigma, 0, tol=1e-10, max_iter=cap)
    assert ref.k == cap + 1 and ref.alpha[1] == 0.5
    got = dropin_solve(A, b, sigma, 0, "dynamic",
                       make_params(tol=1e-10, max_iter=cap, exact_order=True))
    _same(got, ref)
    assert got.report["status"] & QTY


def _breakdown_shift(A, b, sig_s):
    """A shift sigma_j with fl(1 - omega~_1 (sigma_s - sigma_j)) = 0: its
    omega_j = omega~ / 0 and zeta_j = 0 from the first iteration on."""
    w = O.solve(A, b, np.array([sig_s]), 0, tol=1e-10, max_iter=1).omega[1]
    cand = sig_s - 1.0 / w
    for step in range(-200, 201):
        sj = cand
        if step:
            sj = np.nextafter(cand, np.inf if step > 0 else -np.inf)
            for _ in range(abs(step) - 1):
                sj = np.nextafter(sj, np.inf if step > 0 else -np.inf)
        d = np.float64(sig_s) - np.float64(sj)
        if np.float64(1.0) - np.float64(w) * d == 0.0:
            return float(sj)
    raise AssertionError("no shift found")


def test_shift_denominator_zero(c0):
    """A shift with 1 - omega~ (sigma_s - sigma_j) = 0 at the first step
    (src/solver_dynamic.c:247): its coefficients divide by zero, its estimate
    is infinite, so it becomes the switch target once the seed stops, and the
    switch carries the infinity into the seed system; the run goes to its
    cap like the oracle's."""
    sj = _breakdown_shift(c0.A, c0.b, c0.sigma[0])
    sigma = np.array([c0.sigma[0], sj, c0.sigma[3]])
    cap = 150
    print(f"case breakdown 1 - omega d = 0, sigma_j = {sj!r}", flush=True)
    ref = O.solve(c0.A, c0.b, sigma, 0, tol=c0.tol, max_iter=cap)
    got = dropin_solve(c0.A, c0.b, sigma, 0, "static",
                       make_params(tol=c0.tol, max_iter=cap, exact_order=True))
    _same(got, ref)
    assert got.report["status"] & SHIFT


def test_nan_shift_leaves_no_switch_target(c0):
    """A NaN shift never stops and its estimate is never above the seed's;
    when the seed stops there is nothing to switch to.  The reference would
    switch to an uninitialised index (src/solver_dynamic.c:25,318); here the
    solve ends after that iteration (status NO_SWITCH_TARGET) and the shifts
    that converged keep their solutions."""
    sigma = np.array([c0.sigma[0], np.nan, c0.sigma[5]])
    print("case breakdown NaN shift", flush=True)
    ref = O.solve(c0.A, c0.b, sigma, 0, tol=c0.tol)
    assert ref.status & O.NO_SWITCH_TARGET and ref.k < 2001
    assert ref.stop_iter[1] == 0 and ref.stop_iter[0] > 0 and ref.stop_iter[2] > 0
    got = dropin_solve(c0.A, c0.b, sigma, 0, "dynamic",
                       make_params(tol=c0.tol, exact_order=True))
    _same(got, ref)
    assert got.report["status"] & NO_TARGET and got.report["status"] & SHIFT
    got = dropin_solve(c0.A, c0.b, sigma, 0, "dynamic", make_params(tol=c0.tol))
    assert abs(got.k - ref.k) <= 2 and got.report["status"] & NO_TARGET
    ok = [0, 2]
    rel = np.linalg.norm(got.x[ok] - ref.x[ok], axis=1) / np.linalg.norm(ref.x[ok], axis=1)
    assert rel.max() <= 1e-9, rel
    assert np.all(np.isnan(got.x[1]))
