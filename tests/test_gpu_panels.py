"""Column-panel SpMV passes (params.panel_cols) through the C-ABI.

The seed SpMVs can run as one pass per panel of columns, each pass over the
entries of every row whose columns lie in that panel, with the row sums
carried between passes.  A row's columns ascend, so the passes take its
entries in CSR order: in exact order (one lane per row) the solve equals the
one-pass solve and the oracle bit for bit.  The default order sums each
panel's part of a row in a small tree and adds the parts in turn, so it
agrees within rounding.

Cases: a scattered matrix with power-law rows (the 1e6-row workload's
generator at 60k rows) in 15 panels; rows longer than one tile's LDS stage
(6000 entries: over 2048 in one panel too) in 2 and 5 panels; a matrix with
one row whose columns do not ascend (no panels: the report says 1).  The
1e6-row workload itself takes panels by default; test_gpu_fullsize holds it
to the oracle.
"""
import numpy as np
import pytest

import pyoracle as O
from hbicg import dropin_solve, make_params
from hbicg import problems as P

pytestmark = pytest.mark.gpu


def _long_rows(n=40000, seed=7, dom=1.1):
    """Rows of 3 to 40 random ascending columns, every 5000th row holding
    6000; diagonal dom times the row's absolute off-diagonal sum, + 0.1."""
    rng = np.random.default_rng(seed)
    ptr = [0]
    cols, vals = [], []
    for i in range(n):
        m = 6000 if i % 5000 == 17 else int(rng.integers(3, 41))
        c = np.unique(rng.integers(0, n, m))
        c = c[c != i]
        v = rng.uniform(-1, 1, c.size)
        c = np.append(c, i)
        v = np.append(v, dom * np.abs(v).sum() + 0.1)
        o = np.argsort(c)
        cols.append(c[o])
        vals.append(v[o])
        ptr.append(ptr[-1] + c.size)
    return P.CSR(np.array(ptr, dtype=np.uint32), np.concatenate(cols).astype(np.uint32),
                 np.concatenate(vals), n)


def _pair(A, b, sigma, seed, width, exact, cap=2000):
    one = dropin_solve(A, b, sigma, seed, "dynamic",
                       make_params(tol=1e-10, max_iter=cap, exact_order=exact, panel_cols=-1))
    pan = dropin_solve(A, b, sigma, seed, "dynamic",
                       make_params(tol=1e-10, max_iter=cap, exact_order=exact, panel_cols=width))
    assert one.report["panels"] == 1
    assert pan.report["panels"] == (A.n + width - 1) // width
    return one, pan


@pytest.mark.parametrize("case", ["zipf", "long2", "long5"])
def test_panels_exact_order_bitwise(case):
    if case == "zipf":
        A, width = P.zipf_matrix(60000, 30.0, 2000, seed=5), 4096
    else:
        A, width = _long_rows(), (20000 if case == "long2" else 8000)
    b = P.uniform(6, 1, A.n, -1.0, 1.0)
    sigma = P.uniform(6, 2, 12, 0.0, 0.5)
    seed = int(np.argmin(sigma))
    print(f"case panels exact {case} n={A.n} nnz={A.nnz} width={width}", flush=True)
    one, pan = _pair(A, b, sigma, seed, width, exact=True, cap=60)
    assert pan.k == one.k
    assert np.array_equal(pan.stop_iter, one.stop_iter)
    assert pan.switches == one.switches
    assert np.array_equal(pan.x, one.x)
    assert np.array_equal(pan.r, one.r)
    ref = O.solve(A, b, sigma, seed, tol=1e-10, max_iter=60)
    assert pan.k == ref.k and np.array_equal(pan.x, ref.x) and np.array_equal(pan.r, ref.r)


def _unit_diagonal(A, dom):
    """A with every diagonal entry set to dom times its row's absolute
    off-diagonal sum, then every row divided by that diagonal: unit
    diagonal, off-diagonal sums 1/dom."""
    rows = np.repeat(np.arange(A.n), np.diff(A.ptr.astype(np.int64)))
    on = A.col == rows
    off = np.bincount(rows[~on], weights=np.abs(A.val[~on]), minlength=A.n)
    d = dom * off + 0.1
    val = A.val / d[rows]
    val[on] = 1.0
    return P.CSR(A.ptr, A.col, val, A.n)


@pytest.mark.parametrize("case", ["zipf", "long5"])
def test_panels_default_order_within_rounding(case):
    """The default order against one pass.  With the diagonal at 1.1 times
    the off-diagonal sums, and diagonals spread over three decades, these
    matrices take hundreds of iterations, over which a rounding difference
    grows until the reference itself, on 2, 4 or 8 MPI ranks, moves single
    stop iterations by 19 to 100 (the oracle's rank restatement).  Scaled
    to a unit diagonal with off-diagonal sums of 0.98, they converge in 11
    and 28 iterations, where the rank counts agree on every stop iteration
    and a bar of 2 iterations and 1e-9 holds."""
    if case == "zipf":
        A, width = _unit_diagonal(P.zipf_matrix(60000, 30.0, 2000, seed=5), 1.02), 4096
    else:
        A, width = _unit_diagonal(_long_rows(), 1.02), 8000
    b = P.uniform(6, 1, A.n, -1.0, 1.0)
    sigma = P.uniform(6, 2, 12, 0.0, 0.5)
    seed = int(np.argmin(sigma))
    print(f"case panels default {case} n={A.n} width={width}", flush=True)
    one, pan = _pair(A, b, sigma, seed, width, exact=False)
    assert abs(pan.k - one.k) <= 2
    assert np.all(np.abs(pan.stop_iter - one.stop_iter) <= 2)
    rel = np.linalg.norm(pan.x - one.x, axis=1) / np.linalg.norm(one.x, axis=1)
    assert rel.max() <= 1e-9, rel.max()
    tr = O.true_residuals(A, pan.x, b, sigma)
    assert tr[pan.stop_iter > 0].max() <= 10 * 1e-10


def test_unsorted_row_takes_no_panels():
    A = P.zipf_matrix(20000, 10.0, 200, seed=8)
    col = A.col.copy()
    a, e = int(A.ptr[100]), int(A.ptr[101])
    assert e - a >= 2
    col[a:e] = col[a:e][::-1]
    B = P.CSR(A.ptr, col, A.val[np.r_[0:a, np.arange(e - 1, a - 1, -1), e:A.nnz]], A.n)
    b = P.uniform(8, 1, A.n, -1.0, 1.0)
    print("case panels refused: a row with descending columns", flush=True)
    got = dropin_solve(B, b, np.array([0.0, 0.1]), 0, "dynamic",
                       make_params(tol=1e-10, max_iter=5, panel_cols=1024))
    assert got.report["panels"] == 1
