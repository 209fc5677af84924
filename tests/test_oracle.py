"""Pins the oracle (oracle/hbicg_oracle.c) before it is trusted as the
checker.  The reference ships no tests or fixtures for this path, and its
solver sources need a Fujitsu profiler header and MPI that this image lacks,
so the oracle is pinned by:
  * the reference's own BLAS-1 sources compiled under oracle/_ref (bitwise);
  * mathematics that fixes the answer: dense brute-force SpMV, exact
    integer dots, direct sparse solves, a diagonal matrix's closed form, a
    manufactured solution, and the shifted-Krylov collinearity invariant
    b - (A + sigma_j I) x_j = r_seed / (zeta_j pi_j) at every iteration,
    which a dropped term, a wrong sign or a k/k-1 index slip breaks;
  * seed independence and the polynomial-ratio property of seed switching.
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import pyoracle as O
from hbicg import problems as P

_dp = C.POINTER(C.c_double)


def _p(a):
    return a.ctypes.data_as(_dp)


@pytest.fixture(scope="module")
def c0():
    return P.config0()


def small3d(g=10, m=6):
    A = P.conv_diff_3d(g, g, g, 6.0, 0.2)
    b = P.uniform(7, 1, A.n, -1.0, 1.0)
    return A, b, P.logspace_shifts(m, 1e-2, 1.0)


# ---- BLAS-1 against the reference sources -----------------------------------
@pytest.fixture(scope="module")
def refblas():
    if not os.path.exists(O.REF_BLAS_PATH):
        pytest.skip("oracle/_ref/libref_blas.so not built (needs /root/reference at build time)")
    L = C.CDLL(O.REF_BLAS_PATH)
    L.my_daxpy.argtypes = [C.c_int, C.c_double, _dp, _dp]
    L.my_ddot.argtypes = [C.c_int, _dp, _dp]
    L.my_ddot.restype = C.c_double
    L.my_dscal.argtypes = [C.c_int, C.c_double, _dp]
    L.my_openmp_ddot_v3.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
    return L


def test_blas_matches_reference_sources(refblas):
    L = O.lib()
    rng = np.random.default_rng(0)
    n = 10007
    x = rng.standard_normal(n)
    y = rng.standard_normal(n)
    for a in (0.3, -1.7, 1e-5):
        y1, y2 = y.copy(), y.copy()
        L.orc_daxpy(n, a, _p(x), _p(y1))
        refblas.my_daxpy(n, a, _p(x), _p(y2))
        assert np.array_equal(y1, y2)
        z1, z2 = x.copy(), x.copy()
        L.orc_dscal(n, a, _p(z1))
        refblas.my_dscal(n, a, _p(z2))
        assert np.array_equal(z1, z2)
    assert L.orc_ddot(n, _p(x), _p(y)) == refblas.my_ddot(n, _p(x), _p(y))
    tmp1, tmp2 = np.zeros(n), np.zeros(n)
    g = np.zeros(1)
    v3 = L.orc_ddot_v3(n, _p(x), _p(y), _p(tmp1))
    refblas.my_openmp_ddot_v3(n, _p(x), _p(y), _p(tmp2), _p(g))
    assert v3 == g[0]
    # serial ascending order, not numpy's pairwise sum
    acc = 0.0
    for i in range(n):
        acc += float(x[i] * y[i])
    assert v3 == acc


def test_dot_exact_on_integers():
    L = O.lib()
    rng = np.random.default_rng(1)
    x = rng.integers(-1000, 1000, 50000).astype(np.float64)
    y = rng.integers(-1000, 1000, 50000).astype(np.float64)
    assert L.orc_ddot(x.size, _p(x), _p(y)) == float(np.dot(x.astype(np.int64), y.astype(np.int64)))


# ---- SpMV ----------------------------------------------------------------------
def test_mult_dense_bruteforce():
    rng = np.random.default_rng(2)
    for n in (1, 5, 33, 64):
        D = rng.standard_normal((n, n)) * (rng.random((n, n)) < 0.3)
        rows, cols = np.nonzero(D)
        ptr = np.zeros(n + 1, dtype=np.uint32)
        ptr[1:] = np.cumsum(np.bincount(rows, minlength=n))
        A = P.CSR(ptr, cols.astype(np.uint32), D[rows, cols], n)
        x = rng.standard_normal(n)
        np.testing.assert_allclose(O.mult(A, x), D @ x, rtol=1e-13, atol=1e-13)


def test_mult_rowwise_sequential(c0):
    """Each row summed from 0.0 in CSR order (src/matrix.c:506-515)."""
    x = P.uniform(5, 5, c0.n, -1, 1)
    y = O.mult(c0.A, x)
    A = c0.A
    for i in (0, 1, 63, 64, 2000, c0.n - 1):
        t = 0.0
        for e in range(A.ptr[i], A.ptr[i + 1]):
            t += float(A.val[e] * x[A.col[e]])
        assert y[i] == t


def test_stencil_against_scipy(c0):
    x = P.uniform(5, 6, c0.n, -1, 1)
    np.testing.assert_allclose(O.mult(c0.A, x), c0.A.to_scipy() @ x, rtol=1e-14, atol=1e-14)
    assert c0.A.nnz == 20224
    row = 64 * 10 + 10
    s, e = c0.A.ptr[row], c0.A.ptr[row + 1]
    assert list(c0.A.col[s:e]) == [row - 64, row - 1, row, row + 1, row + 64]
    assert list(c0.A.val[s:e]) == [-1.3, -1.3, 4.01, -0.7, -0.7]


# ---- whole solve ------------------------------------------------------------------
def test_c0_converges_true_residual(c0):
    res = O.solve(c0.A, c0.b, c0.sigma, 0, tol=c0.tol)
    assert res.k < 2001
    assert np.all(res.stop_iter > 0)
    tr = O.true_residuals(c0.A, res.x, c0.b, c0.sigma)
    assert tr.max() < 5 * c0.tol


def test_direct_solve_agreement(c0):
    res = O.solve(c0.A, c0.b, c0.sigma, 0, tol=1e-12)
    S = c0.A.to_scipy().tocsc()
    import scipy.sparse as sp
    for j in (0, 7, 15):
        xe = spla.spsolve((S + c0.sigma[j] * sp.identity(c0.n, format="csc")).tocsc(), c0.b)
        assert np.linalg.norm(res.x[j] - xe) / np.linalg.norm(xe) < 1e-9


@pytest.mark.parametrize("seed,cap,tol", [(0, 5, 1e-13), (0, 37, 1e-9), (3, 60, 1e-9),
                                          (15, 100, 1e-9)])
def test_collinearity_invariant(c0, seed, cap, tol):
    """After any number of iterations, with or without seed switches, the true
    residual of every unstopped shift is the seed residual over zeta_j pi_j.
    The deviation allowed is the residual gap roundoff builds up (about
    2e-10*||b|| by iteration 100); a flipped sign, a pi[k]/pi[k-1] slip or a
    dropped zeta in one vector update gives deviations of 40..8e5*||b||."""
    res = O.solve(c0.A, c0.b, c0.sigma, seed, tol=1e-10, max_iter=cap)
    s = res.seed_hist[res.k - 1]
    S = c0.A.to_scipy()
    nb = np.linalg.norm(c0.b)
    active = [j for j in range(c0.M) if res.stop_iter[j] == 0]
    assert len(active) > 0
    biggest = 0.0
    for j in active:
        rj = c0.b - (S @ res.x[j] + c0.sigma[j] * res.x[j])
        if j == s:
            pred = res.r
        else:
            pred = res.r / (res.zeta[j] * res.pi_k[j])
            biggest = max(biggest, np.linalg.norm(pred) / nb)
        assert np.linalg.norm(rj - pred) / nb < tol, (j, s)
    assert biggest > 1e3 * tol   # the check is not vacuous


def test_switching_happens_and_is_seed_independent(c0):
    base = O.solve(c0.A, c0.b, c0.sigma, 0, tol=c0.tol)
    for seed in (5, 15):
        res = O.solve(c0.A, c0.b, c0.sigma, seed, tol=c0.tol)
        assert len(res.switches) >= 1
        k_sw, m = res.switches[0]
        # the new seed is the unconverged shift with the largest estimate
        row = res.res_hist[k_sw]
        cand = [j for j in range(c0.M) if res.stop_iter[j] == 0 or res.stop_iter[j] > k_sw]
        assert m == max(cand, key=lambda j: (row[j], -j))
        rel = np.linalg.norm(res.x - base.x, axis=1) / np.linalg.norm(base.x, axis=1)
        assert rel.max() < 1e-8


def test_diagonal_closed_form():
    n = 7
    d = np.array([1.0, 2.0, 3.5, 4.0, 5.25, 6.0, 8.0])
    A = P.CSR(np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), d, n)
    b = np.array([1.0, -2.0, 0.5, 3.0, -1.0, 2.0, 0.25])
    sigma = np.array([0.0, 0.5, 1.5])
    res = O.solve(A, b, sigma, 1, tol=1e-13, max_iter=50)
    for j in range(3):
        np.testing.assert_allclose(res.x[j], b / (d + sigma[j]), rtol=1e-11)


def test_manufactured_solution():
    """b = (A + sigma_s I) 1 makes the seed solution all ones
    (archive/main_shifted.c:143-151)."""
    A, _, sigma = small3d()
    s = 2
    ones = np.ones(A.n)
    b = O.mult(A, ones) + sigma[s] * ones
    res = O.solve(A, b, sigma, s, tol=1e-13)
    np.testing.assert_allclose(res.x[s], ones, rtol=0, atol=1e-10)


def test_single_shift_is_plain_solve():
    A, b, sigma = small3d()
    res = O.solve(A, b, sigma[:1], 0, tol=1e-12)
    import scipy.sparse as sp
    xe = spla.spsolve((A.to_scipy() + sigma[0] * sp.identity(A.n)).tocsc(), b)
    assert np.linalg.norm(res.x[0] - xe) / np.linalg.norm(xe) < 1e-10


def test_switch_rescale_matches_polynomial_ratio(c0):
    """A switch re-expresses the archives in the new seed's frame; rerunning
    from the start with that seed reproduces the iterates to roundoff: the
    final x of a run that switched 0 -> m agrees with the run seeded at m."""
    res = O.solve(c0.A, c0.b, c0.sigma, 15, tol=c0.tol)
    direct = O.solve(c0.A, c0.b, c0.sigma, res.switches[0][1], tol=c0.tol)
    rel = np.linalg.norm(res.x - direct.x, axis=1) / np.linalg.norm(direct.x, axis=1)
    assert rel.max() < 1e-8


def test_cap_returns_max_iter_plus_one(c0):
    res = O.solve(c0.A, c0.b, c0.sigma, 0, tol=1e-30, max_iter=10)
    assert res.k == 11
    assert np.all(res.stop_iter == 0)


def test_thread_count_independent(c0):
    """Parallel sections never change a rounding (src/openmp_vector.c:68-74
    sums serially); the result is the same at any thread count."""
    import subprocess
    import sys
    code = ("import sys; sys.path[:0]=['oracle','hybrid-bicg_amd'];"
            "import numpy as np, pyoracle as O; from hbicg import problems as P;"
            "p=P.config0(); r=O.solve(p.A,p.b,p.sigma,8,tol=p.tol);"
            "np.save(sys.argv[1], r.x)")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    outs = []
    for t in (1, 3):
        f = f"/tmp/hbicg_thr{t}.npy"
        env = dict(os.environ, OMP_NUM_THREADS=str(t))
        subprocess.check_call([sys.executable, "-c", code, f], cwd=root, env=env)
        outs.append(np.load(f))
    assert np.array_equal(outs[0], outs[1])


# ---- COO -> CSR (src/matrix.c:125-187, 206-232) ------------------------------
def _coo_bruteforce(rows, row, col, val):
    """Row by row, each row's entries in input order: what a stable sort by
    row gives, spelled out with no sort at all."""
    order = [e for r in range(rows) for e in range(len(row)) if row[e] == r]
    ptr = np.zeros(rows + 1, dtype=np.uint32)
    for r in row:
        ptr[r + 1] += 1
    return np.cumsum(ptr).astype(np.uint32), np.asarray(col)[order], np.asarray(val)[order]


@pytest.mark.parametrize("rows,nz,seed", [(1, 5, 0), (7, 60, 1), (33, 500, 2), (5, 0, 3)])
def test_coo2csr_bruteforce(rows, nz, seed):
    """Duplicated (row, col) pairs and many entries per row: a sort that is
    not stable, or a merge that took the right run on ties, reorders them."""
    rng = np.random.default_rng(seed)
    row = rng.integers(0, rows, nz).astype(np.uint32)
    col = rng.integers(0, 4, nz).astype(np.uint32)
    val = rng.standard_normal(nz)
    ptr, oc, ov = O.coo2csr(rows, row, col, val)
    eptr, ec, ev = _coo_bruteforce(rows, row, col, val)
    assert np.array_equal(ptr, eptr)
    assert np.array_equal(oc, ec) and np.array_equal(ov, ev)


def test_coo2csr_numpy_stable_argsort():
    """Larger case against numpy's stable sort (a textbook stable merge
    sort), with empty rows."""
    rng = np.random.default_rng(7)
    rows, nz = 5000, 300_000
    row = (rng.integers(0, rows // 2, nz) * 2).astype(np.uint32)   # odd rows empty
    col = rng.integers(0, rows, nz).astype(np.uint32)
    val = rng.standard_normal(nz)
    ptr, oc, ov = O.coo2csr(rows, row, col, val)
    perm = np.argsort(row, kind="stable")
    assert np.array_equal(oc, col[perm]) and np.array_equal(ov, val[perm])
    assert np.array_equal(ptr, np.concatenate([[0], np.cumsum(np.bincount(row, minlength=rows))]))
    assert np.all(ptr[2::2] == ptr[1:-1:2])


def test_coo2csr_rejects_row_out_of_range():
    with pytest.raises(ValueError):
        O.coo2csr(3, [0, 3], [0, 0], [1.0, 2.0])


# ---- true residual, both directions -------------------------------------------
def test_true_residuals_match_scipy_on_unconverged_x(c0):
    """||b - (A + sigma_j I) x_j|| / ||b|| (src/solver_dynamic.c:405-421) on
    random, far-from-converged x_j, where every residual is O(1), against
    scipy's CSR product and numpy norms to 1e-13: a dropped sigma term, a
    sign error, a square where a root belongs or the norm of b taken
    squared all move the value by O(1).  A 1-D Laplacian with a closed-form
    residual checks the same function without scipy."""
    rng = np.random.default_rng(31)
    x = rng.uniform(-1, 1, (c0.M, c0.n))
    got = O.true_residuals(c0.A, x, c0.b, c0.sigma)
    S = c0.A.to_scipy()
    want = np.array([np.linalg.norm(c0.b - (S @ x[j] + c0.sigma[j] * x[j])) / np.linalg.norm(c0.b)
                     for j in range(c0.M)])
    assert want.min() > 0.5
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=0)
    # tridiag(-1, 2, -1), x = 1: (A x)_i = 1 at both ends, 0 inside; with
    # b = e_0 the residual of shift s is b - A x - s x, whose norm squared is
    # (1 - 1 - s)^2 + (n - 2) s^2 + (1 + s)^2
    n = 50
    rows = np.arange(n)
    ptr = np.zeros(n + 1, dtype=np.uint32)
    cols, vals = [], []
    for i in rows:
        for j, v in ((i - 1, -1.0), (i, 2.0), (i + 1, -1.0)):
            if 0 <= j < n:
                cols.append(j)
                vals.append(v)
        ptr[i + 1] = len(cols)
    T = P.CSR(ptr, np.array(cols, dtype=np.uint32), np.array(vals), n)
    sig = np.array([0.0, 0.25, 2.0])
    b = np.zeros(n)
    b[0] = 1.0
    got = O.true_residuals(T, np.ones((3, n)), b, sig)
    want = np.sqrt(sig ** 2 + (n - 2) * sig ** 2 + (1 + sig) ** 2)
    np.testing.assert_allclose(got, want, rtol=1e-15)


# ---- breakdown semantics -----------------------------------------------------
def test_breakdowns_run_to_the_cap(c0):
    """The reference does not trap a breakdown: b = 0 gives (r^, s) = 0 and
    alpha~ = 0/0 at the first step (src/solver_dynamic.c:167); with A = 2I
    and sigma_s = 0, alpha~ = (b,b)/(b,2b) = 1/2 exactly, q = 0 and
    omega~ = 0/0 (:189).  NaN then meets no stopping rule and the loop runs
    to max_iter + 1."""
    r = O.solve(c0.A, np.zeros(c0.n), c0.sigma, 0, tol=c0.tol, max_iter=25)
    assert r.k == 26 and np.all(np.isnan(r.x)) and r.status == 0
    n = 300
    A = P.CSR(np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.full(n, 2.0), n)
    r = O.solve(A, P.uniform(9, 1, n, -1.0, 1.0), np.array([0.0, 1.0]), 0, tol=1e-10, max_iter=12)
    assert r.k == 13 and r.alpha[1] == 0.5 and np.isnan(r.omega[1])
    assert np.all(np.isnan(r.x))


def test_nan_shift_ends_without_switch_target(c0):
    """A NaN shift's estimate is never above the seed's, so when the seed
    stops there is no shift to switch to (the reference's max_sigma is then
    uninitialised): the solve ends after that iteration with the
    NO_SWITCH_TARGET status, the finite shifts solved as without the NaN
    one."""
    sigma = np.array([c0.sigma[0], np.nan, c0.sigma[5]])
    r = O.solve(c0.A, c0.b, sigma, 0, tol=c0.tol)
    clean = O.solve(c0.A, c0.b, sigma[[0, 2]], 0, tol=c0.tol)
    assert r.status == O.NO_SWITCH_TARGET
    assert r.k == r.stop_iter[0] + 1 and r.stop_iter[1] == 0
    assert r.stop_iter[0] == clean.stop_iter[0] and r.stop_iter[2] == clean.stop_iter[1]
    assert np.array_equal(r.x[[0, 2]], clean.x[:, :])
    assert np.all(np.isnan(r.x[1]))


# ---- the reference on P MPI ranks ---------------------------------------------
def test_ranks_split_row_sums():
    """src/solver_dynamic.c:158-160: a row's diagonal-block entries are summed,
    then its off-diagonal-block sum is added.  Row 0 holds 1e16, 1, -1e16 at
    columns 0, 2, 1: summed in CSR order the 1 is lost (1e16 + 1 rounds to
    1e16), while on 2 ranks (rows and columns 0-1 | 2-3) the diagonal block
    cancels exactly and the off-diagonal 1 survives.  Integer data sums
    exactly in any order, so every rank count gives the dense product."""
    ptr = np.array([0, 3, 4, 5, 6], dtype=np.uint32)
    A = P.CSR(ptr, np.array([0, 2, 1, 1, 2, 3], dtype=np.uint32),
              np.array([1e16, 1.0, -1e16, 2.0, 3.0, 4.0]), 4)
    x = np.ones(4)
    assert O.mult_ranks(A, x, 1)[0] == 0.0 and O.mult(A, x)[0] == 0.0
    assert O.mult_ranks(A, x, 2)[0] == 1.0
    rng = np.random.default_rng(51)
    B = P.conv_diff_2d(9, 7, 4, 1)            # integer entries 4, -2, 0 ... exact products
    xi = rng.integers(-50, 50, B.n).astype(np.float64)
    dense = B.to_scipy() @ xi
    for R in (1, 2, 4, 8, 16):
        assert np.array_equal(O.mult_ranks(B, xi, R), dense)


def test_ranks_dot_partition_and_allreduce():
    """Rows m = 10 on 4 ranks: 3, 3, 2, 2 (src/matrix.c:296-299: the first
    m mod P ranks take one more row); each rank sums its products serially,
    MPI_Allreduce adds the four partials by recursive doubling,
    (p0 + p1) + (p2 + p3).  Random data makes any other partition or
    combination order a different double."""
    rng = np.random.default_rng(52)
    x = rng.standard_normal(10) * 10.0 ** rng.integers(-8, 8, 10)
    y = rng.standard_normal(10)
    t = x * y
    part = [0.0, 0.0, 0.0, 0.0]
    for p, (a, b) in enumerate(((0, 3), (3, 6), (6, 8), (8, 10))):
        acc = 0.0
        for i in range(a, b):
            acc += t[i]
        part[p] = acc
    want = (part[0] + part[1]) + (part[2] + part[3])
    assert O.dot_ranks(x, y, 4) == want
    serial = 0.0
    for v in t:
        serial += v
    assert O.dot_ranks(x, y, 1) == serial
    # crafted: one element per rank, 1e16, 1, -1e16, 1 -- serially 1, by
    # recursive doubling (1e16 + 1) + (-1e16 + 1) = 0
    a = np.array([1e16, 1.0, -1e16, 1.0])
    assert O.dot_ranks(a, np.ones(4), 1) == 1.0
    assert O.dot_ranks(a, np.ones(4), 4) == (a[0] + a[1]) + (a[2] + a[3])


def test_ranks_solves_agree(c0):
    """The 64x64 workload on 1..16 ranks: the same k, stop iterations and
    switches, solutions within 1e-10 of the one-rank run and converged; one
    rank is orc_shifted_lopbicg itself."""
    for seed in (0, 8):
        base = O.solve(c0.A, c0.b, c0.sigma, seed, tol=c0.tol)
        one = O.solve(c0.A, c0.b, c0.sigma, seed, tol=c0.tol, ranks=1)
        assert one.k == base.k and np.array_equal(one.x, base.x)
        for R in (2, 4, 16):
            r = O.solve(c0.A, c0.b, c0.sigma, seed, tol=c0.tol, ranks=R)
            assert r.k == base.k and np.array_equal(r.stop_iter, base.stop_iter)
            assert r.switches == base.switches
            rel = np.linalg.norm(r.x - base.x, axis=1) / np.linalg.norm(base.x, axis=1)
            assert 0 < rel.max() <= 1e-10, rel.max()
            assert O.true_residuals(c0.A, r.x, c0.b, c0.sigma).max() < 5 * c0.tol
