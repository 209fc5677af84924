"""GPU parity of the drop-in entry points against the oracle (the CPU
restatement of src/solver_dynamic.c), called through the C-ABI.

exact_order=1 runs every dot and SpMV row in the reference's summation
order, so the whole GPU solve must equal the oracle bit for bit: returned k,
per-shift stop iterations, seed-switch log, x_set and the returned r.
The default (tree reductions) must stay within the parity bar: k and every
switch iteration within +-2 of the oracle, and solutions within 1e-9
relative (x) of the oracle's, with the true residual of every shift recomputed.
"""
import numpy as np
import pytest

import pyoracle as O
from hbicg import dropin_solve, make_params
from hbicg import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c0():
    return P.config0()


@pytest.fixture(scope="module")
def oracle_c0(c0):
    return {s: O.solve(c0.A, c0.b, c0.sigma, s, tol=c0.tol) for s in (0, 8, 15)}


@pytest.mark.parametrize("value_index", [True, False])
@pytest.mark.parametrize("variant", ["normal", "static", "dynamic"])
@pytest.mark.parametrize("seed", [0, 8, 15])
def test_exact_order_bitwise_c0(c0, oracle_c0, variant, seed, value_index):
    """Exact order equals the oracle bit for bit, with the SpMV reading the
    matrix through its value table and diagonal tiles (the default) and
    reading 8-byte CSR values."""
    print(f"case exact c0 variant={variant} seed={seed} value_index={value_index}", flush=True)
    ref = oracle_c0[seed]
    got = dropin_solve(c0.A, c0.b, c0.sigma, seed, variant,
                       make_params(tol=c0.tol, exact_order=True, value_index=value_index))
    assert got.report["value_table"] == (3 if value_index else 0)
    assert got.k == ref.k
    np.testing.assert_array_equal(got.stop_iter, ref.stop_iter)
    assert got.switches == ref.switches
    assert np.array_equal(got.x, ref.x)
    assert np.array_equal(got.r, ref.r)


@pytest.mark.parametrize("variant", ["normal", "static", "dynamic"])
@pytest.mark.parametrize("seed", [0, 8, 15])
def test_tree_parity_c0(c0, oracle_c0, variant, seed):
    print(f"case tree c0 variant={variant} seed={seed}", flush=True)
    ref = oracle_c0[seed]
    got = dropin_solve(c0.A, c0.b, c0.sigma, seed, variant, make_params(tol=c0.tol))
    assert abs(got.k - ref.k) <= 2
    assert np.all(np.abs(got.stop_iter - ref.stop_iter) <= 2)
    assert len(got.switches) == len(ref.switches)
    for (kg, sg), (kr, sr) in zip(got.switches, ref.switches):
        assert abs(kg - kr) <= 2 and sg == sr
    rel = np.linalg.norm(got.x - ref.x, axis=1) / np.linalg.norm(ref.x, axis=1)
    assert rel.max() <= 1e-9, rel
    tr = O.true_residuals(c0.A, got.x, c0.b, c0.sigma)
    assert tr.max() <= 10 * c0.tol, tr


def _same_up_to_window(got, base, seed, switches):
    """Default mode, two windows: without a seed switch the seed's own
    iteration never reads a shift's vectors, so k, every stop iteration and
    r agree exactly, and every x_j (the seed's too: its x is deferred to the
    window flush like the shifts') only by the roundings of the regrouped
    window sums.  After a switch the new seed continues from its windowed
    p_j, so the rest of the run moves by roundings too."""
    rel = np.linalg.norm(got.x - base.x, axis=1) / np.linalg.norm(base.x, axis=1)
    if not switches:
        assert got.k == base.k
        np.testing.assert_array_equal(got.stop_iter, base.stop_iter)
        assert np.array_equal(got.r, base.r)
        assert rel.max() <= 1e-10, rel.max()
    else:
        assert abs(got.k - base.k) <= 2
        assert np.all(np.abs(got.stop_iter - base.stop_iter) <= 2)
        assert [s for _, s in got.switches] == [s for _, s in base.switches]
        assert rel.max() <= 1e-9, rel.max()


@pytest.mark.parametrize("seed", [0, 8])
def test_variants_equal(c0, oracle_c0, seed):
    """The three entry points compute the same result (src/solver_{normal,
    static,dynamic}.c share their arithmetic).  static and dynamic run the
    same window and agree bit for bit; normal applies every iteration
    (window 1)."""
    outs = [dropin_solve(c0.A, c0.b, c0.sigma, seed, v, make_params(tol=c0.tol)) for v in
            ("normal", "static", "dynamic")]
    for o in outs[1:]:
        _same_up_to_window(o, outs[0], seed, oracle_c0[seed].switches)
    assert outs[1].k == outs[2].k
    assert np.array_equal(outs[1].x, outs[2].x)


@pytest.mark.parametrize("window", [1, 3, 8, 16])
def test_window_bitwise_independent_exact(c0, window):
    """Exact order: deferring the shift updates over a window replays the
    same per-element operations, so x_set must not depend on the window."""
    mk = lambda w: make_params(tol=c0.tol, window=w, exact_order=True)
    base = dropin_solve(c0.A, c0.b, c0.sigma, 8, "static", mk(1))
    got = dropin_solve(c0.A, c0.b, c0.sigma, 8, "static", mk(window))
    assert got.k == base.k
    assert np.array_equal(got.x, base.x)


@pytest.mark.parametrize("seed", [0, 8])
@pytest.mark.parametrize("window", [3, 8, 16, 33, 64, 256, 1024])
def test_window_independent_default(c0, oracle_c0, window, seed):
    """Default mode: the combined-coefficient flush regroups each window's
    sums; results move only by roundings as the window changes.  Windows of
    256 and 1024 outlast the solve (about 120 iterations): one flush at the
    end (or at the seed switch and at the end), x_j only at the last."""
    mk = lambda w: make_params(tol=c0.tol, window=w)
    base = dropin_solve(c0.A, c0.b, c0.sigma, seed, "static", mk(1))
    got = dropin_solve(c0.A, c0.b, c0.sigma, seed, "static", mk(window))
    assert got.report["window"] == window
    _same_up_to_window(got, base, seed, oracle_c0[seed].switches)


def test_default_window_choice(c0):
    """static/dynamic default to a window of min(512, max_iter) when the
    ring buffers fit (they do at n = 4096); normal runs window 1."""
    for variant, mi, w in (("static", 2000, 512), ("dynamic", 300, 300), ("normal", 2000, 1)):
        got = dropin_solve(c0.A, c0.b, c0.sigma, 0, variant, make_params(tol=c0.tol, max_iter=mi))
        assert got.report["window"] == w, (variant, got.report["window"])


@pytest.mark.parametrize("seed", [0, 8])
def test_window_tail_path(c0, oracle_c0, seed):
    """growth_limit below 1 closes every window after its first iteration
    with that iteration applied on its own (the tail path of the flush):
    the result is the per-iteration order, close to window 1 and to the
    oracle, and every shift meets the stopping rule."""
    base = dropin_solve(c0.A, c0.b, c0.sigma, seed, "static", make_params(tol=c0.tol, window=1))
    got = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic",
                       make_params(tol=c0.tol, window=64, growth_limit=0.5))
    _same_up_to_window(got, base, seed, oracle_c0[seed].switches)
    tr = O.true_residuals(c0.A, got.x, c0.b, c0.sigma)
    assert tr.max() <= 10 * c0.tol, tr.max()


@pytest.mark.parametrize("seed", [0, 8])
def test_fused_launches_bitwise_tail_path(c0, seed):
    """The tail path (growth_limit 0.5: every window closes after its first
    iteration, which the flush applies on its own) through the four- and
    five-launch iterations: bit for bit the eight-launch result."""
    print(f"case fused levels, tail path, seed={seed}", flush=True)
    mk = lambda f: make_params(tol=c0.tol, window=64, growth_limit=0.5, fused=f)
    ref = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", mk(0))
    for level in (1, 3):
        a = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", mk(level))
        assert a.k == ref.k and a.switches == ref.switches
        np.testing.assert_array_equal(a.stop_iter, ref.stop_iter)
        assert np.array_equal(a.x, ref.x) and np.array_equal(a.r, ref.r)
        assert not (a.report["status"] & 16)


@pytest.mark.parametrize("seed", [0, 8])
def test_x_zero_never_reads_x(c0, seed):
    """params.x_zero: x_set's incoming contents are never read.  A device
    plan handed an x_set full of NaN returns bit for bit what it returns
    from a zeroed one: in the default mode with the window of 512 (the
    first window's flush writes x_j) and window 1 (a flush every
    iteration), and in exact order (x_j cleared at the start); seed 8
    switches at iteration 102.  The host drop-in with x_zero neither uploads
    nor clears x_loc_set: NaN in, the zeroed run's x out."""
    import torch

    from hbicg import Plan
    b = torch.as_tensor(c0.b).cuda()
    for exact in (False, True):
        for window in (0, 1):
            print(f"case x_zero seed={seed} exact={exact} window={window}", flush=True)
            mk = lambda xz: make_params(tol=c0.tol, exact_order=exact, window=window, graph=1, x_zero=xz)
            pl0 = Plan(c0.A, c0.M, mk(False))
            k0, x0, _ = pl0.solve(b, c0.sigma, seed)
            pl0.close()
            pl1 = Plan(c0.A, c0.M, mk(True))
            xn = torch.full((c0.M, c0.A.n), float("nan"), dtype=torch.float64, device="cuda")
            k1, x1, rep = pl1.solve(b, c0.sigma, seed, x=xn)
            pl1.close()
            assert k1 == k0 and torch.equal(x1, x0), (exact, window)
    ref = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", make_params(tol=c0.tol))
    got = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", make_params(tol=c0.tol, x_zero=True),
                       x0=np.full((c0.M, c0.A.n), np.nan))
    assert got.k == ref.k and np.array_equal(got.x, ref.x)


@pytest.mark.parametrize("seed", [0, 8, 15])
def test_against_golden_fixture(c0, seed):
    """GPU results against the committed oracle fixture
    (tests/golden/c0_oracle.npz): exact order bit for bit, default order
    within the parity bar."""
    import os
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c0_oracle.npz"))
    ex = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", make_params(tol=c0.tol, exact_order=True))
    assert ex.k == int(g[f"k_{seed}"])
    assert np.array_equal(ex.stop_iter, g[f"stop_{seed}"])
    assert np.array_equal(ex.x[:, ::41], g[f"xs_{seed}"])
    assert np.array_equal(ex.x.sum(axis=1), g[f"xsum_{seed}"])
    tr = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", make_params(tol=c0.tol))
    assert abs(tr.k - int(g[f"k_{seed}"])) <= 2
    assert np.all(np.abs(tr.stop_iter - g[f"stop_{seed}"]) <= 2)
    gs = g[f"switch_{seed}"]
    assert len(tr.switches) == len(gs)
    for (kg, sg), (kr, sr) in zip(tr.switches, gs):
        assert abs(kg - kr) <= 2 and sg == sr
    rel = np.abs(tr.x[:, ::41] - g[f"xs_{seed}"]).max() / np.abs(g[f"xs_{seed}"]).max()
    assert rel <= 1e-9


@pytest.mark.parametrize("seed", [0, 8])
def test_residual_trace(c0, oracle_c0, seed):
    """The per-iteration residual estimates recorded on the device
    (src/solver_dynamic.c:286-292) equal the oracle's history: bit for bit in
    exact order, and in the default order within the drift of the tree
    reductions while the estimates are far above the stopping threshold."""
    ref = oracle_c0[seed]
    ex = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic",
                      make_params(tol=c0.tol, exact_order=True, trace=True))
    assert ex.trace.shape == (ref.k, c0.sigma.shape[0])
    assert np.array_equal(ex.trace, ref.res_hist[:ref.k])
    # every shift's last recorded estimate is the one that met the rule
    for j in range(c0.sigma.shape[0]):
        kj = ex.stop_iter[j]
        assert kj > 0 and ex.trace[kj, j] <= c0.tol * (1 + 1e-12) and np.all(ex.trace[kj + 1:, j] == -1.0)
    tr = dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", make_params(tol=c0.tol, trace=True))
    m = min(tr.k, ref.k) - 3
    a, b = tr.trace[1:m], ref.res_hist[1:m]
    live = (a > 0) & (b > 1e-6)
    # the tree reductions perturb alpha and beta by roundings; BiCG's
    # irregular convergence and the 1/(zeta pi) factor of the shifts amplify
    # that in the estimates (3.1e-5 relative at worst on this problem), while
    # a wrong term would move them by O(1)
    dev = np.abs(a[live] - b[live]) / b[live]
    print("worst relative trace drift", dev.max(), flush=True)
    assert dev.max() <= 1e-4
    # a plan without trace=True has none
    assert dropin_solve(c0.A, c0.b, c0.sigma, seed, "dynamic", make_params(tol=c0.tol)).trace is None


def test_sweep_driver_conventions_1024_shifts(c0):
    """The shift-count sweep's largest case (src/main_seed_diff.c:113-155:
    b = 1, sigma_i = (i+1) * 0.01 / M, seed = M/2 - 1) at M = 1024 on the
    64 x 64 operator: exact order equals the oracle bit for bit, the default
    order meets the parity bar.  With 1024 shifts packed into (0, 0.01] one
    shift's pi_k passes close to zero and its true residual drifts above the
    estimate that stopped it: the oracle itself ends with 1.4e-9 at shift 542
    (every other shift below 1e-9), the default order with about 4e-9 at one
    other shift.  So the true residuals are held to 10 tol except for as
    many shifts as the oracle has above it (at least one), and those to ten
    times the oracle's worst."""
    M = 1024
    sigma = (np.arange(M) + 1) * (0.01 / M)
    b = np.ones(c0.n)
    seed = M // 2 - 1
    print(f"case sweep conventions M={M} seed={seed}", flush=True)
    ref = O.solve(c0.A, b, sigma, seed, tol=c0.tol)
    tro = O.true_residuals(c0.A, ref.x, b, sigma)
    n_above = max(1, int((tro > 10 * c0.tol).sum()))
    ex = dropin_solve(c0.A, b, sigma, seed, "dynamic", make_params(tol=c0.tol, exact_order=True))
    assert ex.k == ref.k and ex.switches == ref.switches
    np.testing.assert_array_equal(ex.stop_iter, ref.stop_iter)
    assert np.array_equal(ex.x, ref.x)
    for variant in ("normal", "dynamic"):
        got = dropin_solve(c0.A, b, sigma, seed, variant, make_params(tol=c0.tol))
        assert abs(got.k - ref.k) <= 2
        assert len(got.switches) == len(ref.switches)
        rel = np.linalg.norm(got.x - ref.x, axis=1) / np.linalg.norm(ref.x, axis=1)
        assert rel.max() <= 1e-9, rel.max()
        assert np.all(got.stop_iter > 0)
        tr = O.true_residuals(c0.A, got.x, b, sigma)
        print(variant, "true residual max", tr.max(), "oracle", tro.max(), flush=True)
        assert int((tr > 10 * c0.tol).sum()) <= n_above
        assert tr.max() <= 10 * max(tro.max(), 10 * c0.tol)


@pytest.mark.parametrize("exact", [False, True])
def test_value_table_bitwise(c0, exact):
    """A matrix with few distinct values is read through a one-byte value
    table; the products are the same doubles, so every result is bitwise
    equal to reading the 8-byte values (value_index=0)."""
    mk = lambda vi: make_params(tol=c0.tol, exact_order=exact, value_index=vi)
    a = dropin_solve(c0.A, c0.b, c0.sigma, 8, "dynamic", mk(True))
    b = dropin_solve(c0.A, c0.b, c0.sigma, 8, "dynamic", mk(False))
    assert a.report["value_table"] == 3 and b.report["value_table"] == 0
    assert a.report["diag_tiles"] > 0 and b.report["diag_tiles"] == 0
    assert a.k == b.k and a.switches == b.switches
    assert np.array_equal(a.x, b.x) and np.array_equal(a.r, b.r)


def _rows_csr(rows, n):
    """CSR from per-row lists of (column, value), in the given order."""
    ptr = np.zeros(n + 1, np.uint32)
    ptr[1:] = np.cumsum([len(r) for r in rows])
    col = np.array([c for r in rows for c, _ in r], np.uint32)
    val = np.array([v for r in rows for _, v in r], np.float64)
    return P.CSR(ptr=ptr, col=col, val=val, n=n)


def _mixed_tile_matrix(vary=False):
    """5-point 64x64 operator edited so that its 256-row tiles take every
    path of the value-table SpMV: rows of tile 3 list their entries in
    descending column order (CSR form), one row of tile 5 has entries on 20
    more diagonals (more than 16: CSR form), every row of tile 7 has two more
    diagonals and one of them four more (11 diagonals), boundary rows miss
    diagonals.  Every diagonal tile then holds one value per diagonal (read
    as presence bits and the diagonals' values); vary: the main diagonal of
    tile 9 alternates between two values, so every diagonal tile is read
    through index bytes (16 per row in tile 7)."""
    A = P.conv_diff_2d(64, 64)
    n = A.n
    ptr = np.asarray(A.ptr, np.int64)
    rows = [list(zip(A.col[ptr[i]:ptr[i + 1]].tolist(), A.val[ptr[i]:ptr[i + 1]].tolist())) for i in range(n)]
    for i in range(3 * 256, 4 * 256):
        rows[i] = sorted(rows[i], key=lambda e: -e[0])
    r = 5 * 256 + 100
    rows[r] = sorted(rows[r] + [(r + o, 0.01) for o in range(2, 22) if o != 64], key=lambda e: e[0])
    for i in range(7 * 256, 8 * 256):
        extra = [(i - 2, -0.02), (i + 2, 0.02)] + ([(i + o, 0.01) for o in (-4, -3, 3, 4)] if i == 7 * 256 + 9 else [])
        rows[i] = sorted(rows[i] + extra, key=lambda e: e[0])
    if vary:
        for i in range(9 * 256, 10 * 256, 2):
            rows[i] = [(c, v + 0.5 if c == i else v) for c, v in rows[i]]
    return _rows_csr(rows, n)


@pytest.mark.parametrize("vary", [False, True])
@pytest.mark.parametrize("exact", [False, True])
def test_value_table_diagonal_tiles(exact, vary):
    """Tiles stored by diagonals (presence bits and one value per diagonal,
    or one index byte per row and diagonal; no columns) and tiles kept in
    CSR form in one matrix: results bitwise equal to reading the 8-byte
    values; in exact order also equal to the oracle."""
    A = _mixed_tile_matrix(vary)
    plain = P.conv_diff_2d(64, 64)
    rng = np.random.default_rng(21)
    b = rng.uniform(-1.0, 1.0, A.n)
    sigma = np.logspace(-2, 0, 9)
    print(f"case diagonal tiles exact={exact} vary={vary}", flush=True)
    mk = lambda vi: make_params(tol=1e-10, exact_order=exact, value_index=vi)
    a = dropin_solve(A, b, sigma, 0, "dynamic", mk(True))
    c = dropin_solve(A, b, sigma, 0, "dynamic", mk(False))
    full = dropin_solve(plain, b, sigma, 0, "dynamic", mk(True))
    assert a.report["value_table"] > 0 and c.report["value_table"] == 0
    assert full.report["diag_tiles"] == 16 and full.report["const_tiles"] == 16
    assert 0 < a.report["diag_tiles"] <= 14, a.report["diag_tiles"]
    assert a.report["const_tiles"] == (0 if vary else a.report["diag_tiles"])
    assert a.k == c.k and a.switches == c.switches
    assert np.array_equal(a.x, c.x) and np.array_equal(a.r, c.r)
    if exact:
        ref = O.solve(A, b, sigma, 0, tol=1e-10)
        assert a.k == ref.k and np.array_equal(a.x, ref.x) and np.array_equal(a.r, ref.r)


def test_value_table_diagonal_tiles_with_empty_rows(c0):
    """Rows with no entries at all inside diagonal tiles (every index byte
    marks a missing diagonal): bitwise equal to the 8-byte read, and in
    exact order to the oracle."""
    A = c0.A
    ptr = np.asarray(A.ptr, np.int64)
    rows = [list(zip(A.col[ptr[i]:ptr[i + 1]].tolist(), A.val[ptr[i]:ptr[i + 1]].tolist()))
            for i in range(A.n)]
    for i in (5, 300, 301, 1000, 2047, 4095):
        rows[i] = []
    A = _rows_csr(rows, A.n)
    print("case diagonal tiles with empty rows", flush=True)
    for exact in (False, True):
        mk = lambda vi: make_params(tol=c0.tol, exact_order=exact, value_index=vi)
        a = dropin_solve(A, c0.b, c0.sigma, 0, "dynamic", mk(True))
        b = dropin_solve(A, c0.b, c0.sigma, 0, "dynamic", mk(False))
        assert a.report["diag_tiles"] > 0
        assert a.k == b.k and np.array_equal(a.x, b.x) and np.array_equal(a.r, b.r)
        if exact:
            ref = O.solve(A, c0.b, c0.sigma, 0, tol=c0.tol)
            assert a.k == ref.k and np.array_equal(a.x, ref.x)


def test_value_table_256_values_csr_tiles():
    """256 distinct values fill the table; no index byte is left to mark a
    missing diagonal, so every tile stays in CSR form, bitwise equal to
    reading the 8-byte values."""
    A = P.conv_diff_2d(64, 64)
    ptr = np.asarray(A.ptr, np.int64)
    val = np.empty(len(A.val))
    for i in range(A.n):
        for e in range(ptr[i], ptr[i + 1]):
            val[e] = 4.0 + (i % 128) / 1000.0 if A.col[e] == i else -(0.5 + (e % 128) / 1000.0)
    A = P.CSR(ptr=A.ptr, col=A.col, val=val, n=A.n)
    assert len(np.unique(val)) == 256
    rng = np.random.default_rng(22)
    b = rng.uniform(-1.0, 1.0, A.n)
    sigma = np.logspace(-2, 0, 5)
    a = dropin_solve(A, b, sigma, 0, "dynamic", make_params(tol=1e-10, value_index=True))
    c = dropin_solve(A, b, sigma, 0, "dynamic", make_params(tol=1e-10, value_index=False))
    assert a.report["value_table"] == 256 and a.report["diag_tiles"] == 0
    assert a.k == c.k and np.array_equal(a.x, c.x)


@pytest.mark.parametrize("which", ["stencil", "scattered"])
def test_q_formed_before_or_at_gather_bitwise(c0, which):
    """SpMV 2's input q = r - alpha s formed at every gathered column (q_mode
    1) or for every row first and gathered (q_mode 2): the same expression,
    so every result is bitwise equal.  Auto (q_mode 0) forms it at the
    gathered columns for the stencil (16-bit column offsets) and beforehand
    for a power-law matrix with columns spread uniformly (n = 100000, rows of
    1 to 2000 entries, some longer than one SpMV tile; most entries lie
    beyond 16-bit offsets), which here is capped at 40 iterations (it
    converges slowly) and also compared with the oracle at 12 iterations."""
    if which == "stencil":
        A, b, sigma, seed, cap, auto = c0.A, c0.b, c0.sigma, 8, 2000, 0
    else:
        pr = P.config2(n=100000, m=8)
        A, b, sigma, seed, cap, auto = pr.A, pr.b, pr.sigma, pr.seed, 40, 1
    print(f"case q_mode {which}", flush=True)
    mk = lambda qm, it=cap: make_params(tol=1e-10, max_iter=it, q_mode=qm)
    r0 = dropin_solve(A, b, sigma, seed, "dynamic", mk(0))
    r1 = dropin_solve(A, b, sigma, seed, "dynamic", mk(1))
    r2 = dropin_solve(A, b, sigma, seed, "dynamic", mk(2))
    assert (r0.report["q_unfused"], r1.report["q_unfused"], r2.report["q_unfused"]) == (auto, 0, 1)
    for r in (r1, r2):
        assert r.k == r0.k and r.switches == r0.switches
        assert np.array_equal(r.x, r0.x) and np.array_equal(r.r, r0.r)
    if which == "scattered":
        ref = O.solve(A, b, sigma, seed, tol=1e-10, max_iter=12)
        got = dropin_solve(A, b, sigma, seed, "dynamic", mk(0, 12))
        assert got.k == ref.k
        rel = np.linalg.norm(got.x - ref.x, axis=1) / np.linalg.norm(ref.x, axis=1)
        assert rel.max() <= 1e-9, rel.max()


def test_value_table_with_q_formed_first_bitwise():
    """A scattered-column matrix (power-law rows, n = 100000) whose values
    are rounded to 1/8, the diagonal one constant (18 distinct values): its
    SpMV reads the value table in CSR tiles, and SpMV 2 gathers q formed
    beforehand.  Bitwise equal to the 8-byte read and to forming q at the
    gathered columns."""
    pr = P.config2(n=100000, m=8)
    A = pr.A
    ptr = np.asarray(A.ptr, np.int64)
    rows = np.repeat(np.arange(A.n), np.diff(ptr))
    val = np.round(np.asarray(A.val) * 8.0) / 8.0
    diag = np.asarray(A.col) == rows
    val[diag] = 2400.0   # above every row's off-diagonal sum (rows of at most 2000 entries)
    A = P.CSR(ptr=A.ptr, col=A.col, val=val, n=A.n)
    print(f"case value table + q formed first, {len(np.unique(val))} values", flush=True)
    mk = lambda vi, qm: make_params(tol=1e-10, max_iter=40, value_index=vi, q_mode=qm)
    a = dropin_solve(A, pr.b, pr.sigma, pr.seed, "dynamic", mk(True, 0))
    b = dropin_solve(A, pr.b, pr.sigma, pr.seed, "dynamic", mk(False, 1))
    assert a.report["value_table"] > 0 and a.report["q_unfused"] == 1
    assert b.report["value_table"] == 0 and b.report["q_unfused"] == 0
    assert a.k == b.k and np.array_equal(a.x, b.x) and np.array_equal(a.r, b.r)


def test_value_table_not_used_for_many_values():
    """More than 256 distinct values: the SpMV reads the 8-byte values."""
    rng = np.random.default_rng(3)
    n = 600
    A = P.conv_diff_2d(20, 30)
    A = P.CSR(ptr=A.ptr, col=A.col, val=A.val + 1e-3 * rng.random(len(A.val)), n=A.n)
    b = rng.standard_normal(A.n)
    got = dropin_solve(A, b, np.array([0.0, 0.1]), 0, "dynamic", make_params(tol=1e-10))
    assert got.report["value_table"] == 0
    ref = O.solve(A, b, np.array([0.0, 0.1]), 0, tol=1e-10)
    assert abs(got.k - ref.k) <= 2


@pytest.mark.parametrize("M,seed", [(1, 0), (17, 16), (33, 5)])
def test_ragged_sizes(M, seed):
    """n = 1517 (not a multiple of 64: partial tiles in every kernel) and
    shift counts 1, 17 and 33 (partial chunks of 16 in the flush): exact
    order equals the oracle bit for bit; the default order meets the parity
    bar, with every shift's true residual recomputed."""
    A = P.conv_diff_2d(37, 41, diag=4.01, conv=0.3)
    rng = np.random.default_rng(M)
    b = rng.uniform(-1.0, 1.0, A.n)
    sigma = np.logspace(-2, 0, M) if M > 1 else np.array([0.05])
    print(f"case ragged n={A.n} M={M} seed={seed}", flush=True)
    ref = O.solve(A, b, sigma, seed, tol=1e-10)
    ex = dropin_solve(A, b, sigma, seed, "dynamic", make_params(tol=1e-10, exact_order=True))
    assert ex.k == ref.k and ex.switches == ref.switches
    assert np.array_equal(ex.x, ref.x) and np.array_equal(ex.r, ref.r)
    for variant in ("normal", "static", "dynamic"):
        got = dropin_solve(A, b, sigma, seed, variant, make_params(tol=1e-10))
        assert abs(got.k - ref.k) <= 2 and len(got.switches) == len(ref.switches)
        rel = np.linalg.norm(got.x - ref.x, axis=1) / np.linalg.norm(ref.x, axis=1)
        assert rel.max() <= 1e-9, (variant, rel.max())
        tr = O.true_residuals(A, got.x, b, sigma)
        assert tr.max() <= 10 * 1e-10, (variant, tr.max())


@pytest.mark.parametrize("value_index", [True, False])
@pytest.mark.parametrize("seed", [0, 8])
def test_fused_launches_bitwise(c0, seed, value_index):
    """Default mode: the four-launch iteration (alpha~ finalized by the first
    workgroup of SpMV 2's launch, omega~ by the first of the seed update's,
    the scalar stage by the first of the shift kernel's, each handing its
    values to the launch's other workgroups; combined coefficients formed in
    the shift kernel's next workgroups), the five-launch one (the scalar
    stage on its own) and the six-launch one (alpha~ on its own too) are the
    eight-launch iteration's arithmetic: k, stop iterations, switches, x and
    r bit for bit, and no workgroup timed out waiting.  With the value table
    SpMV 2 reads diagonal tiles (alpha~ read after the gathers are issued);
    without it, 8-byte CSR tiles (alpha~ read first).  `normal` opens and
    flushes a window in every iteration, so the flush reads a list the
    stage wrote in the same launch; seed 8 switches at iteration 102."""
    for variant in ("normal", "static", "dynamic"):
        print(f"case fused vs unfused c0 seed={seed} {variant} value_index={value_index}", flush=True)
        mk = lambda f: make_params(tol=c0.tol, fused=f, value_index=value_index)
        ref = dropin_solve(c0.A, c0.b, c0.sigma, seed, variant, mk(0))
        assert not (ref.report["status"] & 16)
        for level in (1, 2, 3):
            a = dropin_solve(c0.A, c0.b, c0.sigma, seed, variant, mk(level))
            assert a.k == ref.k and a.switches == ref.switches
            np.testing.assert_array_equal(a.stop_iter, ref.stop_iter)
            assert np.array_equal(a.x, ref.x) and np.array_equal(a.r, ref.r)
            assert not (a.report["status"] & 16)


def test_fused_alpha_value_table_csr_tiles():
    """A matrix with a value table whose tiles are not stored by diagonals
    (scattered columns, values from a set of five), with q formed at the
    gathered columns so alpha~ joins SpMV 2: bit for bit the eight-launch
    result."""
    A0 = P.zipf_matrix(20000, 12.0, 300, seed=9)
    rows = np.repeat(np.arange(A0.n), np.diff(A0.ptr.astype(np.int64)))
    vals = np.where(A0.col == rows, 200.0, np.array([-1.0, -0.5, 0.25, 0.5])[A0.col % 4])
    A = P.CSR(A0.ptr, A0.col, vals, A0.n)
    b = P.uniform(9, 2, A.n, -1.0, 1.0)
    sigma = P.uniform(9, 3, 16, 0.05, 0.5)
    seed = int(np.argmin(sigma))
    print("case fused alpha, value table, CSR tiles", flush=True)
    mk = lambda f: make_params(tol=1e-10, fused=f, window=16, q_mode=1)
    a = dropin_solve(A, b, sigma, seed, "dynamic", mk(1))
    c = dropin_solve(A, b, sigma, seed, "dynamic", mk(0))
    assert a.report["value_table"] == 5 and a.report["diag_tiles"] == 0 and a.report["q_unfused"] == 0
    assert a.k == c.k and a.switches == c.switches and a.report["stop_count"] == len(sigma)
    assert np.array_equal(a.x, c.x) and np.array_equal(a.r, c.r)
    assert not (a.report["status"] & 16)


def test_fused_launches_bitwise_gathered_q():
    """A scattered-column matrix whose SpMV 2 gathers q formed beforehand
    (an extra launch before SpMV 2): bit for bit the unfused result, window
    flushes included."""
    A = P.zipf_matrix(20000, 12.0, 300, seed=7)
    b = P.uniform(7, 2, A.n, -1.0, 1.0)
    sigma = P.uniform(7, 3, 24, 0.05, 0.5)
    seed = int(np.argmin(sigma))
    print("case fused vs unfused, gathered q", flush=True)
    mk = lambda f: make_params(tol=1e-10, fused=f, window=16, q_mode=2)
    a = dropin_solve(A, b, sigma, seed, "dynamic", mk(True))
    c = dropin_solve(A, b, sigma, seed, "dynamic", mk(False))
    assert a.report["q_unfused"] == 1
    assert a.k == c.k and a.switches == c.switches and a.report["stop_count"] == len(sigma)
    assert np.array_equal(a.x, c.x) and np.array_equal(a.r, c.r)
    assert not (a.report["status"] & 16)


def test_many_shifts_take_the_five_launch_form_bitwise():
    """1024 shifts on 32^3: the shift kernel with the scalar stage would leave
    too few resident tiles next to its 1024 coefficient workgroups, so the
    default (fused = 1) runs the stage in a launch of its own; results
    bitwise equal to fused = 3 and within the reference drivers'
    conventions (b = 1, sigma_i = (i+1) 0.01 / M, seed M/2 - 1)."""
    import torch
    from hbicg import Plan
    A = P.conv_diff_3d(32, 32, 32, 6.0, 0.2)
    M = 1024
    sigma = (np.arange(M) + 1) * (0.01 / M)
    b = torch.ones(A.n, dtype=torch.float64, device="cuda")
    print("case many shifts fused 1 vs 3, 32^3 M=1024", flush=True)
    out = []
    for fused in (1, 3):
        pl = Plan(A, M, make_params(tol=1e-10, graph=1, fused=fused))
        k, x, _ = pl.solve(b, sigma, M // 2 - 1)
        rep, stop, sw = pl.report()
        out.append((k, x.cpu().numpy(), stop, sw, rep["stop_count"]))
        pl.close()
    (k1, x1, s1, w1, c1), (k3, x3, s3, w3, c3) = out
    assert k1 == k3 and w1 == w3 and c1 == c3 == M
    assert np.array_equal(s1, s3) and np.array_equal(x1, x3)
