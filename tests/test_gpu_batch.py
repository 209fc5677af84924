"""GPU parity of the batched solver (hbicg_batch_*).

Each problem of a batch runs the same tiles and the same partial-sum order
as the plain plan, so the batched results must equal per-problem plain
solves bit for bit; every shift's true residual is recomputed, and in
exact_order mode a batch must equal the oracle itself."""
import numpy as np
import pytest
import torch

import pyoracle as O
from hbicg import Batch, Plan, make_params
from hbicg import problems as P

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    # uneven sizes (2-D and 3-D), 8..64 shifts, per-problem convection
    probs = P.config3(count=10, seed=3, n_lo=300, n_hi=30000)
    for p in probs[::3]:
        p.seed = p.M // 2           # middle seeds: exercises seed switching
    return probs


def test_batch_equals_plain_bitwise(small_batch):
    print("case batch vs plain, 10 problems", flush=True)
    params = make_params(tol=1e-10, graph=1)
    bt = Batch(small_batch, params)
    k, xs, ms = bt.solve()
    for i, p in enumerate(small_batch):
        # the plain plan at the batch's window (the largest its rings hold)
        rep, stop, sw = bt.report(i)
        pl = Plan(p.A, p.M, make_params(tol=1e-10, graph=1, window=rep["window"]))
        kp, xp, _ = pl.solve(torch.as_tensor(p.b).cuda(), p.sigma, p.seed)
        prep, pstop, psw = pl.report()
        assert k[i] == kp, (i, p.name)
        assert np.array_equal(stop, pstop)
        assert sw == psw
        assert torch.equal(xs[i], xp), (i, p.name)
        tr = bt.true_residuals(i, xs[i])
        assert tr.max() <= 10 * p.tol, (i, p.name, tr.max())
        pl.close()
    bt.close()


def test_batch_exact_equals_oracle(small_batch):
    probs = small_batch[:4]
    print("case batch exact vs oracle, 4 problems", flush=True)
    bt = Batch(probs, make_params(tol=1e-10, graph=1, exact_order=True, window=8))
    k, xs, _ = bt.solve()
    for i, p in enumerate(probs):
        ref = O.solve(p.A, p.b, p.sigma, p.seed, tol=1e-10)
        rep, stop, sw = bt.report(i)
        assert k[i] == ref.k
        assert np.array_equal(stop, ref.stop_iter)
        assert sw == ref.switches
        assert np.array_equal(xs[i].cpu().numpy(), ref.x)
    bt.close()


def _rel(a, b):
    return float(((a - b).norm(dim=1) / b.norm(dim=1).clamp_min(1e-300)).max())


def test_autotune_keeps_results(small_batch):
    """Tuning picks a window (default mode: 32/64/128, combined-coefficient
    flush; trials over whole solves) and keeps the result.  Without a seed switch, iteration counts and
    stop iterations come from the seed and the scalar recurrences, which no
    window changes, and x_j move only by roundings of the regrouped window
    sums; after a switch the new seed continues from its windowed p_j, so
    the rest of that run moves by roundings as well."""
    print("case batch autotune", flush=True)
    bt = Batch(small_batch, make_params(tol=1e-10, graph=1))
    k0, x0, _ = bt.solve()
    rep0 = [bt.report(i) for i in range(len(small_batch))]
    best = bt.autotune()
    assert best["window"] in (32, 64, 128) and best["group_cap"] == 0
    assert best["ms_per_iter"] > 0
    k1, x1, _ = bt.solve()
    for i, (a, b) in enumerate(zip(x0, x1)):
        _, stop0, sw0 = rep0[i]
        _, stop1, sw1 = bt.report(i)
        if not sw0:
            assert k0[i] == k1[i]
            assert np.array_equal(stop0, stop1)
            assert _rel(a, b) <= 1e-10, (i, _rel(a, b))
        else:
            assert abs(int(k0[i]) - int(k1[i])) <= 2
            assert np.all(np.abs(stop0 - stop1) <= 2)
            assert _rel(a, b) <= 1e-9, (i, _rel(a, b))
    bt.close()


def test_autotune_exact_keeps_bits(small_batch):
    """Exact order: the window replay with any (window, shifts per row)
    layout is bit-identical."""
    probs = small_batch[:4]
    print("case batch autotune exact", flush=True)
    bt = Batch(probs, make_params(tol=1e-10, graph=1, exact_order=True))
    k0, x0, _ = bt.solve()
    best = bt.autotune(iters=16)
    assert best["window"] in (4, 8, 16) and best["group_cap"] in (0, 4, 16)
    k1, x1, _ = bt.solve()
    assert np.array_equal(k0, k1)
    for a, b in zip(x0, x1):
        assert torch.equal(a, b)
    bt.close()


def test_plan_device_path_matches_dropin():
    """The device-pointer plan and the host drop-in compute the same x."""
    from hbicg import dropin_solve
    p = P.config0()
    got = dropin_solve(p.A, p.b, p.sigma, 8, "static", make_params(tol=p.tol, window=16))
    pl = Plan(p.A, p.M, make_params(tol=p.tol, window=16, graph=1))
    k, x, _ = pl.solve(torch.as_tensor(p.b).cuda(), p.sigma, 8)
    assert k == got.k
    assert np.array_equal(x.cpu().numpy(), got.x)
    pl.close()


def test_batch_reads_index_bytes_where_plans_read_constant_diagonals():
    """Batched plans store their diagonal tiles with index bytes, where a
    plain plan of a constant-coefficient stencil stores presence bits and
    the diagonals' values: two such stencils and one matrix whose main
    diagonal varies in one tile, each bitwise equal to its plain plan, and
    the batch report says which SpMV form the batch read."""
    from test_gpu_solver import _mixed_tile_matrix
    rng = np.random.default_rng(31)
    mats = [P.conv_diff_2d(48, 40), P.conv_diff_3d(14, 12, 16, 6.0, 0.2), _mixed_tile_matrix(vary=True)]
    probs = [P.ShiftedProblem(f"forms-{i}", A, rng.uniform(-1.0, 1.0, A.n), np.logspace(-2, 0, 6 + i), 0, 1e-10)
             for i, A in enumerate(mats)]
    print("case batch diagonal forms, 3 problems", flush=True)
    bt = Batch(probs, make_params(tol=1e-10, graph=1))
    k, xs, _ = bt.solve()
    for i, p in enumerate(probs):
        rep, stop, sw = bt.report(i)
        assert rep["value_table"] > 0 and rep["diag_tiles"] > 0 and rep["const_tiles"] == 0, (i, rep)
        pl = Plan(p.A, p.M, make_params(tol=1e-10, graph=1, window=rep["window"]))
        kp, xp, _ = pl.solve(torch.as_tensor(p.b).cuda(), p.sigma, p.seed)
        prep, pstop, psw = pl.report()
        if i == 2:
            assert prep["const_tiles"] == 0
        else:
            assert prep["const_tiles"] == prep["diag_tiles"] > 0
        assert k[i] == kp and np.array_equal(stop, pstop) and sw == psw, (i, k[i], kp)
        assert torch.equal(xs[i], xp), i
        pl.close()
    bt.close()
