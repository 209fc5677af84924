"""GPU checks of the hand-written adjoint (hbicg.autograd).

* the device transpose equals a stable host transpose bit for bit, in one
  step and as a pattern plus per-call value gather;
* the gradient assembly equals host sums in the kernels' order bit for bit
  (grad_val, grad_b) and to 1e-12 (grad_sigma, tree-summed);
* gradients w.r.t. b, every sigma_k and the nonzeros of A match central
  finite differences (h = 1e-6) to 1e-6 relative, for the collinear loss
  L = sum_k w_k c^T x_k (one shifted adjoint solve) and for a general loss
  (per-shift adjoint solves through the batched solver).  Forward and
  adjoint solves run to 1e-15 so that solver error / h stays far below the
  1e-6 bar;
* the adjoint solutions lam_k equal per-shift transposed solves of the
  oracle (the CPU restatement of the reference solver) to 1e-8 relative.
"""
import numpy as np
import pytest
import torch

import pyoracle as O
from hbicg import problems as P
from hbicg.autograd import shifted_solve, transpose

pytestmark = pytest.mark.gpu

TOL = 1e-15
H = 1e-6


@pytest.fixture(scope="module")
def prob():
    A = P.conv_diff_3d(12, 12, 12, 6.0, 0.2)
    m = 6
    sigma = P.logspace_shifts(m, 1e-2, 1.0)
    b = P.uniform(4, 1, A.n, -1.0, 1.0)
    w = P.uniform(4, 2, m, 0.0, 1.0)
    c = P.uniform(4, 3, A.n, 0.0, 1.0)
    return A, b, sigma, w, c


def dev_csr(A):
    val = torch.as_tensor(A.val).cuda()
    col = torch.as_tensor(A.col.view(np.int32)).cuda()
    ptr = torch.as_tensor(A.ptr.view(np.int32)).cuda()
    return val, col, ptr


def loss_collinear(val, col, ptr, b, sigma, w, c):
    x = shifted_solve(val, ptr, col, b, sigma, seed=0, tol=TOL)
    return (w[:, None] * (x @ c)[:, None]).sum(), x


def test_transpose_matches_host(prob):
    A = prob[0]
    for Am in (A, P.zipf_matrix(3000, 12.0, 200, seed=5)):
        val, col, ptr = dev_csr(Am)
        tval, tcol, tptr = transpose(val, col, ptr)
        T = Am.transpose()
        assert np.array_equal(tptr.cpu().numpy().view(np.uint32), T.ptr)
        assert np.array_equal(tcol.cpu().numpy().view(np.uint32), T.col)
        assert np.array_equal(tval.cpu().numpy(), T.val)


def test_gradients_vs_finite_differences(prob):
    A, b0, sig0, w0, c0 = prob
    print("case adjoint FD collinear 12^3 m=6", flush=True)
    val0, col, ptr = dev_csr(A)
    w = torch.as_tensor(w0).cuda()
    c = torch.as_tensor(c0).cuda()
    val = val0.clone().requires_grad_(True)
    b = torch.as_tensor(b0).cuda().requires_grad_(True)
    sigma = torch.as_tensor(sig0).clone().requires_grad_(True)
    L, _ = loss_collinear(val, col, ptr, b, sigma, w, c)
    L.backward()
    rng = np.random.default_rng(11)

    def Lval(v, bb, ss):
        with torch.no_grad():
            return float(loss_collinear(v, col, ptr, bb, ss, w, c)[0])

    # b along a random direction
    d = torch.as_tensor(rng.uniform(-1, 1, A.n)).cuda()
    fd = (Lval(val0, b.detach() + H * d, sigma.detach()) - Lval(val0, b.detach() - H * d, sigma.detach())) / (2 * H)
    an = float(b.grad @ d)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    # every sigma_k
    for k in range(len(sig0)):
        e = torch.zeros_like(sigma)
        e[k] = H
        fd = (Lval(val0, b.detach(), sigma.detach() + e) - Lval(val0, b.detach(), sigma.detach() - e)) / (2 * H)
        an = float(sigma.grad[k])
        assert abs(fd - an) <= 1e-6 * abs(an), (k, fd, an)
    # nonzeros of A along a random direction on the pattern
    D = torch.as_tensor(rng.uniform(-1, 1, A.nnz)).cuda()
    fd = (Lval(val0 + H * D, b.detach(), sigma.detach()) - Lval(val0 - H * D, b.detach(), sigma.detach())) / (2 * H)
    an = float(val.grad @ D)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_general_loss_uses_per_shift_adjoints(prob):
    A, b0, sig0, _, _ = prob
    print("case adjoint FD general loss", flush=True)
    val0, col, ptr = dev_csr(A)
    m = len(sig0)
    C_ = torch.as_tensor(P.uniform(6, 1, m * A.n, -1.0, 1.0).reshape(m, A.n)).cuda()
    b = torch.as_tensor(b0).cuda().requires_grad_(True)
    sigma = torch.as_tensor(sig0).clone().requires_grad_(True)

    def L(bb, ss):
        x = shifted_solve(val0, ptr, col, bb, ss, seed=0, tol=TOL)
        return (x * C_).sum()

    out = L(b, sigma)
    out.backward()
    d = torch.as_tensor(np.random.default_rng(12).uniform(-1, 1, A.n)).cuda()
    with torch.no_grad():
        fd = (float(L(b.detach() + H * d, sigma.detach())) - float(L(b.detach() - H * d, sigma.detach()))) / (2 * H)
    an = float(b.grad @ d)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    with torch.no_grad():
        e = torch.zeros_like(sigma)
        e[2] = H
        fd = (float(L(b.detach(), sigma.detach() + e)) - float(L(b.detach(), sigma.detach() - e))) / (2 * H)
    assert abs(fd - float(sigma.grad[2])) <= 1e-6 * abs(float(sigma.grad[2]))


def test_adjoint_solutions_match_oracle_transposed_solves(prob):
    """grad_b = sum_k w_k mu_k with mu_k = (A^T + sigma_k)^{-1} c: compare
    against per-shift transposed solves of the oracle."""
    A, b0, sig0, w0, c0 = prob
    val0, col, ptr = dev_csr(A)
    b = torch.as_tensor(b0).cuda().requires_grad_(True)
    L, _ = loss_collinear(val0, col, ptr, b, torch.as_tensor(sig0), torch.as_tensor(w0).cuda(),
                          torch.as_tensor(c0).cuda())
    L.backward()
    T = A.transpose()
    ref = np.zeros(A.n)
    for k in range(len(sig0)):
        r = O.solve(T, c0, sig0[k:k + 1], 0, tol=1e-14)
        ref += w0[k] * r.x[0]
    got = b.grad.cpu().numpy()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) <= 1e-8


def test_values_changed_in_place_between_calls(prob):
    """The differentiable solve keeps its plans across calls and reads A's
    values in place: after an in-place update of val (as an optimiser
    makes), the forward solution and the gradient w.r.t. b (one adjoint
    solve on A^T) equal those of a fresh copy of the updated values."""
    A, b, sigma, w, c = prob
    val, col, ptr = dev_csr(A)
    bt = torch.as_tensor(b).cuda()
    st = torch.as_tensor(sigma)
    wt, ct = torch.as_tensor(w).cuda(), torch.as_tensor(c).cuda()

    def run(v):
        bb = bt.clone().requires_grad_(True)
        L, x = loss_collinear(v, col, ptr, bb, st, wt, ct)
        L.backward()
        return x.detach().cpu().numpy(), bb.grad.cpu().numpy()

    run(val)                                  # plans cached for val's memory
    with torch.no_grad():
        val.mul_(1.25)                        # same memory, new values
    x1, g1 = run(val)
    fresh = val.clone()                       # other memory: new plans
    x2, g2 = run(fresh)
    assert np.array_equal(x1, x2)
    assert np.array_equal(g1, g2)


def test_loss_on_two_shifts_only(prob):
    """A loss that reads only x_1 and x_4 leaves the other rows of dL/dx
    zero.  Those adjoint solves are skipped (a zero right-hand side would
    make BiCG divide 0 by 0), so every gradient is finite and matches
    central differences."""
    A, b0, sig0, _, _ = prob
    print("case adjoint loss on 2 of 6 shifts", flush=True)
    val0, col, ptr = dev_csr(A)
    t1 = torch.as_tensor(P.uniform(7, 1, A.n, -1.0, 1.0)).cuda()
    t4 = torch.as_tensor(P.uniform(7, 2, A.n, -1.0, 1.0)).cuda()
    val = val0.clone().requires_grad_(True)
    b = torch.as_tensor(b0).cuda().requires_grad_(True)
    sigma = torch.as_tensor(sig0).clone().requires_grad_(True)

    def L(v, bb, ss):
        x = shifted_solve(v, ptr, col, bb, ss, seed=0, tol=TOL)
        return ((x[1] - t1) ** 2).sum() + ((x[4] - t4) ** 2).sum()

    L(val, b, sigma).backward()
    for g in (val.grad, b.grad, sigma.grad):
        assert torch.isfinite(g).all()
    assert float(sigma.grad[0]) == 0.0 and float(sigma.grad[5]) == 0.0
    rng = np.random.default_rng(13)
    d = torch.as_tensor(rng.uniform(-1, 1, A.n)).cuda()
    with torch.no_grad():
        fd = (float(L(val0, b.detach() + H * d, sigma.detach())) -
              float(L(val0, b.detach() - H * d, sigma.detach()))) / (2 * H)
    an = float(b.grad @ d)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    for k in (1, 4):
        e = torch.zeros_like(sigma)
        e[k] = H
        with torch.no_grad():
            fd = (float(L(val0, b.detach(), sigma.detach() + e)) -
                  float(L(val0, b.detach(), sigma.detach() - e))) / (2 * H)
        assert abs(fd - float(sigma.grad[k])) <= 1e-6 * abs(float(sigma.grad[k])), (k, fd)
    D = torch.as_tensor(rng.uniform(-1, 1, A.nnz)).cuda()
    with torch.no_grad():
        fd = (float(L(val0 + H * D, b.detach(), sigma.detach())) -
              float(L(val0 - H * D, b.detach(), sigma.detach()))) / (2 * H)
    an = float(val.grad @ D)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_zero_rhs_and_zero_gradient(prob):
    """b = 0: x = 0 without a solve (the solver would carry 0/0 to its cap).
    A loss with an all-zero gradient gives zero gradients."""
    A, _, sig0, _, _ = prob
    val0, col, ptr = dev_csr(A)
    val = val0.clone().requires_grad_(True)
    b = torch.zeros(A.n, dtype=torch.float64, device="cuda", requires_grad=True)
    sigma = torch.as_tensor(sig0).clone().requires_grad_(True)
    x = shifted_solve(val, ptr, col, b, sigma, seed=0, tol=1e-10)
    assert x.shape == (len(sig0), A.n) and not torch.any(x != 0)
    (x.sum() * 0.0).backward()
    for g in (val.grad, b.grad, sigma.grad):
        assert g is not None and not torch.any(g != 0)


def test_transpose_with_duplicate_entries_is_stable():
    """Duplicate (i, j) entries (Matrix Market files may hold them; coo2csr
    keeps them) land in A^T in their order in A, whatever order the
    scatter's atomics place them in: two transposes agree bit for bit with
    each other and with a stable host transpose."""
    rng = np.random.default_rng(21)
    n = 2000
    lens = rng.integers(1, 40, n)
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, 64, rows.size)             # few columns: many duplicates
    order = np.lexsort((cols, rows))                  # row-major, columns ascending
    rows, cols = rows[order], cols[order]
    ptr = np.zeros(n + 1, dtype=np.uint32)
    ptr[1:] = np.cumsum(lens)
    A = P.CSR(ptr, cols.astype(np.uint32), rng.uniform(-1, 1, rows.size), n)
    assert np.any((np.diff(cols) == 0) & (np.diff(rows) == 0))
    val, col, ptr_d = dev_csr(A)
    T = A.transpose()
    outs = [transpose(val, col, ptr_d) for _ in range(3)]
    for tval, tcol, tptr in outs:
        assert np.array_equal(tptr.cpu().numpy().view(np.uint32), T.ptr)
        assert np.array_equal(tcol.cpu().numpy().view(np.uint32), T.col)
        assert np.array_equal(tval.cpu().numpy(), T.val)


def _ragged_csr(seed, n, max_len, ncols=None):
    """CSR with empty rows, rows of up to max_len entries and duplicates."""
    rng = np.random.default_rng(seed)
    lens = rng.integers(0, max_len + 1, n)
    lens[rng.integers(0, n, n // 10)] = 0
    rows = np.repeat(np.arange(n), lens)
    cols = rng.integers(0, ncols or n, rows.size)
    order = np.lexsort((cols, rows))
    rows, cols = rows[order], cols[order]
    ptr = np.zeros(n + 1, dtype=np.uint32)
    ptr[1:] = np.cumsum(lens)
    return P.CSR(ptr, cols.astype(np.uint32), rng.uniform(-1, 1, rows.size), n)


def test_transpose_pattern_then_gather_equals_transpose():
    """The two-step transpose (pattern and source indices once, values
    gathered per call) gives the one-step transpose's arrays bit for bit,
    also after the values change in place."""
    import ctypes as C
    from hbicg import _lib
    from hbicg.autograd import _csr
    L = _lib.load()
    for Am in (P.conv_diff_3d(12, 12, 12, 6.0, 0.2), P.zipf_matrix(3000, 12.0, 200, seed=5),
               _ragged_csr(8, 1500, 30, ncols=64)):
        print("case transpose pattern n", Am.n, "nnz", Am.nnz, flush=True)
        val, col, ptr = dev_csr(Am)
        n = Am.n
        tptr = torch.empty(n + 1, dtype=torch.int32, device="cuda")
        tcol = torch.empty_like(col)
        tsrc = torch.empty_like(col)
        tval = torch.empty_like(val)
        s = torch.cuda.current_stream()
        rc = L.hbicg_csr_transpose_pattern(C.byref(_csr(val, col, ptr)), C.c_void_p(tptr.data_ptr()),
                                           C.c_void_p(tcol.data_ptr()), C.c_void_p(tsrc.data_ptr()),
                                           C.c_void_p(s.cuda_stream))
        assert rc == 0
        src = tsrc.cpu().numpy().view(np.uint32)
        assert np.array_equal(np.sort(src), np.arange(Am.nnz, dtype=np.uint32))
        for step in range(2):
            rc = L.hbicg_gather_values(C.c_void_p(val.data_ptr()), C.c_void_p(tsrc.data_ptr()), Am.nnz,
                                       C.c_void_p(tval.data_ptr()), C.c_void_p(s.cuda_stream))
            assert rc == 0
            rval, rcol, rptr = transpose(val, col, ptr)
            assert torch.equal(rptr, tptr) and torch.equal(rcol, tcol) and torch.equal(rval, tval)
            T = P.CSR(Am.ptr, Am.col, val.cpu().numpy(), n).transpose()
            assert np.array_equal(tval.cpu().numpy(), T.val)
            val.mul_(-1.5).add_(0.25)


def test_gradient_assembly_against_host_sums():
    """grad_val and grad_b bit for bit against host sums in the kernels'
    order (shifts ascending, products rounded on their own), grad_sigma to
    1e-12; the workspace entry point equals the allocating one bit for bit.
    Ragged rows (empty ones included), n not a multiple of 256, and M = 20
    (a partial last round of the kernel's 16 shifts), with and without w."""
    import ctypes as C
    from hbicg import _lib
    from hbicg.autograd import _csr, gradients
    L = _lib.load()
    A = _ragged_csr(9, 3001, 25)
    M = 20
    rng = np.random.default_rng(10)
    mu = rng.standard_normal((M, A.n))
    x = rng.standard_normal((M, A.n))
    wv = rng.uniform(0.5, 2.0, M)
    val, col, ptr = dev_csr(A)
    mu_d, x_d = torch.as_tensor(mu).cuda(), torch.as_tensor(x).cuda()
    rows = np.repeat(np.arange(A.n), np.diff(A.ptr.astype(np.int64)))
    cols = A.col.astype(np.int64)
    for w in (None, wv):
        print("case grads ragged n", A.n, "M", M, "w", w is not None, flush=True)
        lam = mu if w is None else w[:, None] * mu
        gv_ref = np.zeros(A.nnz)
        gb_ref = np.zeros(A.n)
        for k in range(M):
            gv_ref = gv_ref - lam[k][rows] * x[k][cols]
            gb_ref = gb_ref + lam[k]
        gs_ref = -np.einsum("kn,kn->k", lam, x)
        w_d = None if w is None else torch.as_tensor(w).cuda()
        gb, gs, gv = gradients(val, col, ptr, mu_d, w_d, x_d)
        torch.cuda.synchronize()
        assert np.array_equal(gv.cpu().numpy(), gv_ref)
        assert np.array_equal(gb.cpu().numpy(), gb_ref)
        # tree-summed on the GPU: bounded by 1e-12 of the sum of |terms|
        assert np.all(np.abs(gs.cpu().numpy() - gs_ref) <= 1e-12 * np.abs(lam * x).sum(axis=1))
        ob, os_, ov = torch.empty_like(gb), torch.empty_like(gs), torch.empty_like(gv)
        p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
        rc = L.hbicg_adjoint_gradients(C.byref(_csr(val, col, ptr)), M, p(mu_d), p(w_d), p(x_d), p(ob),
                                       p(os_), p(ov), C.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert rc == 0
        assert torch.equal(ob, gb) and torch.equal(os_, gs) and torch.equal(ov, gv)
