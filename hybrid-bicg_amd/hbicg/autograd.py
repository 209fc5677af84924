"""torch.autograd binding of the shifted solver with a hand-written adjoint.

    x = shifted_solve(val, ptr, col, b, sigma, seed)      # x: [M, n]

solves (A + sigma_k I) x_k = b for every shift on the GPU (val holds A's
nonzeros in CSR order; ptr/col are int32 tensors holding the unsigned
32-bit CSR indices).  Backward, for g_k = dL/dx_k:

    lam_k = (A + sigma_k I)^{-T} g_k
    dL/db = sum_k lam_k,  dL/dsigma_k = -lam_k . x_k,
    dL/dA_ij = -sum_k lam_k[i] x_k[j] on A's pattern.

A^T is built on the device (hbicg_csr_transpose).  When the g_k are
collinear, g_k = w_k c (a loss sum_k w_k c^T x_k), one shifted solve on A^T
with right-hand side c gives every mu_k and lam_k = w_k mu_k.  Otherwise
each shift with a nonzero g_k is its own single-shift problem on A^T, all
solved together by the batched solver.  Gradients are assembled by
hbicg_adjoint_gradients.

A zero right-hand side is never handed to the solver: BiCG's first step
divides (r^, s) = 0 by itself, and the solver, like the reference, carries
the NaN to the iteration cap.  Its solution is 0, so a zero b gives x = 0
in the forward pass, a zero g_k gives lam_k = 0, and an all-zero gradient
gives zero gradients.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .solver import Batch, Plan, make_params


def _csr(val, col, ptr):
    n = int(ptr.numel()) - 1
    return _lib.Csr(val.data_ptr(), col.data_ptr(), ptr.data_ptr(), n, int(val.numel()))


def transpose(val, col, ptr):
    """A^T of a device CSR: (tval, tcol, tptr), rows in ascending column order."""
    L = _lib.load()
    n = int(ptr.numel()) - 1
    tptr = torch.empty(n + 1, dtype=torch.int32, device=val.device)
    tcol = torch.empty_like(col)
    tval = torch.empty_like(val)
    s = torch.cuda.current_stream(val.device)
    rc = L.hbicg_csr_transpose(C.byref(_csr(val, col, ptr)), C.c_void_p(tptr.data_ptr()),
                               C.c_void_p(tcol.data_ptr()), C.c_void_p(tval.data_ptr()),
                               C.c_void_p(s.cuda_stream))
    _lib.check(rc, "hbicg_csr_transpose")
    return tval, tcol, tptr


def gradients(val, col, ptr, mu, w, x, need_b=True, need_sigma=True, need_val=True):
    """(grad_b, grad_sigma, grad_val) from adjoint solutions lam_k = w_k mu_k,
    enqueued on the current stream (no allocation in the library, no sync)."""
    L = _lib.load()
    n = int(ptr.numel()) - 1
    M = int(x.shape[0])
    dev = x.device
    gb = torch.empty(n, dtype=torch.float64, device=dev) if need_b else None
    gs = torch.empty(M, dtype=torch.float64, device=dev) if need_sigma else None
    gv = torch.empty_like(val) if need_val else None
    nbytes = int(L.hbicg_adjoint_work_bytes(n, M)) if need_sigma else 0
    work = torch.empty(max(nbytes // 8, 1), dtype=torch.float64, device=dev) if need_sigma else None
    p = lambda t: C.c_void_p(t.data_ptr()) if t is not None else None
    s = torch.cuda.current_stream(dev)
    rc = L.hbicg_adjoint_gradients_ws(C.byref(_csr(val, col, ptr)), M, p(mu), p(w), p(x), p(gb), p(gs),
                                      p(gv), p(work), nbytes, C.c_void_p(s.cuda_stream))
    _lib.check(rc, "hbicg_adjoint_gradients_ws")
    return gb, gs, gv


def collinear_split(g, rtol=1e-13, norms=None):
    """(w, c) with g_k = w_k c when the rows of g are collinear, else None.
    g must have a nonzero row (the caller handles an all-zero gradient).
    norms: the rows' 2-norms, when the caller has them.  One host sync."""
    if norms is None:
        norms = torch.linalg.vector_norm(g, dim=1)
    k0 = torch.argmax(norms)
    c = g.index_select(0, k0.view(1))[0]
    cc = torch.dot(c, c)
    w = (g @ c) / cc
    resid = torch.linalg.vector_norm(g - w[:, None] * c)
    chk = torch.stack([cc, resid, torch.linalg.vector_norm(g)]).cpu()
    if float(chk[0]) == 0.0:
        raise ValueError("collinear_split: every row of g is zero")
    if float(chk[1]) <= rtol * float(chk[2]):
        return w.contiguous(), c.contiguous()
    return None


class _PlanCache:
    """Plans kept across calls, keyed by the device arrays they read (the
    plan reads val in place, so updated values are seen) and the solve
    parameters; a few entries, least recently used dropped first."""

    def __init__(self, size=4):
        self.size = size
        self.items = {}

    def get(self, key, make):
        if key in self.items:
            plan = self.items.pop(key)
        else:
            plan = make()
            while len(self.items) >= self.size:
                old = self.items.pop(next(iter(self.items)))
                old.close()
        self.items[key] = plan
        return plan


_plans = _PlanCache()
_transposes = {}


def _plan_for(val, col, ptr, M, tol, max_iter):
    """A cached plan that reads val in place (value_index=0: a value table
    would hold the values of the plan's creation, and val changes in place
    between calls, e.g. under an optimiser; so does the transposed copy)."""
    key = (val.data_ptr(), col.data_ptr(), ptr.data_ptr(), int(val.numel()), M, tol, max_iter)
    params = make_params(tol=tol, max_iter=max_iter, graph=1, value_index=False)
    return _plans.get(key, lambda: Plan(None, M, params, device=val.device, tensors=(val, col, ptr)))


def _transpose_into(val, col, ptr):
    """A^T in buffers kept per (val, col, ptr), so that the adjoint plan on
    them can be reused.  The pattern of A^T and the source entry of each of
    its entries are built once (the pattern of a cached (col, ptr) is taken
    as fixed, as the cached plans take it); the values are gathered from val
    on every call."""
    key = (val.data_ptr(), col.data_ptr(), ptr.data_ptr(), int(val.numel()))
    L = _lib.load()
    s = torch.cuda.current_stream(val.device)
    bufs = _transposes.get(key)
    if bufs is None:
        if len(_transposes) >= 4:
            _transposes.pop(next(iter(_transposes)))
        n = int(ptr.numel()) - 1
        bufs = (torch.empty_like(val), torch.empty_like(col),
                torch.empty(n + 1, dtype=torch.int32, device=val.device), torch.empty_like(col))
        tval, tcol, tptr, tsrc = bufs
        rc = L.hbicg_csr_transpose_pattern(C.byref(_csr(val, col, ptr)), C.c_void_p(tptr.data_ptr()),
                                           C.c_void_p(tcol.data_ptr()), C.c_void_p(tsrc.data_ptr()),
                                           C.c_void_p(s.cuda_stream))
        _lib.check(rc, "hbicg_csr_transpose_pattern")
        _transposes[key] = bufs
    tval, tcol, tptr, tsrc = bufs
    rc = L.hbicg_gather_values(C.c_void_p(val.data_ptr()), C.c_void_p(tsrc.data_ptr()), int(val.numel()),
                               C.c_void_p(tval.data_ptr()), C.c_void_p(s.cuda_stream))
    _lib.check(rc, "hbicg_gather_values")
    return tval, tcol, tptr


class ShiftedSolve(torch.autograd.Function):
    """x = (A + sigma_k I)^{-1} b for every shift, differentiable in val, b
    and sigma (see the module docstring)."""

    @staticmethod
    def forward(ctx, val, b, sigma, ptr, col, seed, tol, max_iter, adj_tol):
        M = int(sigma.numel())
        sig = sigma.detach().double().cpu().numpy()
        vd = val.detach()
        if not vd.is_contiguous():
            vd = vd.contiguous()
        bd = b.detach().contiguous()
        if bool(torch.any(bd != 0)):
            plan = _plan_for(vd, col, ptr, M, tol, max_iter)
            k, x, rep = plan.solve(bd, sig, int(seed))
        else:
            k, x = 1, torch.zeros((M, bd.numel()), dtype=torch.float64, device=bd.device)
        ctx.save_for_backward(vd, b, x)
        ctx.ptr, ctx.col, ctx.sig, ctx.seed = ptr, col, sig, int(seed)
        ctx.adj_tol, ctx.max_iter = adj_tol, max_iter
        ctx.sigma_device, ctx.sigma_dtype = sigma.device, sigma.dtype
        ctx.k = k
        return x

    @staticmethod
    def backward(ctx, gx):
        vd, b, x = ctx.saved_tensors
        ptr, col, sig = ctx.ptr, ctx.col, ctx.sig
        M = x.shape[0]
        gx = gx.contiguous()
        norms = torch.linalg.vector_norm(gx, dim=1)
        live = norms.cpu().numpy() > 0
        if not live.all():
            # a row whose squares all underflow has norm 0 but is not zero
            live = torch.any(gx != 0, dim=1).cpu().numpy()
        if not live.any():
            zero = lambda t, need: torch.zeros_like(t) if need else None
            gs = zero(torch.empty(M, dtype=ctx.sigma_dtype, device=ctx.sigma_device), ctx.needs_input_grad[2])
            return (zero(vd, ctx.needs_input_grad[0]), zero(b, ctx.needs_input_grad[1]), gs,
                    None, None, None, None, None, None)
        tval, tcol, tptr = _transpose_into(vd, col, ptr)
        params = make_params(tol=ctx.adj_tol, max_iter=ctx.max_iter, graph=1, value_index=False)
        split = collinear_split(gx, norms=norms)
        if split is not None:
            w, c = split
            plan = _plan_for(tval, tcol, tptr, M, ctx.adj_tol, ctx.max_iter)
            _, mu, _ = plan.solve(c, sig, ctx.seed)
        else:
            # one single-shift problem per shift whose g_k is nonzero;
            # lam_k = 0 for the others
            w = None
            rows = [k for k in range(M) if live[k]]
            spec = [{"val": tval, "col": tcol, "ptr": tptr, "b": gx[k], "sigma": sig[k:k + 1],
                     "seed": 0} for k in rows]
            bt = Batch(None, params, device=x.device, device_spec=spec)
            _, lams, _ = bt.solve()
            bt.close()
            mu = torch.zeros_like(x)
            for k, l_ in zip(rows, lams):
                mu[k] = l_[0]
        gb, gs, gv = gradients(vd, col, ptr, mu, w, x, ctx.needs_input_grad[1],
                               ctx.needs_input_grad[2], ctx.needs_input_grad[0])
        if gs is not None:
            gs = gs.to(device=ctx.sigma_device, dtype=ctx.sigma_dtype)
        return gv, gb, gs, None, None, None, None, None, None


def shifted_solve(val, ptr, col, b, sigma, seed=0, tol=1e-12, max_iter=2000, adj_tol=None):
    """Differentiable shifted solve; returns x [M, n] (float64, on val's device).
    sigma may live on the host or the device; its gradient has sigma's device."""
    adj = tol if adj_tol is None else adj_tol
    x = ShiftedSolve.apply(val, b, sigma, ptr, col, seed, tol, max_iter, adj)
    return x
