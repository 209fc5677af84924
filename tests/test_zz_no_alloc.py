"""hbicg_plan_true_residuals_noalloc through the C-ABI against
hbicg_plan_true_residuals: the same kernels with their partial sums in the
GPU's static scratch instead of an allocated buffer, so the residuals agree
bit for bit.  Two calls in a row share the scratch, the second on a smaller
input (fewer rows and shifts) -- it must not read anything the first one
left.  More shifts than the scratch holds are refused with HBICG_E_SCRATCH
before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from hbicg import Plan, _lib, make_params
from hbicg import problems as P

pytestmark = pytest.mark.gpu


def _case(A, b, sigma, seed):
    M = len(sigma)
    rng = np.random.default_rng(seed)
    x = torch.as_tensor(rng.uniform(-1, 1, (M, A.n))).cuda()
    pl = Plan(A, M, make_params(tol=1e-10))
    bt = torch.as_tensor(b).cuda()
    ref = pl.true_residuals(bt, sigma, x)
    out = torch.full((M,), -1.0, dtype=torch.float64, device="cuda")
    got = pl.true_residuals_noalloc(bt, torch.as_tensor(sigma).cuda(), x, out)
    torch.cuda.synchronize()
    pl.close()
    return ref, got.cpu().numpy()


def test_noalloc_equals_allocating_entry_point_twice():
    c0 = P.config0()
    print("case noalloc 1: 64x64, 16 shifts", flush=True)
    ref, got = _case(c0.A, c0.b, c0.sigma, 1)
    assert np.array_equal(got, ref), (got, ref)
    A = P.conv_diff_3d(10, 9, 8, 6.0, 0.2)
    b = P.uniform(5, 1, A.n, -1.0, 1.0)
    sigma = P.logspace_shifts(5, 1e-2, 1.0)
    print("case noalloc 2: 10x9x8, 5 shifts", flush=True)
    ref, got = _case(A, b, sigma, 2)
    assert np.array_equal(got, ref), (got, ref)


def test_noalloc_refuses_more_shifts_than_its_scratch():
    n, M = 4, 65537
    A = P.CSR(np.arange(n + 1, dtype=np.uint32), np.arange(n, dtype=np.uint32), np.full(n, 2.0), n)
    pl = Plan(A, M, make_params(tol=1e-10))
    d = lambda *s: torch.zeros(s, dtype=torch.float64, device="cuda")
    L = _lib.load()
    rc = L.hbicg_plan_true_residuals_noalloc(pl.h, C.c_void_p(d(n).data_ptr()), C.c_void_p(d(M).data_ptr()),
                                             C.c_void_p(d(M, n).data_ptr()), C.c_void_p(d(M).data_ptr()))
    pl.close()
    assert rc == -7
