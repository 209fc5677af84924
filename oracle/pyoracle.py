"""ctypes front end of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU
baseline, never as a product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
REF_BLAS_PATH = os.path.join(HERE, "_ref", "libref_blas.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)

NO_SWITCH_TARGET = 4   # OracleResult.status bit (hbicg_oracle.h ORC_NO_SWITCH_TARGET)


class _Csr(C.Structure):
    _fields_ = [("val", _dp), ("col", _up), ("ptr", _up), ("n", C.c_uint), ("nz", C.c_uint)]


class _Trace(C.Structure):
    _fields_ = [("stop_iter", _ip), ("switch_k", _ip), ("switch_to", _ip), ("n_switch", C.c_int),
                ("alpha", _dp), ("beta", _dp), ("omega", _dp), ("res_hist", _dp),
                ("seed_hist", _ip), ("zeta", _dp), ("pi_k", _dp), ("dots", _dp),
                ("t_seed", C.c_double), ("t_shift", C.c_double), ("t_check", C.c_double),
                ("status", C.c_int)]


_lib = None


def build() -> None:
    subprocess.check_call(["make", "-s", "-C", HERE, "all"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_shifted_lopbicg.restype = C.c_int
        L.orc_shifted_lopbicg.argtypes = [C.POINTER(_Csr), _dp, _dp, _dp, C.c_int, C.c_int,
                                          C.c_double, C.c_int, C.POINTER(_Trace)]
        L.orc_shifted_lopbicg_ranks.restype = C.c_int
        L.orc_shifted_lopbicg_ranks.argtypes = [C.POINTER(_Csr), _dp, _dp, _dp, C.c_int, C.c_int,
                                                C.c_double, C.c_int, C.c_int, C.POINTER(_Trace)]
        L.orc_mult.argtypes = [C.POINTER(_Csr), _dp, _dp]
        L.orc_mult_ranks.argtypes = [C.POINTER(_Csr), C.c_int, _dp, _dp]
        L.orc_dot_ranks.argtypes = [C.c_int, C.c_int, _dp, _dp]
        L.orc_dot_ranks.restype = C.c_double
        L.orc_true_residuals.argtypes = [C.POINTER(_Csr), _dp, _dp, _dp, C.c_int, _dp]
        L.orc_daxpy.argtypes = [C.c_int, C.c_double, _dp, _dp]
        L.orc_ddot.argtypes = [C.c_int, _dp, _dp]
        L.orc_ddot.restype = C.c_double
        L.orc_dscal.argtypes = [C.c_int, C.c_double, _dp]
        L.orc_dcopy.argtypes = [C.c_int, _dp, _dp]
        L.orc_ddot_v3.argtypes = [C.c_int, _dp, _dp, _dp]
        L.orc_ddot_v3.restype = C.c_double
        L.orc_num_threads.restype = C.c_int
        L.orc_coo2csr.restype = C.c_int
        L.orc_coo2csr.argtypes = [C.c_uint, C.c_uint, _up, _up, _dp, _up, _up, _dp]
        _lib = L
    return _lib


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


class _CsrHold:
    """Keeps the numpy arrays behind an orc_csr alive."""

    def __init__(self, A):
        self.val = np.ascontiguousarray(A.val, dtype=np.float64)
        self.col = np.ascontiguousarray(A.col, dtype=np.uint32)
        self.ptr = np.ascontiguousarray(A.ptr, dtype=np.uint32)
        self.s = _Csr(_p(self.val), _p(self.col, _up), _p(self.ptr, _up), A.n, int(self.ptr[-1]))


@dataclass
class OracleResult:
    k: int
    x: np.ndarray            # [M, n]
    r: np.ndarray            # final r (reference r_loc on exit)
    stop_iter: np.ndarray    # [M]
    switches: list           # [(k, new_seed)]
    alpha: np.ndarray
    beta: np.ndarray
    omega: np.ndarray
    res_hist: np.ndarray     # [(max_iter+1), M]
    seed_hist: np.ndarray
    zeta: np.ndarray
    pi_k: np.ndarray
    dots: np.ndarray         # [(max_iter+1), 5]
    t_seed: float = 0.0      # seconds in seed steps
    t_shift: float = 0.0     # seconds in the shift loop
    t_check: float = 0.0     # seconds in check + switch
    status: int = 0          # NO_SWITCH_TARGET (4) when the solve ended for that reason


def solve(A, b, sigma, seed, tol=1e-12, max_iter=2000, x0=None, ranks=1) -> OracleResult:
    """Run the restated reference solver. x0 defaults to zeros.  ranks > 1
    (a power of two): the arithmetic of the reference's run on that many MPI
    ranks (orc_shifted_lopbicg_ranks)."""
    L = lib()
    n = A.n
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    M = sigma.shape[0]
    cap = max_iter + 1
    h = _CsrHold(A)
    x = np.zeros((M, n)) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    r = np.array(b, dtype=np.float64, copy=True)
    stop_iter = np.zeros(M, dtype=np.int32)
    sw_k = np.zeros(max(M, 1), dtype=np.int32)
    sw_to = np.zeros(max(M, 1), dtype=np.int32)
    alpha = np.zeros(cap); beta = np.zeros(cap); omega = np.zeros(cap)
    res_hist = np.zeros(cap * M)
    seed_hist = np.zeros(cap, dtype=np.int32)
    zeta = np.zeros(M); pi_k = np.zeros(M)
    dots = np.zeros(cap * 5)
    tr = _Trace(_p(stop_iter, _ip), _p(sw_k, _ip), _p(sw_to, _ip), 0, _p(alpha), _p(beta),
                _p(omega), _p(res_hist), _p(seed_hist, _ip), _p(zeta), _p(pi_k), _p(dots),
                0.0, 0.0, 0.0, 0)
    k = L.orc_shifted_lopbicg_ranks(C.byref(h.s), _p(x), _p(r), _p(sigma), M, int(seed), float(tol),
                                    int(max_iter), int(ranks), C.byref(tr))
    if k < 0:
        raise ValueError("oracle rejected its arguments")
    sw = [(int(sw_k[i]), int(sw_to[i])) for i in range(tr.n_switch)]
    return OracleResult(k, x, r, stop_iter, sw, alpha, beta, omega, res_hist.reshape(cap, M),
                        seed_hist, zeta, pi_k, dots.reshape(cap, 5), tr.t_seed, tr.t_shift,
                        tr.t_check, tr.status)


def mult(A, x) -> np.ndarray:
    """y = A x with the reference's row-sequential summation."""
    h = _CsrHold(A)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.n)
    lib().orc_mult(C.byref(h.s), _p(x), _p(y))
    return y


def mult_ranks(A, x, ranks) -> np.ndarray:
    """y = A x as the reference computes it on `ranks` MPI ranks."""
    h = _CsrHold(A)
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.zeros(A.n)
    lib().orc_mult_ranks(C.byref(h.s), int(ranks), _p(x), _p(y))
    return y


def dot_ranks(x, y, ranks) -> float:
    """(x, y) as the reference computes it on `ranks` MPI ranks."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_dot_ranks(x.shape[0], int(ranks), _p(x), _p(y))


def true_residuals(A, x, b, sigma) -> np.ndarray:
    h = _CsrHold(A)
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    out = np.zeros(sigma.shape[0])
    lib().orc_true_residuals(C.byref(h.s), _p(x), _p(b), _p(sigma), sigma.shape[0], _p(out))
    return out


def num_threads() -> int:
    return lib().orc_num_threads()


def coo2csr(rows, row, col, val):
    """(ptr, col, val) of the reference's coo2csr (stable merge sort by row)."""
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col, dtype=np.uint32)
    val = np.ascontiguousarray(val, dtype=np.float64)
    nz = row.shape[0]
    ptr = np.zeros(rows + 1, dtype=np.uint32)
    ocol = np.zeros(max(nz, 1), dtype=np.uint32)
    oval = np.zeros(max(nz, 1))
    rc = lib().orc_coo2csr(rows, nz, _p(row, _up), _p(col, _up), _p(val), _p(ptr, _up),
                           _p(ocol, _up), _p(oval))
    if rc != 0:
        raise ValueError("oracle coo2csr rejected its arguments")
    return ptr, ocol[:nz], oval[:nz]
