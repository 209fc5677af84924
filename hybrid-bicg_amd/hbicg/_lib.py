"""ctypes binding of libhbicg.so, the C-ABI library (include/*.h).

The library is the product: this module only marshals pointers.  Loading
fails loudly when the library has not been built; there is no fallback.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "_build", "libhbicg.so")

HBICG_E = {
    -1: "bad argument", -2: "matrix is not square", -3: "non-empty off-process block",
    -4: "HIP runtime error", -5: "allocation failed", -6: "no HIP device",
    -7: "input larger than the no-allocation scratch",
}

_dp = C.POINTER(C.c_double)
_up = C.POINTER(C.c_uint)
_ip = C.POINTER(C.c_int)


class CSR_Matrix(C.Structure):
    """include/matrix.h CSR_Matrix (layout of the reference's src/matrix.h:19-26)."""
    _fields_ = [("val", _dp), ("col", _up), ("ptr", _up), ("nz", C.c_uint), ("rows", C.c_uint),
                ("cols", C.c_uint)]


class COO_Matrix(C.Structure):
    """include/matrix.h COO_Matrix (src/matrix.h:10-17)."""
    _fields_ = [("val", _dp), ("row", _up), ("col", _up), ("nz", C.c_uint), ("rows", C.c_uint),
                ("cols", C.c_uint)]


class INFO_Matrix(C.Structure):
    """include/matrix.h INFO_Matrix (src/matrix.h:28-33)."""
    _fields_ = [("nz", C.c_uint), ("rows", C.c_uint), ("cols", C.c_uint), ("code", C.c_char * 4),
                ("recvcounts", _ip), ("displs", _ip)]


class Params(C.Structure):
    """include/hbicg.h hbicg_params."""
    _fields_ = [("tol", C.c_double), ("max_iter", C.c_int), ("window", C.c_int),
                ("exact_order", C.c_int), ("graph", C.c_int), ("chunk", C.c_int),
                ("device", C.c_int), ("exit_on_error", C.c_int), ("upload_x", C.c_int),
                ("profile", C.c_int), ("growth_limit", C.c_double), ("trace", C.c_int),
                ("value_index", C.c_int), ("q_mode", C.c_int), ("fused", C.c_int),
                ("panel_cols", C.c_int), ("x_zero", C.c_int)]


class KernelStat(C.Structure):
    """include/hbicg.h hbicg_kernel_stat."""
    _fields_ = [("launches", C.c_int), ("total_ms", C.c_double), ("bytes", C.c_double),
                ("replays", C.c_int), ("replay_ms", C.c_double), ("flops", C.c_double)]


class Report(C.Structure):
    """include/hbicg.h hbicg_report."""
    _fields_ = [("k", C.c_int), ("stop_count", C.c_int), ("n_switch", C.c_int),
                ("seed_final", C.c_int), ("status", C.c_int), ("M", C.c_int), ("n", C.c_int),
                ("dot_r", C.c_double), ("dot_zero", C.c_double), ("rTr", C.c_double),
                ("t_setup_ms", C.c_double), ("t_solve_ms", C.c_double), ("t_total_ms", C.c_double),
                ("window", C.c_int), ("value_table", C.c_int),
                ("diag_tiles", C.c_int), ("q_unfused", C.c_int), ("panels", C.c_int),
                ("const_tiles", C.c_int)]


class Tuning(C.Structure):
    """include/hbicg.h hbicg_tuning."""
    _fields_ = [("window", C.c_int), ("group_cap", C.c_int), ("ms_per_iter", C.c_double)]


class Csr(C.Structure):
    """include/hbicg.h hbicg_csr."""
    _fields_ = [("val", C.c_void_p), ("col", C.c_void_p), ("ptr", C.c_void_p), ("n", C.c_uint),
                ("nz", C.c_uint)]


# Every symbol the headers under include/ declare (checked by the CPU tests).
EXPORTS = {
    "solver.h": ["shifted_lopbicg_dynamic", "shifted_lopbicg_static", "shifted_lopbicg_normal"],
    "vector.h": ["my_daxpy", "my_ddot", "my_dscal", "my_dcopy"],
    "matrix.h": ["coo_init_matrix", "coo_free_matrix", "csr_init_matrix", "csr_free_matrix",
                 "coo_load_matrix", "coo2csr", "csr_load_matrix", "MPI_csr_load_matrix_block",
                 "mult"],
    "hbicg.h": ["hbicg_default_params", "hbicg_set_params", "hbicg_get_params",
                "hbicg_last_report", "hbicg_last_residual_trace", "hbicg_plan_create", "hbicg_plan_destroy", "hbicg_errno",
                "hbicg_plan_solve", "hbicg_plan_true_residuals",
                "hbicg_plan_true_residuals_noalloc", "hbicg_plan_spmv",
                "hbicg_plan_report", "hbicg_plan_residual_trace", "hbicg_plan_kernel_stats",
                "hbicg_batch_create",
                "hbicg_batch_destroy", "hbicg_batch_solve", "hbicg_batch_report",
                "hbicg_batch_true_residuals", "hbicg_batch_autotune", "hbicg_csr_transpose",
                "hbicg_csr_transpose_pattern", "hbicg_gather_values", "hbicg_adjoint_gradients",
                "hbicg_adjoint_work_bytes", "hbicg_adjoint_gradients_ws", "hbicg_coo2csr_device", "hbicg_csr_load_device",
                "hbicg_csr_free_device"],
}

_lib = None


def load():
    """Load libhbicg.so and declare its signatures (raises if not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libhbicg.so not built at {LIB_PATH}; run __graft_entry__.build()")
    # torch carries its own HIP runtime under the same soname
    # (libamdhip64.so.7) as the one the library links: whichever is loaded
    # first serves the whole process, and torch's device initialisation
    # fails on the other one ("No HIP GPUs are available").  Load torch's
    # first, when torch is installed, so that both work in one process.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    sig = [C.POINTER(CSR_Matrix), C.POINTER(CSR_Matrix), C.POINTER(INFO_Matrix), _dp, _dp, _dp,
           C.c_int, C.c_int]
    for name in EXPORTS["solver.h"]:
        f = getattr(L, name)
        f.argtypes = sig
        f.restype = C.c_int
    L.my_daxpy.argtypes = [C.c_int, C.c_double, _dp, _dp]
    L.my_ddot.argtypes = [C.c_int, _dp, _dp]
    L.my_ddot.restype = C.c_double
    L.my_dscal.argtypes = [C.c_int, C.c_double, _dp]
    L.my_dcopy.argtypes = [C.c_int, _dp, _dp]
    L.csr_init_matrix.argtypes = [C.POINTER(CSR_Matrix)]
    L.csr_free_matrix.argtypes = [C.POINTER(CSR_Matrix)]
    L.coo_init_matrix.argtypes = [C.POINTER(COO_Matrix)]
    L.coo_free_matrix.argtypes = [C.POINTER(COO_Matrix)]
    L.coo_load_matrix.argtypes = [C.c_char_p, C.POINTER(COO_Matrix)]
    L.coo_load_matrix.restype = C.c_int
    L.coo2csr.argtypes = [C.POINTER(COO_Matrix), C.POINTER(CSR_Matrix)]
    L.csr_load_matrix.argtypes = [C.c_char_p, C.POINTER(CSR_Matrix)]
    L.csr_load_matrix.restype = C.c_int
    L.MPI_csr_load_matrix_block.argtypes = [C.c_char_p, C.POINTER(CSR_Matrix),
                                            C.POINTER(CSR_Matrix), C.POINTER(INFO_Matrix)]
    L.mult.argtypes = [C.POINTER(CSR_Matrix), _dp, _dp]
    L.hbicg_default_params.argtypes = [C.POINTER(Params)]
    L.hbicg_set_params.argtypes = [C.POINTER(Params)]
    L.hbicg_set_params.restype = C.c_int
    L.hbicg_get_params.argtypes = [C.POINTER(Params)]
    L.hbicg_get_params.restype = C.c_int
    L.hbicg_last_report.argtypes = [C.POINTER(Report), _ip, C.c_int, _ip, _ip, C.c_int]
    L.hbicg_last_report.restype = C.c_int
    L.hbicg_last_residual_trace.argtypes = [_dp, C.c_int, C.c_int]
    L.hbicg_last_residual_trace.restype = C.c_int
    L.hbicg_plan_residual_trace.argtypes = [C.c_void_p, _dp, C.c_int]
    L.hbicg_plan_residual_trace.restype = C.c_int
    L.hbicg_plan_create.argtypes = [C.POINTER(Csr), C.c_int, C.c_int, C.POINTER(Params), C.c_void_p]
    L.hbicg_plan_create.restype = C.c_void_p
    L.hbicg_plan_destroy.argtypes = [C.c_void_p]
    L.hbicg_errno.restype = C.c_int
    L.hbicg_plan_solve.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_int, C.POINTER(Report)]
    L.hbicg_plan_solve.restype = C.c_int
    L.hbicg_plan_true_residuals.argtypes = [C.c_void_p, C.c_void_p, _dp, C.c_void_p, C.c_int, _dp]
    L.hbicg_plan_true_residuals.restype = C.c_int
    L.hbicg_plan_true_residuals_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_void_p]
    L.hbicg_plan_true_residuals_noalloc.restype = C.c_int
    L.hbicg_plan_spmv.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]
    L.hbicg_plan_spmv.restype = C.c_int
    L.hbicg_plan_report.argtypes = [C.c_void_p, C.POINTER(Report), _ip, _ip, _ip, C.c_int]
    L.hbicg_plan_report.restype = C.c_int
    L.hbicg_plan_kernel_stats.argtypes = [C.c_void_p, C.POINTER(KernelStat), C.c_int]
    L.hbicg_plan_kernel_stats.restype = C.c_int
    vpp = C.POINTER(C.c_void_p)
    L.hbicg_batch_create.argtypes = [C.POINTER(Csr), C.c_int, C.c_int, _ip, C.POINTER(Params), C.c_void_p]
    L.hbicg_batch_create.restype = C.c_void_p
    L.hbicg_batch_destroy.argtypes = [C.c_void_p]
    L.hbicg_batch_solve.argtypes = [C.c_void_p, vpp, vpp, _ip, vpp, _ip, _dp]
    L.hbicg_batch_solve.restype = C.c_int
    L.hbicg_batch_report.argtypes = [C.c_void_p, C.c_int, C.POINTER(Report), _ip, _ip, _ip, C.c_int]
    L.hbicg_batch_report.restype = C.c_int
    L.hbicg_batch_true_residuals.argtypes = [C.c_void_p, C.c_int, C.c_void_p, _dp, C.c_void_p, _dp]
    L.hbicg_batch_true_residuals.restype = C.c_int
    L.hbicg_batch_autotune.argtypes = [C.c_void_p, vpp, vpp, _ip, vpp, C.c_int, C.POINTER(Tuning)]
    L.hbicg_batch_autotune.restype = C.c_int
    L.hbicg_csr_transpose.argtypes = [C.POINTER(Csr), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.hbicg_csr_transpose.restype = C.c_int
    L.hbicg_adjoint_gradients.argtypes = [C.POINTER(Csr), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.hbicg_adjoint_gradients.restype = C.c_int
    L.hbicg_csr_transpose_pattern.argtypes = [C.POINTER(Csr), C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.hbicg_csr_transpose_pattern.restype = C.c_int
    L.hbicg_gather_values.argtypes = [C.c_void_p, C.c_void_p, C.c_uint, C.c_void_p, C.c_void_p]
    L.hbicg_gather_values.restype = C.c_int
    L.hbicg_adjoint_work_bytes.argtypes = [C.c_uint, C.c_int]
    L.hbicg_adjoint_work_bytes.restype = C.c_size_t
    L.hbicg_adjoint_gradients_ws.argtypes = [C.POINTER(Csr), C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t,
                                             C.c_void_p]
    L.hbicg_adjoint_gradients_ws.restype = C.c_int
    L.hbicg_coo2csr_device.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint, C.c_uint,
                                       C.c_uint, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.hbicg_coo2csr_device.restype = C.c_int
    L.hbicg_csr_load_device.argtypes = [C.c_char_p, C.POINTER(Csr), C.POINTER(INFO_Matrix), C.c_void_p]
    L.hbicg_csr_load_device.restype = C.c_int
    L.hbicg_csr_free_device.argtypes = [C.POINTER(Csr)]
    L.hbicg_csr_free_device.restype = None
    _lib = L
    return L


def check(rc: int, what: str) -> int:
    if rc < 0:
        raise RuntimeError(f"{what} failed: {HBICG_E.get(rc, rc)}")
    return rc


def default_params() -> Params:
    p = Params()
    load().hbicg_default_params(C.byref(p))
    return p
