"""Python front end of the C-ABI library.

* ``dropin_solve`` calls the reference entry points ``shifted_lopbicg_*``
  exactly as a C driver would (host CSR_Matrix, host x_set and r).
* ``Plan`` wraps the device-resident ``hbicg_plan`` for solves on torch
  tensors that already live in HBM (benchmarks, PyTorch binding).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib

VARIANTS = ("normal", "static", "dynamic")


def make_params(tol=1e-12, max_iter=2000, window=0, exact_order=False, graph=-1, chunk=16,
                device=0, exit_on_error=False, upload_x=True, profile=False,
                growth_limit=1e8, trace=False, value_index=True, q_mode=0, fused=True,
                panel_cols=0, x_zero=False) -> _lib.Params:
    p = _lib.default_params()
    p.tol = float(tol)
    p.max_iter = int(max_iter)
    p.window = int(window)
    p.exact_order = int(bool(exact_order))
    p.graph = int(graph)
    p.chunk = int(chunk)
    p.device = int(device)
    p.exit_on_error = int(bool(exit_on_error))
    p.upload_x = int(bool(upload_x))
    p.profile = int(bool(profile))
    p.growth_limit = float(growth_limit)
    p.trace = int(bool(trace))
    p.value_index = int(bool(value_index))
    p.q_mode = int(q_mode)
    p.fused = int(fused) if not isinstance(fused, bool) else int(fused)
    p.panel_cols = int(panel_cols)
    p.x_zero = int(bool(x_zero))
    return p


@dataclass
class SolveResult:
    k: int
    x: object                 # [M, n] numpy (drop-in) or torch tensor (plan)
    r: object
    stop_iter: np.ndarray
    switches: list            # [(k, new_seed)]
    report: dict
    trace: object = None      # [k, M] residual estimates (params.trace), -1 = stopped


def _report_dict(rep: _lib.Report) -> dict:
    return {name: getattr(rep, name) for name, _ in rep._fields_}


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def dropin_solve(A, b, sigma, seed, variant="dynamic", params=None, x0=None) -> SolveResult:
    """Run shifted_lopbicg_<variant> on host arrays, like the reference's main.c."""
    L = _lib.load()
    if variant not in VARIANTS:
        raise ValueError(variant)
    p = params if params is not None else make_params()
    _lib.check(L.hbicg_set_params(C.byref(p)), "hbicg_set_params")
    n = A.n
    sigma = np.ascontiguousarray(sigma, dtype=np.float64)
    M = sigma.shape[0]
    val = np.ascontiguousarray(A.val, dtype=np.float64)
    col = np.ascontiguousarray(A.col, dtype=np.uint32)
    ptr = np.ascontiguousarray(A.ptr, dtype=np.uint32)
    up = C.POINTER(C.c_uint)
    diag = _lib.CSR_Matrix(_dp(val), col.ctypes.data_as(up), ptr.ctypes.data_as(up), int(ptr[-1]), n, n)
    offd_ptr = np.zeros(n + 1, dtype=np.uint32)
    offd = _lib.CSR_Matrix(None, None, offd_ptr.ctypes.data_as(up), 0, n, n)
    rc = np.array([n], dtype=np.int32)
    dp = np.array([0], dtype=np.int32)
    info = _lib.INFO_Matrix(int(ptr[-1]), n, n, b"MCRG", rc.ctypes.data_as(C.POINTER(C.c_int)),
                            dp.ctypes.data_as(C.POINTER(C.c_int)))
    x = np.zeros((M, n)) if x0 is None else np.array(x0, dtype=np.float64, copy=True)
    r = np.array(b, dtype=np.float64, copy=True)
    fn = getattr(L, f"shifted_lopbicg_{variant}")
    k = fn(C.byref(diag), C.byref(offd), C.byref(info), _dp(x), _dp(r), _dp(sigma), M, int(seed))
    _lib.check(k, f"shifted_lopbicg_{variant}")
    rep = _lib.Report()
    stop = np.zeros(M, dtype=np.int32)
    swk = np.zeros(M, dtype=np.int32)
    swt = np.zeros(M, dtype=np.int32)
    ip = C.POINTER(C.c_int)
    ns = L.hbicg_last_report(C.byref(rep), stop.ctypes.data_as(ip), M, swk.ctypes.data_as(ip),
                             swt.ctypes.data_as(ip), M)
    trace = None
    if p.trace:
        trace = np.empty((k, M))
        rows = _lib.check(L.hbicg_last_residual_trace(_dp(trace), M, k), "hbicg_last_residual_trace")
        trace = trace[:rows]
    return SolveResult(k, x, r, stop, [(int(swk[i]), int(swt[i])) for i in range(ns)],
                       _report_dict(rep), trace)


class Plan:
    """Device-resident solver plan over a CSR matrix held in torch tensors."""

    def __init__(self, A, M, params=None, device="cuda", stream=None, tensors=None):
        """A: problems.CSR on the host, or None with tensors=(val, col, ptr)
        already on the device (float64, int32, int32; col/ptr hold the
        unsigned 32-bit indices).  With params.value_index (the default) a
        qualifying matrix is read through a value table built here: values
        changed in place later need a new plan, or value_index=False, which
        reads val at every solve."""
        import torch
        self._torch = torch
        L = _lib.load()
        self.L = L
        dev = torch.device(device)
        self.device = dev
        if tensors is not None:
            self.val, self.col, self.ptr = tensors
            self.n = int(self.ptr.numel()) - 1
            self.nnz = int(self.val.numel())
        else:
            self.n = A.n
            self.val = torch.as_tensor(np.ascontiguousarray(A.val, dtype=np.float64)).to(dev)
            self.col = torch.as_tensor(np.ascontiguousarray(A.col, dtype=np.uint32).view(np.int32)).to(dev)
            self.ptr = torch.as_tensor(np.ascontiguousarray(A.ptr, dtype=np.uint32).view(np.int32)).to(dev)
            self.nnz = int(A.ptr[-1])
        self.M = int(M)
        self.params = params if params is not None else make_params()
        # A stream of the plan's own (default priority); every call orders it
        # after the caller's current stream and the caller's stream after it.
        s = stream if stream is not None else torch.cuda.Stream(dev)
        self.stream = s
        csr = _lib.Csr(self.val.data_ptr(), self.col.data_ptr(), self.ptr.data_ptr(), self.n, self.nnz)
        torch.cuda.synchronize(dev)
        self.h = L.hbicg_plan_create(C.byref(csr), 1, self.M, C.byref(self.params),
                                     C.c_void_p(s.cuda_stream))
        if not self.h:
            _lib.check(L.hbicg_errno() or -4, "hbicg_plan_create")

    def close(self):
        if getattr(self, "h", None):
            self.L.hbicg_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, b, sigma, seed, x=None, r_out=None):
        """b: [n] device tensor; sigma: host array [M]; x: [M, n] device tensor
        accumulated in place (zeros if None).  Returns (k, x, report)."""
        torch = self._torch
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        assert sigma.shape[0] == self.M
        if x is None:
            x = torch.zeros((self.M, self.n), dtype=torch.float64, device=self.device)
        assert x.is_contiguous() and x.dtype == torch.float64 and b.dtype == torch.float64
        rep = _lib.Report()
        self._enter()
        k = self.L.hbicg_plan_solve(self.h, C.c_void_p(b.data_ptr()), _dp(sigma), int(seed),
                                    C.c_void_p(x.data_ptr()),
                                    C.c_void_p(r_out.data_ptr()) if r_out is not None else None,
                                    1, C.byref(rep))
        self._leave()
        _lib.check(k, "hbicg_plan_solve")
        return k, x, _report_dict(rep)

    def _enter(self):
        self.stream.wait_stream(self._torch.cuda.current_stream(self.device))

    def _leave(self):
        self._torch.cuda.current_stream(self.device).wait_stream(self.stream)

    def report(self):
        rep = _lib.Report()
        stop = np.zeros(self.M, dtype=np.int32)
        swk = np.zeros(self.M, dtype=np.int32)
        swt = np.zeros(self.M, dtype=np.int32)
        ip = C.POINTER(C.c_int)
        ns = self.L.hbicg_plan_report(self.h, C.byref(rep), stop.ctypes.data_as(ip),
                                      swk.ctypes.data_as(ip), swt.ctypes.data_as(ip), self.M)
        return _report_dict(rep), stop, [(int(swk[i]), int(swt[i])) for i in range(ns)]

    def residual_trace(self):
        """[k, M] residual estimates of the last solve (plan made with
        trace=True): row k is iteration k, -1 where a shift had stopped."""
        k = self.report()[0]["k"]
        out = np.empty((k, self.M))
        rows = _lib.check(self.L.hbicg_plan_residual_trace(self.h, _dp(out), k),
                          "hbicg_plan_residual_trace")
        return out[:rows]

    KERNELS = ("spmv_seed", "fin_spmv_seed", "spmv_q_fused", "fin_spmv_q", "seed_update",
               "scalar_stage", "flush_coef", "seed_p_update", "window_flush")

    def kernel_stats(self):
        """Per-kernel {launches, total_ms, bytes, gbps} of the last profiled
        solve, one entry per launch of an iteration (include/hbicg.h)."""
        n = len(self.KERNELS)
        arr = (_lib.KernelStat * n)()
        m = self.L.hbicg_plan_kernel_stats(self.h, arr, n)
        out = {}
        for i in range(m):
            st = arr[i]
            gbps = st.bytes / (st.total_ms * 1e-3) / 1e9 if st.total_ms > 0 else 0.0
            out[self.KERNELS[i]] = {"launches": st.launches, "total_ms": st.total_ms,
                                    "bytes": st.bytes, "gbps": gbps}
            if st.replays:
                out[self.KERNELS[i]].update(replays=st.replays, replay_ms=st.replay_ms)
            if st.flops:
                out[self.KERNELS[i]].update(flops=st.flops)
        return out

    def true_residuals(self, b, sigma, x):
        sigma = np.ascontiguousarray(sigma, dtype=np.float64)
        out = np.zeros(self.M)
        self._enter()
        rc = self.L.hbicg_plan_true_residuals(self.h, C.c_void_p(b.data_ptr()), _dp(sigma),
                                              C.c_void_p(x.data_ptr()), 1, _dp(out))
        self._leave()
        _lib.check(rc, "hbicg_plan_true_residuals")
        return out

    def true_residuals_noalloc(self, b, sigma, x, out):
        """hbicg_plan_true_residuals_noalloc: sigma [M] and out [M] are float64
        device tensors; nothing is allocated by the library.  Returns out
        once the plan's stream has written it."""
        torch = self._torch
        for t in (b, sigma, x, out):
            assert t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
        assert sigma.numel() == self.M and out.numel() == self.M
        self._enter()
        rc = self.L.hbicg_plan_true_residuals_noalloc(self.h, C.c_void_p(b.data_ptr()),
                                                      C.c_void_p(sigma.data_ptr()),
                                                      C.c_void_p(x.data_ptr()),
                                                      C.c_void_p(out.data_ptr()))
        self._leave()
        _lib.check(rc, "hbicg_plan_true_residuals_noalloc")
        return out

    def spmv(self, x, shift=0.0):
        y = self._torch.empty_like(x)
        self._enter()
        rc = self.L.hbicg_plan_spmv(self.h, C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()),
                                    float(shift))
        self._leave()
        _lib.check(rc, "hbicg_plan_spmv")
        return y


class Batch:
    """Many independent shifted problems solved together on one GPU
    (hbicg_batch_*).  problems: list of problems.ShiftedProblem."""

    def __init__(self, problems, params=None, device="cuda", stream=None, device_spec=None):
        """problems: list of problems.ShiftedProblem (host data), or None with
        device_spec = list of dicts {val, col, ptr, b (device tensors),
        sigma (host array), seed} describing problems already on the GPU."""
        import torch
        self._torch = torch
        L = _lib.load()
        self.L = L
        dev = torch.device(device)
        self.device = dev
        self.params = params if params is not None else make_params(graph=1)
        self.stream = stream if stream is not None else torch.cuda.Stream(dev)
        self._keep = []
        if device_spec is None:
            spec = []
            for p in problems:
                spec.append({
                    "val": torch.as_tensor(np.ascontiguousarray(p.A.val, dtype=np.float64)).to(dev),
                    "col": torch.as_tensor(np.ascontiguousarray(p.A.col, dtype=np.uint32).view(np.int32)).to(dev),
                    "ptr": torch.as_tensor(np.ascontiguousarray(p.A.ptr, dtype=np.uint32).view(np.int32)).to(dev),
                    "b": torch.as_tensor(p.b).to(dev), "sigma": p.sigma, "seed": p.seed})
            self.problems = list(problems)
        else:
            spec = list(device_spec)
            self.problems = None
        self._spec = spec
        nprob = len(spec)
        self.n = [int(d["ptr"].numel()) - 1 for d in spec]
        csrs = (_lib.Csr * nprob)()
        for i, d in enumerate(spec):
            self._keep += [d["val"], d["col"], d["ptr"]]
            csrs[i] = _lib.Csr(d["val"].data_ptr(), d["col"].data_ptr(), d["ptr"].data_ptr(),
                               self.n[i], int(d["val"].numel()))
        self.sig = [np.ascontiguousarray(d["sigma"], dtype=np.float64) for d in spec]
        self.M = np.array([s_.shape[0] for s_ in self.sig], dtype=np.int32)
        torch.cuda.synchronize(dev)
        self.h = L.hbicg_batch_create(csrs, 1, nprob, self.M.ctypes.data_as(C.POINTER(C.c_int)),
                                      C.byref(self.params), C.c_void_p(self.stream.cuda_stream))
        if not self.h:
            _lib.check(L.hbicg_errno() or -4, "hbicg_batch_create")
        self.b = [d["b"] for d in spec]
        self.seed = np.array([d["seed"] for d in spec], dtype=np.int32)

    def close(self):
        if getattr(self, "h", None):
            self.L.hbicg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptrs(self, xs):
        nb = len(self.b)
        VP = C.c_void_p * nb
        b = VP(*[t.data_ptr() for t in self.b])
        s = VP(*[a.ctypes.data for a in self.sig])
        x = VP(*[t.data_ptr() for t in xs])
        return b, s, x

    def new_x(self):
        return [self._torch.zeros((int(m), n), dtype=self._torch.float64, device=self.device)
                for m, n in zip(self.M, self.n)]

    def solve(self, xs=None):
        """Solves every problem (xs accumulated in place; zeros if None).
        Returns (k per problem, xs, device ms)."""
        if xs is None:
            xs = self.new_x()
        b, s, x = self._ptrs(xs)
        k = np.zeros(len(self.b), dtype=np.int32)
        ms = C.c_double(0.0)
        vpp = C.POINTER(C.c_void_p)
        self.stream.wait_stream(self._torch.cuda.current_stream(self.device))
        rc = self.L.hbicg_batch_solve(self.h, C.cast(b, vpp), C.cast(s, vpp),
                                      self.seed.ctypes.data_as(C.POINTER(C.c_int)), C.cast(x, vpp),
                                      k.ctypes.data_as(C.POINTER(C.c_int)), C.byref(ms))
        self._torch.cuda.current_stream(self.device).wait_stream(self.stream)
        _lib.check(rc, "hbicg_batch_solve")
        return k, xs, ms.value

    def autotune(self, iters=2000):
        """hbicg_batch_autotune: each candidate layout timed over `iters`
        iterations (the default covers whole solves, flushes included)."""
        xs = self.new_x()
        b, s, x = self._ptrs(xs)
        t = _lib.Tuning()
        vpp = C.POINTER(C.c_void_p)
        self.stream.wait_stream(self._torch.cuda.current_stream(self.device))
        rc = self.L.hbicg_batch_autotune(self.h, C.cast(b, vpp), C.cast(s, vpp),
                                         self.seed.ctypes.data_as(C.POINTER(C.c_int)), C.cast(x, vpp),
                                         int(iters), C.byref(t))
        self._torch.cuda.current_stream(self.device).wait_stream(self.stream)
        _lib.check(rc, "hbicg_batch_autotune")
        return {"window": t.window, "group_cap": t.group_cap, "ms_per_iter": t.ms_per_iter}

    def report(self, i):
        m = int(self.M[i])
        rep = _lib.Report()
        stop = np.zeros(m, dtype=np.int32)
        swk = np.zeros(m, dtype=np.int32)
        swt = np.zeros(m, dtype=np.int32)
        ip = C.POINTER(C.c_int)
        ns = self.L.hbicg_batch_report(self.h, i, C.byref(rep), stop.ctypes.data_as(ip),
                                       swk.ctypes.data_as(ip), swt.ctypes.data_as(ip), m)
        return _report_dict(rep), stop, [(int(swk[j]), int(swt[j])) for j in range(ns)]

    def true_residuals(self, i, x):
        out = np.zeros(int(self.M[i]))
        self.stream.wait_stream(self._torch.cuda.current_stream(self.device))
        rc = self.L.hbicg_batch_true_residuals(self.h, i, C.c_void_p(self.b[i].data_ptr()),
                                               _dp(self.sig[i]), C.c_void_p(x.data_ptr()), _dp(out))
        self._torch.cuda.current_stream(self.device).wait_stream(self.stream)
        _lib.check(rc, "hbicg_batch_true_residuals")
        return out
