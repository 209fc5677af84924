"""GPU COO -> CSR (Matrix Market ingest) against the oracle's restatement of
the reference's coo2csr (src/matrix.c:125-187, 206-232: stable merge sort by
row), called through the C-ABI.  Integer/index work: results must be equal
bit for bit."""
import ctypes as C

import numpy as np
import pytest

import pyoracle as O
from hbicg import _lib
from hbicg import problems as P

pytestmark = pytest.mark.gpu

E_ARG = -1   # HBICG_E_ARG


def _d2h(dst, src_ptr):
    """hipMemcpy device -> host of dst.nbytes bytes (device memory the
    library allocated, not a torch tensor)."""
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip.hipMemcpy(dst.ctypes.data_as(C.c_void_p), C.c_void_p(src_ptr), dst.nbytes, 2)


def _dev(a, torch):
    a = np.ascontiguousarray(a)
    if a.dtype == np.uint32:
        a = a.view(np.int32)
    return torch.as_tensor(a).to("cuda")


def gpu_coo2csr(rows, cols, row, col, val):
    """Returns (rc, ptr, col, val) of hbicg_coo2csr_device."""
    import torch
    L = _lib.load()
    nz = len(row)
    drow, dcol, dval = _dev(np.asarray(row, np.uint32), torch), _dev(np.asarray(col, np.uint32), torch), \
        _dev(np.asarray(val, np.float64), torch)
    ptr = torch.full((rows + 1,), -7, dtype=torch.int32, device="cuda")
    ocol = torch.full((max(nz, 1),), -7, dtype=torch.int32, device="cuda")
    oval = torch.full((max(nz, 1),), -7.0, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    rc = L.hbicg_coo2csr_device(drow.data_ptr(), dcol.data_ptr(), dval.data_ptr(), nz, rows, cols,
                                ptr.data_ptr(), ocol.data_ptr(), oval.data_ptr(), None)
    return (rc, ptr.cpu().numpy().view(np.uint32), ocol.cpu().numpy().view(np.uint32)[:nz],
            oval.cpu().numpy()[:nz])


def _check(rows, row, col, val, cols=None):
    cols = rows if cols is None else cols
    rc, ptr, oc, ov = gpu_coo2csr(rows, cols, row, col, val)
    assert rc == 0
    eptr, ec, ev = O.coo2csr(rows, row, col, val)
    assert np.array_equal(ptr, eptr)
    assert np.array_equal(oc, ec)
    assert np.array_equal(ov, ev)


@pytest.mark.parametrize("rows,nz,seed", [
    (1, 1000, 0),            # one row: no radix pass
    (7, 0, 1),               # no entries
    (300, 5000, 2),          # two passes, many entries per row
    (70_000, 200_000, 3),    # three passes, empty rows, partial last tile
    ((1 << 24) + 5, 100_000, 4),   # four passes
])
def test_random_against_oracle(rows, nz, seed):
    print(f"case coo2csr rows={rows} nz={nz} seed={seed}", flush=True)
    rng = np.random.default_rng(seed)
    row = rng.integers(0, rows, nz).astype(np.uint32)
    col = rng.integers(0, min(rows, 50), nz).astype(np.uint32)   # many duplicated (row, col)
    val = rng.standard_normal(nz)
    _check(rows, row, col, val)


def test_long_row_across_tiles():
    """One row with 100k entries spread over many sort tiles, interleaved
    with short rows: its entries must stay in input order."""
    print("case coo2csr long row", flush=True)
    rng = np.random.default_rng(11)
    nz = 150_000
    row = np.where(rng.random(nz) < 0.66, 5, rng.integers(0, 1000, nz)).astype(np.uint32)
    col = np.arange(nz, dtype=np.uint32) % 1000
    val = np.arange(nz, dtype=np.float64)
    _check(1000, row, col, val)


def test_rejects_out_of_range_indices():
    print("case coo2csr bad indices", flush=True)
    rc, ptr, _, _ = gpu_coo2csr(4, 4, [0, 4], [0, 0], [1.0, 2.0])
    assert rc == E_ARG and np.all(ptr.view(np.int32) == -7)
    rc, _, _, _ = gpu_coo2csr(4, 4, [0, 1], [0, 9], [1.0, 2.0])
    assert rc == E_ARG


@pytest.mark.parametrize("cfg", ["config1", "config2"])
def test_fullsize_shuffled_roundtrip(cfg):
    """The 128^3 stencil and the 1e6-row power-law matrix, their entries
    shuffled: the GPU result equals the oracle's, and sorting each row's
    entries by column gives back the generator's CSR."""
    print(f"case coo2csr full size {cfg}", flush=True)
    A = getattr(P, cfg)().A
    n = A.n
    ptr0 = np.asarray(A.ptr, dtype=np.uint32)
    row = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ptr0))
    perm = np.random.default_rng(5).permutation(len(row))
    row, col, val = row[perm], np.asarray(A.col, np.uint32)[perm], np.asarray(A.val)[perm]
    rc, ptr, oc, ov = gpu_coo2csr(n, n, row, col, val)
    assert rc == 0
    eptr, ec, ev = O.coo2csr(n, row, col, val)
    assert np.array_equal(ptr, eptr) and np.array_equal(oc, ec) and np.array_equal(ov, ev)
    assert np.array_equal(ptr, ptr0)
    r_of = np.repeat(np.arange(n), np.diff(ptr.astype(np.int64)))
    o = np.lexsort((oc, r_of))
    assert np.array_equal(oc[o], np.asarray(A.col, np.uint32)) and np.array_equal(ov[o], np.asarray(A.val))


def _write_mtx(path, n, row, col, val):
    with open(path, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n% shuffled entries\n")
        f.write(f"{n} {n} {len(row)}\n")
        for i, j, v in zip(row, col, val):
            f.write(f"{int(i) + 1} {int(j) + 1} {float(v)!r}\n")


def test_matrix_market_file_to_device_solve(tmp_path):
    """A Matrix Market file read into device CSR by hbicg_csr_load_device
    gives the oracle's coo2csr of the file's entries, and a plan built on
    that device CSR solves exactly like the drop-in on the host CSR."""
    print("case csr_load_device + plan solve", flush=True)
    c0 = P.config0()
    A = c0.A
    n = A.n
    ptr0 = np.asarray(A.ptr, dtype=np.uint32)
    row = np.repeat(np.arange(n, dtype=np.uint32), np.diff(ptr0))
    perm = np.random.default_rng(9).permutation(len(row))
    row, col, val = row[perm], np.asarray(A.col, np.uint32)[perm], np.asarray(A.val)[perm]
    path = str(tmp_path / "c0.mtx")
    _write_mtx(path, n, row, col, val)
    L = _lib.load()
    csr = _lib.Csr()
    rcounts = (C.c_int * 1)(0)
    displs = (C.c_int * 1)(-1)
    info = _lib.INFO_Matrix(0, 0, 0, b"", rcounts, displs)
    assert L.hbicg_csr_load_device(path.encode(), C.byref(csr), C.byref(info), None) == 0
    try:
        assert (csr.n, csr.nz, info.rows, info.cols, info.nz) == (n, len(row), n, n, len(row))
        assert info.code[:3] == b"MCR" and rcounts[0] == n and displs[0] == 0
        import torch
        hptr = np.empty(n + 1, np.uint32)
        hcol = np.empty(csr.nz, np.uint32)
        hval = np.empty(csr.nz)
        torch.cuda.synchronize()
        for dst, src in ((hptr, csr.ptr), (hcol, csr.col), (hval, csr.val)):
            assert _d2h(dst, src) == 0
        eptr, ec, ev = O.coo2csr(n, row, col, val)
        assert np.array_equal(hptr, eptr) and np.array_equal(hcol, ec) and np.array_equal(hval, ev)

        # the device CSR drives a plan; the file's row order changes the
        # within-row summation order, so compare against the drop-in on the
        # same (file-ordered) host CSR, in exact order: bit for bit
        from hbicg import dropin_solve, make_params
        p = make_params(tol=c0.tol, exact_order=True)
        A_file = P.CSR(ptr=hptr, col=hcol, val=hval, n=n)
        ref = dropin_solve(A_file, c0.b, c0.sigma, 0, "dynamic", p)
        pl = L.hbicg_plan_create(C.byref(csr), 1, c0.M, C.byref(p), None)
        assert pl
        try:
            x = np.zeros((c0.M, n))
            b = np.ascontiguousarray(c0.b, dtype=np.float64)
            sig = np.ascontiguousarray(c0.sigma, dtype=np.float64)
            dp = C.POINTER(C.c_double)
            k = L.hbicg_plan_solve(pl, b.ctypes.data, sig.ctypes.data_as(dp), 0, x.ctypes.data,
                                   None, 0, None)
            assert k == ref.k
            assert np.array_equal(x, ref.x)
        finally:
            L.hbicg_plan_destroy(pl)
    finally:
        L.hbicg_csr_free_device(C.byref(csr))
