"""A compiled C caller of the C-ABI (tests/c/driver.c, built by
__graft_entry__.build() with include/ and -lhbicg, as a reference driver
would be): the CPU check runs it where there is no GPU; the GPU test runs it
as a child process on the 64x64 workload in exact order and holds its x_set,
r, k, stop iterations and switch log to the oracle and its committed fixture
(tests/golden/c0_oracle.npz) bit for bit."""
import os
import subprocess

import numpy as np
import pytest

import pyoracle as O
from hbicg import problems as P

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "c", "_build", "driver")
GOLD = os.path.join(HERE, "golden", "c0_oracle.npz")


def _inputs(tmp_path, prob):
    A = prob.A
    mtx = str(tmp_path / "a.mtx")
    rows = np.repeat(np.arange(A.n), np.diff(A.ptr.astype(np.int64)))
    with open(mtx, "w") as f:
        f.write("%%MatrixMarket matrix coordinate real general\n")
        f.write(f"{A.n} {A.n} {A.nnz}\n")
        f.write("".join(f"{r + 1} {c + 1} {v!r}\n" for r, c, v in zip(rows, A.col, A.val.tolist())))
    bpath, spath = str(tmp_path / "b.bin"), str(tmp_path / "s.bin")
    np.ascontiguousarray(prob.b, dtype=np.float64).tofile(bpath)
    np.ascontiguousarray(prob.sigma, dtype=np.float64).tofile(spath)
    return mtx, bpath, spath


def _run(tmp_path, prob, seed, variant, exact):
    mtx, bpath, spath = _inputs(tmp_path, prob)
    out = str(tmp_path / f"out_{seed}_{variant}_{exact}.bin")
    res = subprocess.run([DRIVER, mtx, bpath, spath, str(prob.M), str(seed), variant, str(int(exact)), out],
                         capture_output=True, text=True, timeout=120)
    return res, out


def _parse(out, M, n):
    raw = np.fromfile(out, dtype=np.uint8)
    ints = raw[:4 * (2 + 3 * M)].view(np.int32)
    k, ns = int(ints[0]), int(ints[1])
    stop = ints[2:2 + M]
    sw = [(int(a), int(b)) for a, b in zip(ints[2 + M:2 + M + ns], ints[2 + 2 * M:2 + 2 * M + ns])]
    dbl = raw[4 * (2 + 3 * M):].view(np.float64)
    return k, stop, sw, dbl[:M * n].reshape(M, n), dbl[M * n:M * n + n]


def test_driver_built_and_reports_no_device(tmp_path):
    assert os.path.exists(DRIVER), "tests/c/_build/driver is built by __graft_entry__.build()"
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    res, _ = _run(tmp_path, P.config0(), 0, "dynamic", True)
    assert res.returncode == 1
    assert "returned -6" in res.stderr


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 8, 15])
def test_driver_exact_order_equals_oracle(tmp_path, seed):
    p = P.config0()
    print(f"case C driver exact order seed {seed}", flush=True)
    res, out = _run(tmp_path, p, seed, "dynamic", True)
    assert res.returncode == 0, res.stderr
    k, stop, sw, x, r = _parse(out, p.M, p.n)
    assert res.stdout.strip() == f"k={k}"
    gold = np.load(GOLD)
    assert k == int(gold[f"k_{seed}"])
    assert np.array_equal(stop, gold[f"stop_{seed}"])
    assert np.array_equal(np.array(sw, dtype=np.int32).reshape(-1, 2), gold[f"switch_{seed}"])
    assert np.array_equal(x[:, ::41], gold[f"xs_{seed}"])
    assert np.array_equal(x.sum(axis=1), gold[f"xsum_{seed}"])
    ref = O.solve(p.A, p.b, p.sigma, seed, tol=p.tol)
    assert np.array_equal(x, ref.x) and np.array_equal(r, ref.r)


@pytest.mark.gpu
def test_driver_default_order_within_bar(tmp_path):
    p = P.config0()
    ref = O.solve(p.A, p.b, p.sigma, 8, tol=p.tol)
    for variant in ("normal", "static"):
        print(f"case C driver default order {variant}", flush=True)
        res, out = _run(tmp_path, p, 8, variant, False)
        assert res.returncode == 0, res.stderr
        k, stop, sw, x, r = _parse(out, p.M, p.n)
        assert abs(k - ref.k) <= 2 and np.all(np.abs(stop - ref.stop_iter) <= 2)
        assert [s for _, s in sw] == [s for _, s in ref.switches]
        rel = np.linalg.norm(x - ref.x, axis=1) / np.linalg.norm(ref.x, axis=1)
        assert rel.max() <= 1e-9, rel.max()
