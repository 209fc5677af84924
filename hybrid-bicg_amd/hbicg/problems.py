"""Deterministic synthetic shifted-system families.

The reference ships no data (its matrices are SuiteSparse files loaded by
src/matrix.c:402-419), so every workload here is generated.  All randomness
comes from SplitMix64 over a (seed, stream) key, so a problem is a pure
function of its parameters on any machine.  Matrices are CSR with 32-bit
indices, rows in order and columns ascending inside a row: the layout the
reference's coo2csr produces from a row-major, column-sorted Matrix Market
file (src/matrix.c:206-232).

Families:
  * 2-D 5-point and 3-D 7-point convection-diffusion stencils on n_x*n_y(*n_z)
    grids with Dirichlet boundaries (missing neighbours dropped);
  * a random nonsymmetric matrix with power-law row lengths (irregular rows);
  * the batched and adjoint workloads built from those.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_M64 = (1 << 64) - 1
_GAMMA = 0x9E3779B97F4A7C15


def _mix64_scalar(z: int) -> int:
    z &= _M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
    return z ^ (z >> 31)


def stream_key(seed: int, stream: int) -> int:
    """64-bit key of an independent random stream (seed, stream)."""
    return _mix64_scalar((seed * 0x100000001B3 + _mix64_scalar(stream + 0x632BE59BD9B4E019)) & _M64)


def splitmix64(key: int, count: int, offset: int = 0) -> np.ndarray:
    """count outputs of SplitMix64 started at state key, skipping offset."""
    with np.errstate(over="ignore"):
        z = np.arange(offset + 1, offset + count + 1, dtype=np.uint64)
        z *= np.uint64(_GAMMA)
        z += np.uint64(key)
        t = z >> np.uint64(30)
        z ^= t
        z *= np.uint64(0xBF58476D1CE4E5B9)
        np.right_shift(z, np.uint64(27), out=t)
        z ^= t
        z *= np.uint64(0x94D049BB133111EB)
        np.right_shift(z, np.uint64(31), out=t)
        z ^= t
        return z


def uniform(seed: int, stream: int, count: int, lo: float = 0.0, hi: float = 1.0,
            offset: int = 0) -> np.ndarray:
    """count doubles uniform in [lo, hi) from stream (seed, stream)."""
    z = splitmix64(stream_key(seed, stream), count, offset)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


@dataclass
class CSR:
    """0-based CSR, float64 values, uint32 column indices and row pointers."""
    ptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    n: int

    @property
    def nnz(self) -> int:
        return int(self.ptr[-1])

    def to_scipy(self):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col.astype(np.int64), self.ptr.astype(np.int64)),
                             shape=(self.n, self.n))

    def transpose(self) -> "CSR":
        """A^T in CSR (stable: column order inside each row follows row order of A)."""
        rows = np.repeat(np.arange(self.n, dtype=np.int64), np.diff(self.ptr.astype(np.int64)))
        order = np.argsort(self.col, kind="stable")
        tcol = rows[order].astype(np.uint32)
        tval = self.val[order]
        counts = np.bincount(self.col.astype(np.int64), minlength=self.n)
        tptr = np.zeros(self.n + 1, dtype=np.uint32)
        tptr[1:] = np.cumsum(counts)
        return CSR(tptr, tcol, tval, self.n)


@dataclass
class ShiftedProblem:
    """One family (A + sigma_j I) x_j = b with its seed index and stopping rule."""
    name: str
    A: CSR
    b: np.ndarray
    sigma: np.ndarray
    seed: int
    tol: float = 1e-10
    max_iter: int = 2000
    extra: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return self.A.n

    @property
    def M(self) -> int:
        return int(self.sigma.shape[0])


def _stencil(dims, offsets_coefs, diag) -> CSR:
    """Build a constant-coefficient stencil matrix in CSR, columns ascending.

    dims: grid extents (fastest first); offsets_coefs: list of
    (axis, step(+1/-1), coefficient); diag: diagonal value."""
    dims = tuple(int(d) for d in dims)
    n = int(np.prod(dims))
    strides = [1]
    for d in dims[:-1]:
        strides.append(strides[-1] * d)
    idx = np.arange(n, dtype=np.int64)
    coords = []
    rem = idx
    for d in dims:
        coords.append(rem % d)
        rem = rem // d
    entries = []  # (column offset, coefficient, valid mask)
    for axis, step, coef in offsets_coefs:
        c = coords[axis] + step
        valid = (c >= 0) & (c < dims[axis])
        entries.append((step * strides[axis], coef, valid))
    entries.append((0, diag, np.ones(n, dtype=bool)))
    entries.sort(key=lambda e: e[0])
    k = len(entries)
    cols = np.empty((n, k), dtype=np.int64)
    vals = np.empty((n, k), dtype=np.float64)
    mask = np.empty((n, k), dtype=bool)
    for t, (off, coef, valid) in enumerate(entries):
        cols[:, t] = idx + off
        vals[:, t] = coef
        mask[:, t] = valid
    counts = mask.sum(axis=1)
    ptr = np.zeros(n + 1, dtype=np.uint32)
    ptr[1:] = np.cumsum(counts)
    return CSR(ptr, cols[mask].astype(np.uint32), vals[mask], n)


def conv_diff_2d(nx: int, ny: int, diag: float = 4.0, conv: float = 0.3) -> CSR:
    """5-point convection-diffusion: diag on the diagonal, -1+conv on the
    east/north neighbours and -1-conv on the west/south neighbours."""
    return _stencil((nx, ny), [(0, +1, -1.0 + conv), (0, -1, -1.0 - conv),
                               (1, +1, -1.0 + conv), (1, -1, -1.0 - conv)], diag)


def conv_diff_3d(nx: int, ny: int, nz: int, diag: float = 6.0, conv: float = 0.2) -> CSR:
    """7-point convection-diffusion: -1+conv on the +axis neighbours, -1-conv
    on the -axis neighbours."""
    terms = []
    for ax in range(3):
        terms.append((ax, +1, -1.0 + conv))
        terms.append((ax, -1, -1.0 - conv))
    return _stencil((nx, ny, nz), terms, diag)


def _zipf_exponent(mean: float, cap: int) -> float:
    ls = np.arange(1, cap + 1, dtype=np.float64)

    def m(a):
        w = ls ** (-a)
        return float((ls * w).sum() / w.sum())

    lo, hi = 0.5, 4.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if m(mid) > mean:
            lo = mid
        else:
            hi = mid
    return 0.5 * (lo + hi)


def zipf_matrix(n: int, mean: float = 30.0, cap: int = 2000, seed: int = 2) -> CSR:
    """Random nonsymmetric matrix: off-diagonal row lengths from a truncated
    power law on [1, cap] with the given mean, columns uniform without repeats
    (and never the diagonal), values uniform in [-1, 1], diagonal 1.1 times the
    row's absolute off-diagonal sum."""
    a = _zipf_exponent(mean, cap)
    ls = np.arange(1, cap + 1, dtype=np.float64)
    cdf = np.cumsum(ls ** (-a))
    cdf /= cdf[-1]
    u = uniform(seed, 11, n)
    lens = (np.searchsorted(cdf, u, side="right") + 1).clip(1, cap).astype(np.int64)
    lens = np.minimum(lens, n - 1)
    tot = int(lens.sum())
    rows = np.repeat(np.arange(n, dtype=np.int64), lens)
    rown = rows * n
    key = rown + (splitmix64(stream_key(seed, 12), tot) % np.uint64(n)).astype(np.int64)
    key.sort()                           # row-major keys: rows stay grouped, columns ascend
    cols = key - rown
    del key
    bad = cols == rows
    bad[1:] |= (cols[1:] == cols[:-1]) & (rows[1:] == rows[:-1])   # a repeat in the row
    starts = np.zeros(n, dtype=np.int64)
    starts[1:] = np.cumsum(lens)[:-1]
    for r in np.unique(rows[bad]):
        # rows with a repeated or diagonal column draw replacements from a
        # stream of their own until they hold lens[r] distinct columns
        a, m = int(starts[r]), int(lens[r])
        have = sorted(set(int(c) for c in cols[a:a + m]) - {int(r)})
        extra, off = [], 0
        while len(have) + len(extra) < m:
            cand = splitmix64(stream_key(seed, 1000 + int(r)), 64, off) % np.uint64(n)
            off += 64
            taken = set(have) | set(extra) | {int(r)}
            for c in cand:
                c = int(c)
                if c not in taken and len(have) + len(extra) < m:
                    extra.append(c)
                    taken.add(c)
        cols[a:a + m] = np.sort(np.array(have + extra, dtype=np.int64))
    off_val = uniform(seed, 13, tot, -1.0, 1.0)
    absum = np.bincount(rows, weights=np.abs(off_val), minlength=n)
    # the diagonal goes after the row's columns smaller than the row index
    below = np.bincount(rows, weights=(cols < rows), minlength=n).astype(np.int64)
    at = starts + below
    all_cols = np.insert(cols, at, np.arange(n, dtype=np.int64))
    all_vals = np.insert(off_val, at, 1.1 * absum)
    ptr = np.zeros(n + 1, dtype=np.uint32)
    ptr[1:] = np.cumsum(lens + 1)
    return CSR(ptr, all_cols.astype(np.uint32), all_vals, n)


def logspace_shifts(m: int, lo: float, hi: float) -> np.ndarray:
    if m == 1:
        return np.array([lo], dtype=np.float64)
    return np.array([lo * (hi / lo) ** (k / (m - 1)) for k in range(m)], dtype=np.float64)


# ---- the five named workloads ------------------------------------------------

def config0(grid: int = 64, m: int = 16, seed_index: int = 0) -> ShiftedProblem:
    """2-D 5-point convection-diffusion on a 64x64 grid (n=4096, 20,224 nnz):
    diagonal 4 + 0.01, off-diagonals -1 +- 0.3, b uniform in [-1,1] from seed 0,
    16 shifts 0.01*(k+1), tolerance 1e-10, seed = smallest shift."""
    A = conv_diff_2d(grid, grid, diag=4.0 + 0.01, conv=0.3)
    b = uniform(0, 1, A.n, -1.0, 1.0)
    sigma = np.array([0.01 * (k + 1) for k in range(m)], dtype=np.float64)
    return ShiftedProblem("conv-diff-2d-64x64-m16", A, b, sigma, seed_index, 1e-10)


def config1(grid: int = 128, m: int = 64, seed_index: int = 0) -> ShiftedProblem:
    """3-D 7-point convection-diffusion on a 128^3 grid (n=2,097,152,
    14,581,760 nnz): diagonal 6, off-diagonals -1 +- 0.2, b uniform in [-1,1]
    from seed 1, 64 shifts log-spaced in [1e-3, 1], tolerance 1e-10."""
    A = conv_diff_3d(grid, grid, grid, diag=6.0, conv=0.2)
    b = uniform(1, 1, A.n, -1.0, 1.0)
    sigma = logspace_shifts(m, 1e-3, 1.0)
    return ShiftedProblem(f"conv-diff-3d-{grid}^3-m{m}", A, b, sigma, seed_index, 1e-10)


def config2(n: int = 1_000_000, m: int = 128) -> ShiftedProblem:
    """Irregular rows: power-law row lengths (mean 30, cap 2000), n=1e6,
    b uniform from seed 2, 128 shifts uniform in [0, 0.5], tolerance 1e-10,
    seed = smallest shift."""
    A = zipf_matrix(n, 30.0, 2000, seed=2)
    b = uniform(2, 1, n, -1.0, 1.0)
    sigma = uniform(2, 2, m, 0.0, 0.5)
    return ShiftedProblem(f"zipf-n{n}-m{m}", A, b, sigma, int(np.argmin(sigma)), 1e-10)


def config3(count: int = 256, seed: int = 3, n_lo: int = 1000, n_hi: int = 200_000):
    """Batched workload: count independent problems, n log-uniform in
    [n_lo, n_hi] (rounded to a square or cubic grid), alternating 2-D 5-point
    and 3-D 7-point convection-diffusion with per-problem convection in
    [0, 0.4], 8..64 shifts log-spaced in [1e-3, 1], b uniform in [-1,1]."""
    un = uniform(seed, 1, count)
    uc = uniform(seed, 2, count, 0.0, 0.4)
    um = uniform(seed, 3, count)
    probs = []
    for p in range(count):
        n_target = math.exp(math.log(n_lo) + (math.log(n_hi) - math.log(n_lo)) * float(un[p]))
        m = 8 + int(float(um[p]) * 57)  # 8..64
        conv = float(uc[p])
        if p % 2 == 0:
            side = max(4, int(round(math.sqrt(n_target))))
            A = conv_diff_2d(side, side, diag=4.0 + 0.01, conv=conv)
            name = f"batch{p}-2d-{side}^2-m{m}"
        else:
            side = max(4, int(round(n_target ** (1.0 / 3.0))))
            A = conv_diff_3d(side, side, side, diag=6.0, conv=conv)
            name = f"batch{p}-3d-{side}^3-m{m}"
        b = uniform(seed, 1000 + p, A.n, -1.0, 1.0)
        sigma = logspace_shifts(m, 1e-3, 1.0)
        probs.append(ShiftedProblem(name, A, b, sigma, 0, 1e-10))
    return probs


def config4(grid: int = 64, m: int = 32) -> ShiftedProblem:
    """Adjoint workload: 3-D 7-point convection-diffusion on 64^3
    (n=262,144), 32 shifts log-spaced in [1e-2, 1]; loss weights w and
    direction c uniform in [0,1] from seed 4 (in extra)."""
    A = conv_diff_3d(grid, grid, grid, diag=6.0, conv=0.2)
    b = uniform(4, 1, A.n, -1.0, 1.0)
    sigma = logspace_shifts(m, 1e-2, 1.0)
    w = uniform(4, 2, m, 0.0, 1.0)
    c = uniform(4, 3, A.n, 0.0, 1.0)
    return ShiftedProblem(f"conv-diff-3d-{grid}^3-m{m}-adjoint", A, b, sigma, 0, 1e-10,
                          extra={"w": w, "c": c})
