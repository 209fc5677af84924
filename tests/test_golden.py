"""The oracle against its committed fixtures (tests/golden/c0_oracle.npz,
written by tests/golden/make_golden.py): any change to the restatement that
moves a bit of these results fails here."""
import os

import numpy as np
import pytest

import pyoracle as O
from hbicg import problems as P

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "c0_oracle.npz")


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


@pytest.mark.parametrize("seed", [0, 8, 15])
def test_oracle_reproduces_fixture(gold, seed):
    p = P.config0()
    r = O.solve(p.A, p.b, p.sigma, seed, tol=p.tol)
    assert r.k == int(gold[f"k_{seed}"])
    assert np.array_equal(r.stop_iter, gold[f"stop_{seed}"])
    assert np.array_equal(np.array(r.switches, dtype=np.int32).reshape(-1, 2), gold[f"switch_{seed}"])
    assert np.array_equal(r.x[:, ::41], gold[f"xs_{seed}"])
    assert np.array_equal(r.x.sum(axis=1), gold[f"xsum_{seed}"])
    if seed:
        assert len(r.switches) >= 1   # the fixture exercises seed switching
