"""The committed golden fixtures (tests/golden/) against the oracle that wrote them.

Regenerating every fixture with the current oracle must reproduce it bit for bit (NaN
positions included), so a change in the oracle's arithmetic cannot pass unnoticed, and the
fixtures the device tests read stay the oracle's. The residual histories must fall and the
stored solutions must satisfy their systems.
"""
import os
import sys

import numpy as np
import pytest

import oracle as orc
from conftest import bits_equal

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

import make_golden  # noqa: E402


@pytest.mark.parametrize("n", [8, 16])
def test_oracle_reproduces_golden_fixtures(n):
    with np.load(os.path.join(HERE, "golden", f"golden_{n}.npz"), allow_pickle=False) as z:
        stored = {k: z[k] for k in z.files}
    fresh = make_golden.generate(n)
    assert sorted(stored) == sorted(fresh)
    for k, v in fresh.items():
        v = np.asarray(v)
        if v.dtype.kind == "f":
            assert bits_equal(stored[k], v), k
        else:
            assert np.array_equal(stored[k], v), k


@pytest.mark.parametrize("n", [8, 16])
def test_golden_solutions_solve_their_systems(n):
    with np.load(os.path.join(HERE, "golden", f"golden_{n}.npz"), allow_pickle=False) as z:
        for name, *_ in make_golden.operators(n):
            rp, col, val = z[f"{name}_rp"], z[f"{name}_col"], z[f"{name}_val"]
            rhs, x, hist = z[f"{name}_rhs"], z[f"{name}_x"], z[f"{name}_hist"]
            res = np.linalg.norm(rhs - orc.spmv(rp, col, val, x))
            assert res <= 2 * z[f"{name}_tol"][0], name
            assert len(hist) == min(3, int(z[f"{name}_it"][0])) and np.all(np.diff(hist) < 0), name
