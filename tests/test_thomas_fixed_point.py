"""The multipliers of the z Thomas solve reach a fixed point and keep it.

The solve (src/solver.f90:90-135, coefficient tables src/initsolver.f90:17-56)
computes, per (x-mode, y-mode) column, d(1) = c/(b(1) + lambda) and then
d(k) = c * (1 / ((b + lambda) - a d(k-1))) for k = 2 .. nz-1, where a, b, c
are the same in every interior row.  The HIP solve relies on two facts about
the rounded sequence:

- once d(k) == d(k-1) for some k >= 2, every later row repeats it, so z and d
  of later rows need no division and d need not be stored again (by itself
  this follows from the rows sharing one map; the test checks it anyway);
- in most columns that happens within a few dozen rows.

Its byte model (bench.py's thomas_d_rows, and ipg_substep_bytes, which repeats
the same count in C) counts the rows up to the last change.  Here the full
recurrence is run in numpy over every column at the benchmark grids, with no
early exit. Checks:

- the sequence is monotone;
- the first repeat ends the changes;
- the count matches thomas_d_rows.

A scalar Python loop on one small grid checks the vectorised recurrence.
"""
import math

import numpy as np
import pytest

import bench
from interparts_mpich_amd.inputs import CONFIGS


def eigenvalues(n, dinv):
    """lambda per folded wavenumber 0 .. n/2 (src/initsolver.f90:41-56)."""
    f = np.arange(1, n // 2, dtype=np.float64)
    return np.concatenate(([0.0], -4.0 * dinv * dinv * np.sin(f * np.pi / n) ** 2, [-4.0 * dinv * dinv]))


def multipliers(case):
    """d(k), k = 1 .. nz-1, for every (x-mode, |y-mode|) column: array [nz-1, ncol]."""
    dxi, dyi, dzi = case.itot / case.lx, case.jtot / case.ly, case.ktot / case.lz
    xy = (eigenvalues(case.itot, dxi)[:, None] + eigenvalues(case.jtot, dyi)[None, :]).ravel()
    a = c = dzi * dzi
    b = -(a + c)
    K = case.ktot - 1
    d = np.empty((K, xy.size))
    d[0] = c * (1.0 / ((b + a) + xy))   # row 1: Neumann b(1) = b + a
    for k in range(1, K):
        d[k] = c * (1.0 / ((b + xy) - a * d[k - 1]))
    return d, xy


@pytest.mark.parametrize("name", ["t24x16x32", "ch64_8sph", "ch256_1000sph"])
def test_multipliers_repeat_from_their_last_change(name):
    case = CONFIGS[name]
    d, _ = multipliers(case)
    K = d.shape[0]
    steps = np.diff(d[1:], axis=0)          # rows 2 .. K: one map
    # monotone in every column (all steps of one sign, or zero)
    assert np.all((steps >= 0).all(axis=0) | (steps <= 0).all(axis=0))
    change = np.vstack([np.ones((1, d.shape[1]), bool), d[1:] != d[:-1]])
    last = K - np.argmax(change[::-1], axis=0)   # 1-based row of the last change
    # the first repeat (rows k >= 2 with d(k) == d(k-1)) ends the changes
    repeat = ~change
    repeat[0] = False
    has = repeat.any(axis=0)
    first = np.argmax(repeat, axis=0) + 1        # 1-based row of the first repeat
    assert np.all(last[has] == first[has] - 1)
    assert np.all(last[~has] == K)
    assert bench.thomas_d_rows(case) == float(last.sum())
    if K >= 255:
        # most columns settle early; only the lowest modes run the whole column
        assert last.mean() < 20
        assert np.mean(last == K) < 0.02


def test_vectorised_recurrence_matches_a_scalar_loop():
    case = CONFIGS["t24x16x32"]
    d, xy = multipliers(case)
    dzi = case.ktot / case.lz
    a = c = dzi * dzi
    b = -(a + c)
    for col in (0, 1, 7, xy.size - 1):
        lam = float(xy[col])
        dk = c * (1.0 / ((b + a) + lam))
        assert dk == d[0, col]
        for k in range(1, case.ktot - 1):
            dk = c * (1.0 / ((b + lam) - a * dk))
            assert dk == d[k, col]
    assert math.isfinite(float(d.sum()))
