"""Pins of the oracle's phase indicator 'sol' (orc_phase_sol: post3d's call
of src/phase_indicator.f90:11-559 with itype = 1, src/output.f90:1068-1074),
each independent of the restated loop:

 * classification: with the eight corners of each pressure cell measured
   against every periodic image of every sphere (numpy, all nine images), a
   cell with all corners of some image inside the radius has sol = 0, a cell
   with no corner inside any image has sol = 1, and every other cell lies
   strictly between.  A wrong image slot, a missing image, a box bound cut
   short or a wrong corner set fails it;
 * a cut cell's value: a sphere centred on the middle of a cell's bottom face
   with radius between the two corner layers' distances a < R < b gives
   sol = 1 - (R - a)/(b - a) (four corners at R - a > 0, four at R - b < 0);
 * volume: the sum of (1 - sol) dV over the grid converges to the spheres'
   volume at second order (measured -2.0 % at 4 cells radius, -0.5 % at 8);
 * translation: moving a sphere by whole cells across the periodic boundary
   rolls the field by as many cells.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest

from interparts_mpich_amd.inputs import Case
from oracle.pyoracle import Oracle


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def sol_of(case, xs, ys, zs):
    o = Oracle(case)
    x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (xs, ys, zs))
    o.lib.orc_init_particles(o.s, _ptr(x), _ptr(y), _ptr(z))
    s = o.post3d("sol")
    o.close()
    return s


def classify(case, xs, ys, zs):
    """-1: no corner inside any image; 1: all corners inside some image; 0: else."""
    nx, ny, nz = case.itot, case.jtot, case.ktot
    dx, dy, dz = case.lx / nx, case.ly / ny, case.lz / nz
    R = case.radius
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    all_in = np.zeros(k.shape, bool)
    any_in = np.zeros(k.shape, bool)
    for X, Y, Z in zip(xs, ys, zs):
        for sx, sy in itertools.product((-1, 0, 1), repeat=2):
            Xi, Yi = X + sx * case.lx, Y + sy * case.ly
            inside = []
            for ci, cj, ck in itertools.product((0, 1), repeat=3):
                d2 = ((i - 1 + ci) * dx - Xi) ** 2 + ((j - 1 + cj) * dy - Yi) ** 2 + ((k - 1 + ck) * dz - Z) ** 2
                inside.append(d2 < R * R)
            inside = np.array(inside)
            all_in |= inside.all(axis=0)
            any_in |= inside.any(axis=0)
    return np.where(all_in, 1, np.where(any_in, 0, -1))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_sol_classifies_cells_by_their_corners(seed):
    rng = np.random.default_rng(seed)
    case = Case(name="phase", itot=32, jtot=24, ktot=20, lx=1.0, ly=0.75, lz=0.625, np=4, diameter_h=7.0)
    R = case.radius
    xs = rng.uniform(0, case.lx, 4)
    ys = rng.uniform(0, case.ly, 4)
    zs = rng.uniform(R, case.lz - R, 4)
    # one sphere across the x = 0 plane, one across the corner x = lx, y = 0
    xs[0] = 0.3 * R
    xs[1], ys[1] = case.lx - 0.4 * R, 0.1 * R
    s = sol_of(case, xs, ys, zs)
    cls = classify(case, xs, ys, zs)
    assert np.all(s[cls == 1] == 0.0)
    assert np.all(s[cls == -1] == 1.0)
    cut = s[cls == 0]
    assert cut.size > 100 and np.all((cut > 0.0) & (cut < 1.0))


def test_cut_cell_value_closed_form():
    case = Case(name="phase1", itot=16, jtot=16, ktot=16, np=1, diameter_h=2.4)
    h = case.lx / case.itot
    i0, j0, k0 = 7, 9, 6                           # the cell [i0-1, i0] x [j0-1, j0] x [k0-1, k0] (cells of h)
    X, Y, Z = (i0 - 0.5) * h, (j0 - 0.5) * h, (k0 - 1) * h   # centre of its bottom face
    a, b = math.sqrt(0.5) * h, math.sqrt(1.5) * h  # bottom and top corner distances
    R = case.radius
    assert a < R < b
    s = sol_of(case, [X], [Y], [Z])
    got = s[k0 - 1, j0 - 1, i0 - 1]
    assert abs(got - (1.0 - (R - a) / (b - a))) <= 1e-14, got


def test_sol_volume_converges_to_the_sphere_volume():
    """Relative volume error of the indicator at radius 4 and 8 cells (one
    sphere across the periodic boundary, two inside): small, one-signed and of
    second order (the error drops about fourfold when the radius in cells
    doubles)."""
    err = {}
    for rc in (4.0, 8.0):
        case = Case(name="phasev", itot=64, jtot=64, ktot=64, np=3, diameter_h=2 * rc)
        R = case.radius
        xs, ys, zs = np.array([0.01, 0.45, 0.8]), np.array([0.3, 0.95, 0.6]), np.array([0.3, 0.5, 0.72])
        s = sol_of(case, xs, ys, zs)
        vol = float(np.sum(1.0 - s)) / case.itot ** 3
        want = 3 * 4.0 / 3.0 * math.pi * R ** 3
        err[rc] = (vol - want) / want
    print(err)
    assert -0.03 < err[4.0] < -0.01 and -0.008 < err[8.0] < -0.003, err
    assert 3.0 < err[4.0] / err[8.0] < 5.0, err


def test_sol_rolls_with_the_sphere():
    case = Case(name="phaser", itot=24, jtot=20, ktot=16, lx=1.0, ly=20 / 24, lz=16 / 24, np=1, diameter_h=6.0)
    h = case.lx / case.itot
    X, Y, Z = 0.37 * h, 0.81 * h + 0.5 * case.ly, 0.5 * case.lz + 0.13 * h
    base = sol_of(case, [X], [Y], [Z])
    for m, n in ((3, 0), (-5, 2), (11, -7)):
        Xm = (X + m * h) % case.lx
        Yn = (Y + n * h) % case.ly
        got = sol_of(case, [Xm], [Yn], [Z])
        want = np.roll(np.roll(base, m, axis=2), n, axis=1)
        assert np.max(np.abs(got - want)) <= 1e-12, (m, n)
