"""The launch shapes of the sphere integrals and of the Poisson solve leave the
arithmetic unchanged.

The sphere integrals (src/intgr_over_sphere.f90:12-625) run one block per
sphere or, on a grid smaller than the spheres, as blocks striding over them
(by default one per CU when the update runs beside the RK stage and there are
more spheres than CUs); each sphere's sums are formed by one block the same
way either way.  The block queues its cut cells with LDS atomics, so the order
of their terms, and the last bits of an integral, vary from run to run with
either grid (as the IBM spreads' do): the bar is rounding, <= 1e-13 relative.
"""
import os

import numpy as np
import pytest

from interparts_mpich_amd.inputs import CONFIGS, Case, initial_fields, markers_for, place_spheres
from interparts_mpich_amd.driver import Simulation
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "PNEW", "C")
PARTS = ("X", "Y", "Z", "U", "V", "W", "OMX", "OMY", "OMZ", "INTU", "INTV", "INTW", "INTOMX", "INTOMY", "INTOMZ",
         "FXLTOT", "FYLTOT", "FZLTOT")
INTS = ("INTU", "INTV", "INTW", "INTOMX", "INTOMY", "INTOMZ")


def rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def worst(a, b):
    w = {n: rel(a.get(n), b.get(n)) for n in FIELDS}
    w.update({"p" + n: rel(a.pget(n), b.pget(n)) for n in PARTS})
    return w


def make(case, inputs, sphere_blocks=None, **kw):
    """A context; sphere_blocks sets its sphere-integral grid at init (the
    IPG_SPH_GRID setting, read by ipg_init only)."""
    old = os.environ.get("IPG_SPH_GRID")
    if sphere_blocks is not None:
        os.environ["IPG_SPH_GRID"] = str(sphere_blocks)
    try:
        return Simulation(case, **kw).initialize(*inputs)
    finally:
        if old is None:
            os.environ.pop("IPG_SPH_GRID", None)
        else:
            os.environ["IPG_SPH_GRID"] = old


def test_sphere_integrals_on_a_striding_grid():
    """3 blocks over 8 spheres (and 5, not dividing them): the integrals at
    initialisation, and three substeps (graphs, side stream), agree with one
    block per sphere to rounding, and with the oracle."""
    case = CONFIGS["ch64_8sph"]
    inputs = (initial_fields(case), place_spheres(case), markers_for(case))
    print(f"[case {case.name}] np={case.np}", flush=True)
    a = make(case, inputs, use_graphs=True)
    o = Oracle(case).initialize(*inputs)
    for blocks in (3, 5):
        b = make(case, inputs, sphere_blocks=blocks, use_graphs=True)
        assert b.tuning()[1] == blocks
        for n in INTS:
            r = rel(b.pget(n), a.pget(n))
            assert r <= 1e-13, (blocks, n, r)
        b.close()
    b = make(case, inputs, sphere_blocks=3, use_graphs=True)
    for rk in (1, 2, 3):
        a.substep(rk); b.substep(rk); o.substep(rk)
    a.sync(); b.sync()
    for n, r in worst(b, a).items():
        assert r <= 1e-13, ("striding vs one block per sphere", n, r)
    for n in FIELDS:
        r = rel(b.get(n), o.get(n))
        assert r <= 1e-12, ("vs oracle", n, r)
    for n in PARTS:
        r = rel(b.pget(n), o.pget(n))
        assert r <= 1e-12, ("vs oracle", n, r)
    a.close(); b.close()


# more spheres than the MI355X's 256 CUs, so that the default grid strides
MANY = Case(name="many_128_300sph", itot=128, jtot=128, ktot=128, np=300, diameter_h=6.0, ratiorho=1.5,
            re_bulk=1000.0, seed=5)


def test_default_striding_grid_matches_one_block_per_sphere():
    """300 spheres: the default grid (one block per CU, with the side stream
    on) against one block per sphere, over two steps with the RK stage tuned."""
    case = MANY
    inputs = (initial_fields(case), place_spheres(case), markers_for(case))
    print(f"[case {case.name}] np={case.np} nl={case.nl}", flush=True)
    a = make(case, inputs, sphere_blocks=0, use_graphs=True)
    b = make(case, inputs, use_graphs=True)
    kb = b.autotune()
    rk_kb, blocks = b.tuning()
    print(f"[launch shapes] rk z planes {kb}, sphere-integral blocks {blocks}", flush=True)
    assert rk_kb == kb and kb in (8, 16, 32, 64)
    assert a.tuning()[1] == 0 and 0 < blocks < case.np
    for n in INTS:
        r = rel(b.pget(n), a.pget(n))
        assert r <= 1e-13, (n, r)
    for q in range(6):
        rk = q % 3 + 1
        a.substep(rk); b.substep(rk)
    a.sync(); b.sync()
    for n, r in worst(b, a).items():
        assert r <= 1e-12, (n, r)
    a.close(); b.close()


def test_poisson_in_two_mode_ranges_matches_one():
    """The Poisson solve's y passes and Thomas solve in two x-mode ranges on
    two streams (the default once the side stream exists) against one range
    on one stream (IPG_POISSON_SPLIT=0): each mode's arithmetic is the same,
    so three steps agree to the IBM spreads' rounding (FP64: 1e-13; mixed,
    where such differences can flip an FP32 rounding: 1e-6, as two mixed runs
    are held)."""
    case = CONFIGS["ch64_8sph"]
    inputs = (initial_fields(case), place_spheres(case), markers_for(case))
    for precision in (0, 1):
        print(f"[case {case.name}] precision {precision}", flush=True)
        a = make(case, inputs, use_graphs=True, precision=precision)
        old = os.environ.get("IPG_POISSON_SPLIT")
        os.environ["IPG_POISSON_SPLIT"] = "0"
        try:
            b = make(case, inputs, use_graphs=True, precision=precision)
        finally:
            if old is None:
                os.environ.pop("IPG_POISSON_SPLIT", None)
            else:
                os.environ["IPG_POISSON_SPLIT"] = old
        for q in range(9):
            rk = q % 3 + 1
            a.substep(rk); b.substep(rk)
        a.sync(); b.sync()
        for n, r in worst(a, b).items():
            assert r <= (1e-13 if precision == 0 else 1e-6), (precision, n, r)
        a.close(); b.close()
