"""The batched ensemble (ipg_batch_*): cases of uneven grids, sphere counts and
parameters advanced through each substep together, one merged launch per
kernel and stage for all cases, against the same cases advanced alone by
ipg_substep on identical seeded inputs.

A merged launch runs each case's blocks with the case's own arguments, so a
case's arithmetic is that of its own substep; only the IBM spreads, which sum
with LDS atomics in any order, differ from one run to the next.  The bar is
therefore the substep parity bar, <= 1e-12 relative on every field and
particle quantity after a few substeps (FP64); mixed-precision cases round
their RK history and spectra to FP32, where such differences can flip a
rounding, and are held to 1e-6 as two mixed runs are.
"""
import dataclasses

import numpy as np
import pytest

from interparts_mpich_amd.inputs import CONFIGS, Case, ensemble_cases, initial_fields, markers_for, place_spheres
from interparts_mpich_amd.driver import Batch, Simulation

pytestmark = pytest.mark.gpu

FIELDS = ("U", "V", "W", "PNEW", "C", "P")
PARTS = ("X", "Y", "Z", "U", "V", "W", "OMX", "OMY", "OMZ", "FXLTOT", "FYLTOT", "FZLTOT")


def rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def make(case, precision=0, graphs=True, solve=None):
    print(f"[case {case.name}] {case.itot}x{case.jtot}x{case.ktot} np={case.np} precision={precision} solve={solve}",
          flush=True)
    return Simulation(case, use_graphs=graphs, side_stream=False, precision=precision, solve=solve).initialize(
        initial_fields(case), place_spheres(case), markers_for(case) if case.np > 0 else None)


def worst(a, b):
    w = 0.0
    for n in FIELDS:
        w = max(w, rel(a.get(n), b.get(n)))
    if a.case.np > 0:
        for n in PARTS:
            w = max(w, rel(a.pget(n), b.pget(n)))
    return w


# a case without spheres: its substep has no IBM or particle stages, so its
# records are a subsequence of the others'
NOSPH = Case(name="nosph_32x48x24", itot=32, jtot=48, ktot=24, lx=1.0, ly=1.5, lz=0.75, np=0, re_bulk=800.0, seed=11)


def cases_small():
    return [CONFIGS["t16"], CONFIGS["t32"], CONFIGS["t24x16x32"], CONFIGS["ch64_8sph"], NOSPH,
            dataclasses.replace(CONFIGS["t32"], name="t32_seed9", seed=9, re_bulk=700.0)]


@pytest.mark.parametrize("graphs,carry", [(True, True), (False, True), (True, False)])
def test_batch_matches_cases_alone(graphs, carry):
    """With carry on (the default) each substep's Newton-Euler update runs
    beside the next substep's RK stage; off, it closes its substep."""
    cases = cases_small()
    bat = [make(c) for c in cases]
    alone = [make(c) for c in cases]
    b = Batch(bat, use_graphs=graphs)
    b.set_carry(carry)
    for step in range(2):
        for rk in (1, 2, 3):
            b.substep(rk)
            for s in alone:
                s.substep(rk)
        b.sync()
        for s in alone:
            s.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        print(f"[batch vs alone] {x.case.name}: {w:.3e}", flush=True)
        assert w <= 1e-12, (x.case.name, w)
    n, _ = b.launches(1)
    print(f"[batch] {n} merged launches per substep for {len(cases)} cases", flush=True)
    b.close()


@pytest.mark.parametrize("ngroups", [2, 3])
def test_batch_stream_groups(ngroups):
    """The cases split into stream groups whose merged launches run side by
    side in one graph, regrouped between steps."""
    cases = cases_small()
    bat = [make(c) for c in cases]
    alone = [make(c) for c in cases]
    b = Batch(bat)
    b.set_groups(ngroups)
    assert b.groups() == ngroups
    for step in range(2):
        if step == 1:
            b.set_groups(ngroups - 1)
        for rk in (1, 2, 3):
            b.substep(rk)
            for s in alone:
                s.substep(rk)
    b.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        print(f"[batch in {ngroups} groups vs alone] {x.case.name}: {w:.3e}", flush=True)
        assert w <= 1e-12, (x.case.name, w)
    b.close()


def test_batch_with_per_case_checks_between_steps():
    """chkdiv, chkdt and set_dt on each case between batch steps: per-case
    entry points and the batch's merged launches stay ordered."""
    cases = cases_small()[:4]
    bat = [make(c) for c in cases]
    alone = [make(c) for c in cases]
    b = Batch(bat)
    for step in range(3):
        for rk in (1, 2, 3):
            b.substep(rk)
            for s in alone:
                s.substep(rk)
        for x, y in zip(bat, alone):
            dx, dy = x.chkdt(), y.chkdt()
            assert dx == pytest.approx(dy, rel=1e-12)
            x.set_dt(0.75 * dx)
            y.set_dt(0.75 * dx)
            assert x.chkdiv()[1] <= 1e-10
    b.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        print(f"[batch vs alone, dt updates] {x.case.name}: {w:.3e}", flush=True)
        assert w <= 1e-12, (x.case.name, w)
    b.close()


def test_batch_interleaved_with_cases_stepping_alone():
    """A batch step, then every case steps on its own (ipg_substep on its own
    stream), then a batch step again: each kind of work waits for the other."""
    cases = cases_small()[:4]
    bat = [make(c) for c in cases]
    alone = [make(c) for c in cases]
    b = Batch(bat)
    for step in range(3):
        for rk in (1, 2, 3):
            if step == 1:
                for s in bat:
                    s.substep(rk)
            else:
                b.substep(rk)
            for s in alone:
                s.substep(rk)
    b.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        print(f"[batch and own steps vs alone] {x.case.name}: {w:.3e}", flush=True)
        assert w <= 1e-12, (x.case.name, w)
    b.close()


def test_batch_tuner_keeps_results():
    """The tuner times whole RK3 steps (RK launch shapes, persistent grids);
    the steps it runs are ordinary substeps, and the tuned graphs replay the
    same arithmetic."""
    cases = cases_small()[:5]
    bat = [make(c) for c in cases]
    alone = [make(c) for c in cases]
    b = Batch(bat)
    nsub = b.autotune()
    assert nsub % 3 == 0 and nsub > 0
    n, npers = b.launches(1)
    print(f"[tuner] {nsub} substeps; {npers} of {n} merged launches persistent; {b.groups()} stream groups; "
          f"carried updates {b.carry()}", flush=True)
    for q in range(nsub + 6):
        rk = q % 3 + 1
        if q >= nsub:
            b.substep(rk)
        for s in alone:
            s.substep(rk)
    b.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        print(f"[tuned batch vs alone] {x.case.name}: {w:.3e}", flush=True)
        assert w <= 1e-12, (x.case.name, w)
    b.close()


@pytest.mark.parametrize("solve", [64, 32])
def test_batch_mixed_precision(solve):
    cases = cases_small()[:4]
    bat = [make(c, precision=1, solve=solve) for c in cases]
    alone = [make(c, precision=1, solve=solve) for c in cases]
    b = Batch(bat)
    for step in range(2):
        for rk in (1, 2, 3):
            b.substep(rk)
            for s in alone:
                s.substep(rk)
    b.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        _, mdiv = x.chkdiv()
        print(f"[mixed batch vs alone] {x.case.name}: {w:.3e}, max|div| {mdiv:.2e}", flush=True)
        assert w <= 1e-6, (x.case.name, w)
        assert mdiv <= 1e-10
    b.close()


def test_batch_of_fp64_and_mixed_cases():
    """FP64 and mixed-precision cases in one batch, the mixed ones with each
    Poisson solve: the FP64 substep's stages are a subsequence of the mixed
    one's (whose FP32-stored solve adds the refinement solve), and each stage
    merges the cases of each body (FP64 and FP32 spectra apart)."""
    cases = cases_small()[:4]
    precs = [0, 1, 0, 1]
    solves = [None, 64, None, 32]
    bat = [make(c, precision=p, solve=v) for c, p, v in zip(cases, precs, solves)]
    alone = [make(c, precision=p, solve=v) for c, p, v in zip(cases, precs, solves)]
    b = Batch(bat)
    for step in range(2):
        for rk in (1, 2, 3):
            b.substep(rk)
            for s in alone:
                s.substep(rk)
    b.sync()
    n, _ = b.launches(1)
    print(f"[fp64 + mixed batch] {n} merged launches per substep", flush=True)
    for x, y, p in zip(bat, alone, precs):
        w = worst(x, y)
        print(f"[fp64 + mixed batch vs alone] {x.case.name} precision {p}: {w:.3e}", flush=True)
        assert w <= (1e-6 if p else 1e-12), (x.case.name, w)
    b.close()


def test_batch_ensemble_grid_shapes():
    """Cases of the 64-case ensemble's draw (grids 64^3, 96x96x64, 128^3,
    128x128x192, up to 200 spheres): FFT lengths 64, 96 and 128 and 192 z
    planes in one batch."""
    cases = ensemble_cases(64)
    picked, seen = [], set()
    for c in cases:
        shape = (c.itot, c.jtot, c.ktot)
        if shape not in seen or len(picked) < 6:
            picked.append(c)
            seen.add(shape)
        if len(seen) == 4 and len(picked) >= 6:
            break
    assert len(seen) == 4, seen
    bat = [make(c) for c in picked]
    alone = [make(c) for c in picked]
    b = Batch(bat)
    for rk in (1, 2, 3):
        b.substep(rk)
        for s in alone:
            s.substep(rk)
    b.sync()
    for x, y in zip(bat, alone):
        w = worst(x, y)
        print(f"[ensemble batch vs alone] {x.case.name} np={x.case.np}: {w:.3e}", flush=True)
        assert w <= 1e-12, (x.case.name, w)
    b.close()


def test_batch_rejects_cases_out_of_step():
    """Cases at different c/dcdt buffer parities cannot share a merged
    substep: the batch refuses rather than run one of them with the other's
    buffers."""
    a, c = make(CONFIGS["t16"]), make(CONFIGS["t32"])
    a.substep(1)
    b = Batch([a, c])
    with pytest.raises(Exception, match="parit"):
        b.substep(2)
    b.close()
