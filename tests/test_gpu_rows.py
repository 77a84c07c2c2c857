"""GPU parity for paths the seeded cases do not reach, against the FP64 oracle:

- spheres placed by hand at the walls: markers within h/2 of the top wall,
  whose w stencils reach the k1 = kmax+1 plane (src/interp_spread.f90:100 has
  no upper clamp; the plane keeps its spread until the next one), and markers
  near the bottom wall, where the stencils are clamped to k >= 1;
- chkdiv (src/chkdiv.f90:23-39), the max-divergence metric, on identical
  fields and after a step;
- the ensemble grid shapes 96x96x64, 128^3 and 128x128x192 (x/y FFT lengths
  96, 128; nz = 192, where the singular kx = ky = 0 Thomas pivot of
  src/solver.f90:114-125 is not exactly zero);
- ensemble cases advanced concurrently equal each case advanced alone;
- the device error word stops every output entry after a dropped contribution;
- post3d's phase indicator 'sol' (src/phase_indicator.f90:11-559, itype = 1)
  on spheres across the periodic edges and corners and at both walls.
"""
import numpy as np
import pytest

from interparts_mpich_amd.inputs import Case, ensemble_cases, initial_fields, markers_for, place_spheres
from interparts_mpich_amd.driver import Simulation
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

FLUID = ("U", "V", "W", "PNEW", "C")
PART = ("X", "Y", "Z", "U", "V", "W", "INTU", "INTV", "INTW", "FXLTOT", "FYLTOT", "FZLTOT", "COLFZ")


def rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def wall_case():
    return Case(name="walls", itot=32, jtot=32, ktot=32, np=2, diameter_h=8.0, ratiorho=1.5, re_bulk=1000.0, seed=23,
                ioutchk=1)


def wall_spheres(case):
    """Sphere 0 just below the top wall, sphere 1 just above the bottom wall:
    z = lz - R - h/20 puts the top markers at lz - 0.35 h (radfp = R - 0.3 h),
    within h/2 of the wall; z = R + h/20 puts the bottom ones at 0.35 h."""
    h, R = case.lz / case.ktot, case.radius
    return (np.array([0.31, 0.69]) * case.lx, np.array([0.42, 0.58]) * case.ly,
            np.array([case.lz - R - 0.05 * h, R + 0.05 * h]))


def test_wall_spheres_top_plane_matches_oracle():
    case = wall_case()
    fields, spheres, markers = initial_fields(case), wall_spheres(case), markers_for(case)
    print(f"[case {case.name}] spheres at the walls, z = {spheres[2]}", flush=True)
    g = Simulation(case).initialize(fields, spheres, markers)
    o = Oracle(case).initialize(fields, spheres, markers)
    # the routine sequence first: the IBM block alone, with the oracle's k1
    # plane of dwdtf carrying a spread the pass gathers from
    g.rk(1); o.rk(1)
    g.ibm(); o.ibm()
    k1 = o.get("DWDTF")[case.ktot + 2]       # dwdtf(:, :, k1), 2 ghost layers
    print("oracle dwdtf(k1) max", float(np.max(np.abs(k1))), flush=True)
    assert np.max(np.abs(k1)) > 0.0
    for n in ("DUDT", "DVDT", "DWDT"):
        r = rel(g.get(n)[1:-1, 1:-1, 1:-1], o.get(n)[1:-1, 1:-1, 1:-1])
        print("ibm", n, r)
        assert r <= 1e-13, (n, r)
    for n in ("FXLTOT", "FYLTOT", "FZLTOT"):
        r = rel(g.pget(n), o.pget(n))
        print("ibm", n, r)
        assert r <= 1e-11, (n, r)
    # then whole time steps through the fused substep, where the k1 planes
    # carry the last spread of one substep into the next
    g2 = Simulation(case, use_graphs=True).initialize(fields, spheres, markers)
    o2 = Oracle(case).initialize(fields, spheres, markers)
    for s in range(3):
        g2.step(); o2.step()
        for n in FLUID:
            r = rel(g2.get(n), o2.get(n))
            assert r <= (1e-12 if s == 0 else 1e-11), (n, r, s)
        for n in PART:
            r = rel(g2.pget(n), o2.pget(n))
            print(s + 1, n, r)
            assert r <= (1e-12 if s == 0 else 1e-11), (n, r, s)
        top = o2.get("DWDTF")[case.ktot + 2]
        print("step", s + 1, "oracle dwdtf(k1) max", float(np.max(np.abs(top))), flush=True)


@pytest.mark.parametrize("name", ["t32", "ch64_8sph"])
def test_chkdiv_matches_oracle(name):
    from interparts_mpich_amd.inputs import CONFIGS
    case = CONFIGS[name]
    fields, spheres, markers = initial_fields(case), place_spheres(case), markers_for(case)
    g = Simulation(case).initialize(fields, spheres, markers)
    o = Oracle(case).initialize(fields, spheres, markers)
    # the unprojected initial field: identical inputs, the same per-cell
    # expression (bitwise maximum; the total is a sum in another order)
    tg, mg = g.chkdiv()
    to, mo = o.chkdiv()
    assert mg == mo, (mg, mo)
    assert mo > 1.0
    div_abs = mo * case.cells
    assert abs(tg - to) <= 1e-13 * div_abs, (tg, to)
    g.step(); o.step()
    tg, mg = g.chkdiv()
    to, mo = o.chkdiv()
    print(name, "after a step: max|div| gpu", mg, "oracle", mo, "total", tg, to)
    assert mg <= 1e-10 and mo <= 1e-10
    for n in ("U", "V", "W"):
        g.put(n, o.get(n))
    tg, mg = g.chkdiv()
    assert mg == mo, (mg, mo)
    assert abs(tg - to) <= 1e-13 * max(mo * case.cells, 1e-300), (tg, to)


def ensemble_shape(nx, ny, nz, npart=6):
    """An ensemble case of the given grid (inputs.ensemble_cases' parameters),
    with few spheres so that the oracle finishes in seconds."""
    base = [c for c in ensemble_cases(64) if (c.itot, c.jtot, c.ktot) == (nx, ny, nz)][0]
    base.np = npart
    base.name = f"{base.name}_np{npart}"
    return base


@pytest.mark.parametrize("shape", [(96, 96, 64), (128, 128, 128), (128, 128, 192)])
def test_ensemble_grid_one_step_matches_oracle(shape):
    case = ensemble_shape(*shape)
    fields, spheres, markers = initial_fields(case), place_spheres(case), markers_for(case)
    print(f"[case {case.name}] D={case.diameter_h:.3f}h nl={case.nl} Re={case.re_bulk:.1f}", flush=True)
    g = Simulation(case, use_graphs=True).initialize(fields, spheres, markers)
    o = Oracle(case).initialize(fields, spheres, markers)
    g.step(); o.step()
    for n in FLUID:
        a, b = g.get(n), o.get(n)
        if n == "PNEW":
            # the singular mode's pressure level is the reference's arbitrary
            # constant; it must still be reproduced, to rounding
            print("pnew mean gpu", a[1:-1, 1:-1, 1:-1].mean(), "oracle", b[1:-1, 1:-1, 1:-1].mean())
        r = rel(a, b)
        print(case.name, n, r)
        assert r <= 1e-12, (n, r)
    for n in PART:
        r = rel(g.pget(n), o.pget(n))
        print(case.name, n, r)
        assert r <= 1e-12, (n, r)
    _, mg = g.chkdiv()
    assert mg <= 1e-10, mg


def test_concurrent_ensemble_equals_each_case_alone():
    """Cases whose substeps are issued round-robin (each context on its own
    stream, captured graphs, no side stream) end where each ends when run on
    its own (<= 1e-13: the IBM's atomic sums are reordered run to run)."""
    cases = [ensemble_shape(64, 64, 64, 8), ensemble_shape(96, 96, 64, 5), ensemble_shape(128, 128, 128, 4)]
    inputs = [(initial_fields(c), place_spheres(c), markers_for(c)) for c in cases]
    conc = [Simulation(c, use_graphs=True, side_stream=False).initialize(*i) for c, i in zip(cases, inputs)]
    for s in range(6):
        for sim in conc:
            sim.substep(s % 3 + 1)
    for sim in conc:
        sim.sync()
    for c, i, sim in zip(cases, inputs, conc):
        alone = Simulation(c, use_graphs=True).initialize(*i)
        for s in range(6):
            alone.substep(s % 3 + 1)
        alone.sync()
        for n in FLUID:
            r = rel(sim.get(n), alone.get(n))
            assert r <= 1e-13, (c.name, n, r)
        for n in PART:
            r = rel(sim.pget(n), alone.pget(n))
            assert r <= 1e-13, (c.name, n, r)
        alone.close()
        sim.close()


def test_error_word_stops_outputs():
    """A fused substep with more than 48 spheres within collision reach of one
    (set up from the host: 63 centres inside 1.9 R of sphere 0) flags the
    device error word; every entry a driver writes output from returns an
    error afterwards (chkdiv, chkdt, field_get, post3d, particles_get)."""
    case = Case(name="crowd", itot=32, jtot=32, ktot=32, np=64, diameter_h=8.0, re_bulk=1000.0, seed=5)
    R = case.radius
    rng = np.random.default_rng(5)
    d = rng.standard_normal((case.np, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d *= 1.9 * R * np.cbrt(rng.uniform(0.05, 1.0, case.np))[:, None]
    d[0] = 0.0
    xs, ys, zs = (0.5 + d[:, 0]) * case.lx, (0.5 + d[:, 1]) * case.ly, (0.5 + d[:, 2]) * case.lz
    g = Simulation(case, use_graphs=True, side_stream=False).initialize(initial_fields(case), (xs, ys, zs),
                                                                        markers_for(case))
    g.substep(1)
    # the crowd overflows the IBM's neighbour-marker list as well as the
    # collision candidates: either message
    for call in (g.chkdiv, g.chkdt, lambda: g.get("U"), lambda: g.post3d("pre"), lambda: g.pget("U")):
        with pytest.raises(RuntimeError, match="more than"):
            call()


def test_phase_indicator_matches_oracle():
    """'sol' of post3d bit for bit against the oracle, on the seeded 32^3
    placement and on spheres placed across the x and y periodic edges, across
    a corner of the periodic plane and against both walls (images through the
    neighbour flags, boxes clamped at the walls); <= 1e-12 after a step, as
    the positions then differ by rounding."""
    from interparts_mpich_amd.inputs import CONFIGS
    edge = Case(name="edges", itot=32, jtot=32, ktot=32, np=5, diameter_h=8.0, ratiorho=1.5, re_bulk=1000.0, seed=9)
    h, R = edge.lx / edge.itot, edge.radius
    edge_spheres = (np.array([0.2 * R, 0.5, edge.lx - 0.3 * R, 0.3, 0.7]),
                    np.array([0.5, edge.ly - 0.1 * R, 0.4 * R, 0.25, 0.75]),
                    np.array([0.5, 0.4, 0.6, R + 0.05 * h, edge.lz - R - 0.05 * h]))
    for case, spheres in ((CONFIGS["t32"], None), (edge, edge_spheres)):
        spheres = spheres if spheres is not None else place_spheres(case)
        fields, markers = initial_fields(case), markers_for(case)
        print(f"[case {case.name}] sol", flush=True)
        g = Simulation(case).initialize(fields, spheres, markers)
        o = Oracle(case).initialize(fields, spheres, markers)
        a, b = g.post3d("sol"), o.post3d("sol")
        print(case.name, "cells inside", int(np.sum(b == 0.0)), "cut", int(np.sum((b > 0) & (b < 1))))
        assert np.any(b < 1.0)
        assert np.array_equal(a, b)
        g.step(); o.step()
        d = float(np.max(np.abs(g.post3d("sol") - o.post3d("sol"))))
        print(case.name, "after a step", d)
        assert d <= 1e-12, d
        g.close(); o.close()


@pytest.mark.parametrize("graphs", [True, False])
def test_ibm_pass_forces_match_oracle(graphs):
    """updtintermediatevel's values (src/forcing.f90:440-458): the sums over
    the spheres of fxltot, fyltot, fzltot after each of the three IBM passes,
    from the fused substep (graph replays or eager launches), against the
    oracle after every substep of two steps; the last pass's sums equal the
    sums of the per-sphere totals the library keeps."""
    from interparts_mpich_amd.inputs import CONFIGS
    case = CONFIGS["t32"]
    fields, spheres, markers = initial_fields(case), place_spheres(case), markers_for(case)
    g = Simulation(case, use_graphs=graphs).initialize(fields, spheres, markers)
    o = Oracle(case).initialize(fields, spheres, markers)
    for s in range(6):
        rk = s % 3 + 1
        g.substep(rk); o.substep(rk)
        a, b = g.ibm_forces(), o.ibm_forces()
        scale = max(float(np.max(np.abs(b))), 1e-300)
        d = float(np.max(np.abs(a - b)))
        print("substep", s + 1, "pass totals", b[:, 1], "max |difference|", d, "of", scale)
        assert np.all(b != 0.0)
        assert d <= 1e-12 * scale, (s, d, scale)
        for c, n in enumerate(("FXLTOT", "FYLTOT", "FZLTOT")):
            t = g.pget(n)
            assert abs(a[2, c] - sum(float(x) for x in t)) <= 1e-13 * scale
