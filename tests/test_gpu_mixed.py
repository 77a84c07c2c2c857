"""The mixed-precision path (ipg_params.precision = 1) against the FP64 oracle
and the FP64 GPU path, on identical seeded inputs.

Mixed precision keeps u, v, w, pnew, c, dudt..dcdt, the particles, the
sphere integrals and the IBM in FP64 and stores the RK history in FP32.  Its
Poisson solve either stores the spectra and the pressure correction's pieces
in FP32 (the arithmetic stays FP64) and refines with FP64 residuals (solve
32, the default) or is the FP64 path's (solve 64); the tests run both.  The bars, as the metric states them per result:

- max |div u| (src/chkdiv.f90:23-39) <= 1e-10 after every substep;
- velocity L2 relative error against the FP64 oracle <= 1e-6 after the
  64^3 case's 100 steps;
- particle-velocity error against the oracle, printed with the others (held
  here to <= 1e-5 of the largest particle speed).

Comparisons between two mixed runs are held to 1e-6, the FP32 storage's
rounding, not to FP64 rounding.
"""
import numpy as np
import pytest

from interparts_mpich_amd.inputs import CONFIGS, initial_fields, markers_for, place_spheres
from interparts_mpich_amd.driver import Simulation
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

FLUID = ("U", "V", "W", "PNEW", "C")
PVEL = ("U", "V", "W")


def rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def vel_l2(a, b):
    """velocity L2 relative error over the interior: |u_a - u_b| / |u_b|, all
    three components together"""
    num = sum(float(np.sum((a.get(n)[1:-1, 1:-1, 1:-1] - b.get(n)[1:-1, 1:-1, 1:-1]) ** 2)) for n in PVEL)
    den = sum(float(np.sum(b.get(n)[1:-1, 1:-1, 1:-1] ** 2)) for n in PVEL)
    return (num / den) ** 0.5


def part_err(a, b):
    """largest particle-velocity difference over the largest particle speed"""
    num = max(float(np.max(np.abs(a.pget(n) - b.pget(n)))) for n in PVEL)
    den = max(float(np.max(np.abs(b.pget(n)))) for n in PVEL)
    return num / den


def inputs(case):
    print(f"[case {case.name}] {case.itot}x{case.jtot}x{case.ktot} np={case.np} nl={case.nl}", flush=True)
    return initial_fields(case), place_spheres(case), markers_for(case)


SOLVES = [64, 32]


@pytest.mark.parametrize("solve", SOLVES)
def test_mixed_ch64_100_steps_against_fp64_oracle(solve):
    case = CONFIGS["ch64_8sph"]
    fields, spheres, markers = inputs(case)
    m = Simulation(case, use_graphs=True, precision=1, solve=solve).initialize(fields, spheres, markers)
    o = Oracle(case).initialize(fields, spheres, markers)
    worst_div = 0.0
    for s in range(case.steps):
        m.step(); o.step()
        if (s + 1) % case.ioutchk == 0:
            # step() ran chkdiv on the substep's result
            worst_div = max(worst_div, m.divergence[1])
            assert m.divergence[1] <= 1e-10, (s + 1, m.divergence)
        if (s + 1) % case.ioutchk == 0 or s + 1 > case.steps - 10:
            print(f"step {s + 1}: max|div| {m.divergence[1]:.3e}, velocity L2 {vel_l2(m, o):.3e}, "
                  f"particle velocity {part_err(m, o):.3e}, min wall gap / R "
                  f"{float(np.min(np.minimum(o.pget('Z'), case.lz - o.pget('Z')))) / case.radius - 1.0:.4f}",
                  flush=True)
    l2, pe = vel_l2(m, o), part_err(m, o)
    print(f"ch64_8sph mixed (solve {solve}) after {case.steps} steps: max|div| {worst_div:.3e} (worst check), velocity L2 {l2:.3e}, "
          f"particle velocity {pe:.3e}")
    assert l2 <= 1e-6, l2
    assert pe <= 1e-5, pe


@pytest.mark.parametrize("solve", SOLVES)
@pytest.mark.parametrize("name", ["t32", "ch64_8sph"])
def test_mixed_one_step_against_fp64_gpu(name, solve):
    case = CONFIGS[name]
    fields, spheres, markers = inputs(case)
    m = Simulation(case, precision=1, solve=solve).initialize(fields, spheres, markers)
    d = Simulation(case).initialize(fields, spheres, markers)
    m.step(); d.step()
    _, mdiv = m.chkdiv()
    l2, pe = vel_l2(m, d), part_err(m, d)
    print(name, f"solve {solve}", "max|div|", mdiv, "velocity L2", l2, "particle velocity", pe)
    assert mdiv <= 1e-10
    assert l2 <= 1e-7, l2
    assert pe <= 1e-6, pe
    # the pressure level is the solve's free constant, which the FP32-stored
    # solve pins at p(kmax) = 0 for the mean mode: compared without its mean
    pm, pd = m.get("PNEW")[1:-1, 1:-1, 1:-1], d.get("PNEW")[1:-1, 1:-1, 1:-1]
    r = rel(pm - pm.mean(), pd - pd.mean())
    print(name, "pnew (less its mean)", r)
    assert r <= 1e-5, r


@pytest.mark.parametrize("solve", SOLVES)
def test_mixed_split_entries_match_fused(solve):
    """The reference-shaped entries one by one (rk, IBM, bounduvw, fillps,
    solver2d, boundp, correc, pnew_update, intgr_nwtn_eulr) on the mixed path
    equal its fused substep."""
    case = CONFIGS["t32"]
    fields, spheres, markers = inputs(case)
    a = Simulation(case, precision=1, solve=solve).initialize(fields, spheres, markers)
    b = Simulation(case, precision=1, side_stream=False, solve=solve).initialize(fields, spheres, markers)
    for rk in (1, 2, 3):
        a.rk(rk); a.ibm(); a.bounduvw(1); a.fillps(rk); a.solver2d(); a.boundp(1); a.correc(rk); a.pnew_update()
        a.intgr_nwtn_eulr(rk)
        b.substep(rk)
    b.sync()
    _, adiv = a.chkdiv()
    assert adiv <= 1e-10, adiv
    assert vel_l2(a, b) <= 1e-6
    assert part_err(a, b) <= 1e-6
    # p after the fused substep is read as the sum of its pieces
    pa, pb = a.get("P"), b.get("P")
    r = rel(pa - pa[1:-1, 1:-1, 1:-1].mean(), pb - pb[1:-1, 1:-1, 1:-1].mean())
    assert r <= 1e-5, r


@pytest.mark.parametrize("solve", SOLVES)
def test_mixed_graph_replay_matches_eager(solve):
    case = CONFIGS["t32"]
    fields, spheres, markers = inputs(case)
    a = Simulation(case, precision=1, solve=solve).initialize(fields, spheres, markers)
    b = Simulation(case, use_graphs=True, precision=1, solve=solve).initialize(fields, spheres, markers)
    for _ in range(3):
        a.step(); b.step()
        assert vel_l2(b, a) <= 1e-6
        assert part_err(b, a) <= 1e-6
        assert b.divergence[1] <= 1e-10 or b.istep % b.case.ioutchk != 0


def test_mixed_rk_history_is_fp32():
    """The RK history is stored FP32: what is put comes back rounded to FP32."""
    case = CONFIGS["t16"]
    m = Simulation(case, precision=1)
    a = np.random.default_rng(3).standard_normal(m.shape)
    for n in ("DURK", "DVRK", "DWRK", "DCRK"):
        m.put(n, a)
        assert np.array_equal(m.get(n), a.astype(np.float32).astype(np.float64)), n
    m.put("U", a)
    assert np.array_equal(m.get("U"), a)


@pytest.mark.parametrize("name", ["ch256_1000sph", "ch512_8000sph"])
def test_mixed_fullsize_divergence_and_refinement(name):
    """At the headline size and at 512^3: one substep with no refinement leaves
    max |div u| well above 1e-10 (the FP32-stored solve alone), one FP64
    residual refinement brings it under 1e-10; the velocity stays within FP32
    rounding of the FP64 GPU path."""
    case = CONFIGS[name]
    fields, spheres, markers = inputs(case)
    d = Simulation(case).initialize(fields, spheres, markers)
    d.substep(1)
    d.sync()
    ref = {n: d.get(n) for n in PVEL}
    d.close()
    divs = {}
    for nref in (0, 1, 2):
        m = Simulation(case, precision=1, nref=nref).initialize(fields, spheres, markers)
        m.substep(1)
        m.sync()
        divs[nref] = m.chkdiv()[1]
        l2 = (sum(float(np.sum((m.get(n) - ref[n])[1:-1, 1:-1, 1:-1] ** 2)) for n in PVEL) /
              sum(float(np.sum(ref[n][1:-1, 1:-1, 1:-1] ** 2)) for n in PVEL)) ** 0.5
        print(name, f"nref {nref}: max|div| {divs[nref]:.3e}, velocity L2 vs FP64 {l2:.3e}", flush=True)
        m.close()
        if nref > 0:
            assert l2 <= 1e-6, (nref, l2)
    assert divs[0] > 1e-10
    assert divs[1] <= 1e-10, divs
    assert divs[2] <= 1e-10, divs


def test_mixed_fp64_solve_is_the_fp64_solve():
    """With the FP64 solve, the mixed path's fillps and solver2d are the FP64
    path's: from the same dudt, dvdt, dwdt, p bit for bit."""
    case = CONFIGS["ch64_8sph"]
    fields, spheres, markers = inputs(case)
    d = Simulation(case).initialize(fields, spheres, markers)
    m = Simulation(case, precision=1, solve=64).initialize(fields, spheres, markers)
    rng = np.random.default_rng(5)
    for n in ("DUDT", "DVDT", "DWDT"):
        a = rng.standard_normal(d.shape)
        d.put(n, a)
        m.put(n, a)
    for s in (d, m):
        s.bounduvw(1)
        s.fillps(2)
        s.solver2d()
    assert np.array_equal(m.get("P"), d.get("P"))
    d.close()
    m.close()


@pytest.mark.parametrize("graphs", [False, True])
def test_mixed_fp64_solve_first_substep_against_fp64(graphs):
    """With the FP64 solve, the mixed path's first RK substep only writes the
    FP32 history (it reads none) and runs the FP64 path's arithmetic: its
    fields and particles equal the FP64 path's to rounding (the RK stage is a
    kernel compiled apart, for FP32 history stores); the second substep reads
    the history and differs at FP32 rounding.  Also the solve selection, its
    refusals, and switching it between substeps."""
    import ctypes as C
    case = CONFIGS["ch64_8sph"]
    fields, spheres, markers = inputs(case)
    d = Simulation(case, use_graphs=graphs).initialize(fields, spheres, markers)
    m = Simulation(case, use_graphs=graphs, precision=1).initialize(fields, spheres, markers)
    bits = C.c_int()
    assert m.lib.ipg_mixed_solve(m.ctx, C.byref(bits)) == 0 and bits.value == 32   # the default
    assert m.lib.ipg_set_mixed_solve(m.ctx, 64) == 0
    assert m.lib.ipg_mixed_solve(m.ctx, C.byref(bits)) == 0 and bits.value == 64
    assert d.lib.ipg_mixed_solve(d.ctx, C.byref(bits)) == 0 and bits.value == 0
    assert d.lib.ipg_set_mixed_solve(d.ctx, 64) != 0        # not a mixed context
    assert m.lib.ipg_set_mixed_solve(m.ctx, 16) != 0        # not 32 or 64
    d.substep(1); m.substep(1)
    d.sync(); m.sync()
    worst = max(rel(m.get(n)[1:-1, 1:-1, 1:-1], d.get(n)[1:-1, 1:-1, 1:-1]) for n in FLUID)
    pw = max(rel(m.pget(n), d.pget(n)) for n in PVEL + ("X", "Y", "Z"))
    print(f"graphs {graphs}: first substep, fields {worst:.3e}, particles {pw:.3e}", flush=True)
    assert worst <= 1e-12, worst
    assert pw <= 1e-12, pw
    d.substep(2); m.substep(2)
    d.sync(); m.sync()
    l2 = vel_l2(m, d)
    print(f"graphs {graphs}: second substep velocity L2 vs FP64 {l2:.3e}", flush=True)
    assert l2 <= 1e-7, l2
    # switching the solve between substeps: 32 from here on, then back
    assert m.lib.ipg_set_mixed_solve(m.ctx, 32) == 0
    assert m.lib.ipg_mixed_solve(m.ctx, C.byref(bits)) == 0 and bits.value == 32
    d.substep(3); m.substep(3)
    d.sync(); m.sync()
    assert m.chkdiv()[1] <= 1e-10
    assert vel_l2(m, d) <= 1e-6
    assert m.lib.ipg_set_mixed_solve(m.ctx, 64) == 0
    d.substep(1); m.substep(1)
    d.sync(); m.sync()
    assert m.chkdiv()[1] <= 1e-10
    assert vel_l2(m, d) <= 1e-6
