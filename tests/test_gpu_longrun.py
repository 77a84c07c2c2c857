"""Long runs of the 256^3 channel with 1000 spheres on the GPU (the headline
case; the reference's own run is 1000 steps):

- 100 steps (300 substeps, dt from chkdt every 10 steps) against a digest of
  the FP64 oracle's state after 10, 50 and 100 steps, computed in the build
  container by tests/golden/make_ch256_digest.py: the particle state, the
  plane means and 4096 sample cells of u, v, w, pnew and c, dt and chkdiv,
  to <= 1e-8 (the IBM's atomic sums differ from the oracle's order by
  rounding, and the flow amplifies it);
- 850 steps: the device error word stays clear (no neighbour, marker or
  collision list overflows while the spheres settle on the bottom wall),
  max |div u| <= 1e-10 at every check (src/chkdiv.f90:23-39), the particle
  state finite and every centre inside the periodic x, y extent.  The spheres
  (density ratio 1.5) settle into the bottom wall, to 0.18 R by step 850:
  the soft-sphere wall contact is tuned for a contact time of
  Nstretch dt_estim = 0.4 (src/init.f90:137-160), long against the settling
  speed, so no bound on z is asserted beyond finiteness.  Near step 890 of
  this case the flow next to a sphere sunk to 0.16 R diverges (the run's
  velocity reaches 10^3 by step 900), and the IBM then reports a stencil
  outside its box: the test stops short of that.
"""
import os

import numpy as np
import pytest

from interparts_mpich_amd.driver import Simulation
from interparts_mpich_amd.inputs import CONFIGS, initial_fields, markers_for, place_spheres

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DIGEST = os.path.join(HERE, "golden", "ch256_1000sph_digest.npz")
PART = ("X", "Y", "Z", "U", "V", "W", "OMX", "OMY", "OMZ", "INTU", "INTV", "INTW", "FXLTOT", "FYLTOT", "FZLTOT")
FIELDS = ("U", "V", "W", "PNEW", "C")


def case_and_inputs():
    case = CONFIGS["ch256_1000sph"]
    return case, (initial_fields(case), place_spheres(case), markers_for(case))


def scaled_err(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def test_ch256_100_steps_against_oracle_digest():
    if not os.path.exists(DIGEST):
        pytest.skip("digest not generated (tests/golden/make_ch256_digest.py)")
    ref = np.load(DIGEST)
    case, inputs = case_and_inputs()
    g = Simulation(case, use_graphs=True).initialize(*inputs)
    cells = (ref["cells_k"], ref["cells_j"], ref["cells_i"])
    steps = [int(s) for s in ref["steps"]]
    worst = 0.0
    for s in range(1, steps[-1] + 1):
        g.step()
        if s not in steps:
            continue
        errs = {}
        for n in PART:
            errs["p_" + n] = scaled_err(g.pget(n), ref[f"s{s}_p_{n}"])
        for n in FIELDS:
            a = g.get(n)
            scale = max(float(np.max(np.abs(a))), 1e-300)
            errs["mean_" + n] = float(np.max(np.abs(a[1:-1, 1:-1, 1:-1].mean(axis=(1, 2)) - ref[f"s{s}_mean_{n}"]))) / scale
            errs["cell_" + n] = float(np.max(np.abs(a[cells] - ref[f"s{s}_cell_{n}"]))) / scale
        errs["dt"] = abs(g.dt - float(ref[f"s{s}_dt"][0])) / float(ref[f"s{s}_dt"][0])
        tot, mx = g.chkdiv()
        big = max(errs, key=errs.get)
        print(f"step {s}: worst {big} {errs[big]:.3e}; max|div| {mx:.3e} (oracle {ref[f's{s}_chkdiv'][1]:.3e})", flush=True)
        assert mx <= 1e-10
        worst = max(worst, errs[big])
        assert errs[big] <= 1e-8, (s, big, errs[big])
    print("worst scaled error over the recorded steps", worst)


def test_ch256_850_steps_error_word_and_divergence():
    case, inputs = case_and_inputs()
    g = Simulation(case, use_graphs=True).initialize(*inputs)
    R = case.radius
    worst = 0.0
    for s in range(1, 851):
        g.step()   # chkdiv every ioutchk steps: it fails on a set error word
        if s % case.ioutchk == 0:
            worst = max(worst, g.divergence[1])
            assert g.divergence[1] <= 1e-10, (s, g.divergence)
        if s % 100 == 0:
            z = g.pget("Z")
            print(f"step {s}: t = {g.time:.4f}, max|div| so far {worst:.3e}, z / R in [{z.min() / R:.3f}, "
                  f"{z.max() / R:.3f}], max |w_p| {np.max(np.abs(g.pget('W'))):.4f}", flush=True)
            for n in ("Z", "U", "V", "W", "OMX", "OMY", "OMZ"):
                assert np.all(np.isfinite(g.pget(n))), n
            for n in ("X", "Y"):
                a = g.pget(n)
                assert np.all(np.isfinite(a)) and np.all(a >= 0.0) and np.all(a <= case.lx)
    g.sync()
