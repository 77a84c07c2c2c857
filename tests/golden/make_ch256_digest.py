"""Writes tests/golden/ch256_1000sph_digest.npz: a digest of the FP64 oracle's
state after N time steps of the 256^3 channel with 1000 spheres (the seeded
inputs of interparts_mpich_amd.inputs, the case's dt policy: chkdt every
ioutchk steps), for the GPU long-run test (tests/test_gpu_longrun.py).

The digest holds, at each of the recorded steps: the particle state (all
spheres: positions, velocities, angular velocities, integrals), the
plane means over x and y of u, v, w, pnew and c, the values of u, v, w, pnew
and c at 4096 seeded sample cells, dt, and chkdiv's total and maximum.

    OMP_NUM_THREADS=8 python tests/golden/make_ch256_digest.py [steps]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from interparts_mpich_amd.inputs import CONFIGS, initial_fields, markers_for, place_spheres  # noqa: E402
from oracle.pyoracle import Oracle  # noqa: E402

PART = ("X", "Y", "Z", "U", "V", "W", "OMX", "OMY", "OMZ", "INTU", "INTV", "INTW", "FXLTOT", "FYLTOT", "FZLTOT")
FIELDS = ("U", "V", "W", "PNEW", "C")
RECORD = (10, 50, 100)
NSAMPLE = 4096


def sample_cells(case, n=NSAMPLE, seed=20261204):
    """Interior cells (k, j, i) of the ghosted (ktot+2, jtot+2, itot+2) array."""
    rng = np.random.default_rng(seed)
    return (rng.integers(1, case.ktot + 1, n), rng.integers(1, case.jtot + 1, n), rng.integers(1, case.itot + 1, n))


def digest(o, case, cells):
    out = {}
    for n in PART:
        out["p_" + n] = o.pget(n)
    for n in FIELDS:
        a = o.view(n)
        out["mean_" + n] = a[1:-1, 1:-1, 1:-1].mean(axis=(1, 2))
        out["cell_" + n] = a[cells]
    out["dt"] = np.array([o.dt])
    out["chkdiv"] = np.array(o.chkdiv())
    return out


def main():
    steps = int(sys.argv[1]) if len(sys.argv) > 1 else RECORD[-1]
    case = CONFIGS["ch256_1000sph"]
    cells = sample_cells(case)
    o = Oracle(case).initialize(initial_fields(case), place_spheres(case), markers_for(case))
    rec = {"cells_k": cells[0], "cells_j": cells[1], "cells_i": cells[2], "steps": np.array([s for s in RECORD if s <= steps])}
    t0 = time.time()
    for s in range(1, steps + 1):
        o.step()
        if s in RECORD:
            for k, v in digest(o, case, cells).items():
                rec[f"s{s}_{k}"] = v
        print(f"step {s} dt {o.dt:.6e} t {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "ch256_1000sph_digest.npz"), **rec)


if __name__ == "__main__":
    main()
