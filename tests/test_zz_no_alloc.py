"""ipg_chkdiv_noalloc (the divergence check on caller-owned device arrays,
with no allocation of its own) against ipg_chkdiv on the same fields: the
maximum bit for bit (the same per-cell expression, src/chkdiv.f90:23-39), the
total to 1e-12 of sum |div| (the partial sums are grouped differently).  Two
calls in a row, the second on a smaller grid; then a grid beyond its 512^3
maximum is refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from interparts_mpich_amd.inputs import CONFIGS, initial_fields, markers_for, place_spheres
from interparts_mpich_amd.driver import Simulation

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip_runtime():
    """The HIP runtime libipg.so itself uses: the test allocates the inputs."""
    h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    return h


def to_device(h, arrays):
    ptrs = []
    for a in arrays:
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = C.c_void_p()
        assert h.hipMalloc(C.byref(p), a.nbytes) == 0
        assert h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        ptrs.append(p)
    return ptrs


def noalloc(lib, ptrs, case, dims=None):
    nx, ny, nz = dims or (case.itot, case.jtot, case.ktot)
    tot, mx = C.c_double(), C.c_double()
    rc = lib.ipg_chkdiv_noalloc(ptrs[0], ptrs[1], ptrs[2], nx, ny, nz, case.lx, case.ly, case.lz, C.byref(tot),
                                C.byref(mx))
    return rc, tot.value, mx.value


def compare(sim, h, case):
    u, v, w = sim.get("U"), sim.get("V"), sim.get("W")
    ptrs = to_device(h, (u, v, w))
    try:
        rc, tot, mx = noalloc(sim.lib, ptrs, case)
        assert rc == 0, sim.lib.ipg_last_error()
        ref_tot, ref_mx = sim.chkdiv()
        print(case.name, "noalloc", tot, mx, "chkdiv", ref_tot, ref_mx, flush=True)
        assert mx == ref_mx
        assert abs(tot - ref_tot) <= 1e-12 * max(ref_mx * case.cells, 1e-300)
    finally:
        for p in ptrs:
            h.hipFree(p)
    return ptrs


def test_chkdiv_noalloc_matches_chkdiv():
    h = hip_runtime()
    for name in ("ch64_8sph", "t16"):
        case = CONFIGS[name]
        fields = initial_fields(case)
        sim = Simulation(case).initialize(fields, place_spheres(case), markers_for(case))
        # the initial field, not divergence-free (max |div| of order 10), then
        # the projected one after a time step
        compare(sim, h, case)
        sim.step()
        compare(sim, h, case)
        sim.close()
    # a grid beyond the maximum is refused before any launch
    case = CONFIGS["t16"]
    sim = Simulation(case).initialize(initial_fields(case), place_spheres(case), markers_for(case))
    ptrs = to_device(h, (sim.get("U"), sim.get("V"), sim.get("W")))
    rc, _, _ = noalloc(sim.lib, ptrs, case, dims=(1024, 512, 512))
    assert rc == 2
    assert b"exceed" in sim.lib.ipg_last_error()
    for p in ptrs:
        h.hipFree(p)
    sim.close()
