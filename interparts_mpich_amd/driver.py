"""Host driver of the GPU substep, mirroring program interpt
(src/main.f90:30-541) for one rank holding the whole domain.

The method names follow the reference's module procedures (rk1/rk2/rk3,
ibm = the eulr2lagr/complagrforces/updtlagrforces/lagr2eulr block, bounduvw,
boundp, boundscal, fillps, solver2d, correc, intgr_over_sphere,
intgr_nwtn_eulr, chkdt, chkdiv); each forwards to the C-ABI of libipg.so.
Fields stay resident on the GPU between calls; host copies are made only by
get()/put().
"""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import F, PQ, check, load
from .inputs import Case, initial_fields, place_spheres, markers_for


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Simulation:
    def __init__(self, case: Case, use_graphs=False, side_stream=True, precision=0, nref=None, solve=None):
        """precision 0: the FP64 path; 1: mixed precision (ipg_params.precision).
        solve (mixed only): 32, the Poisson solve with FP32-stored spectra and
        nref FP64-residual refinement sweeps (the library's default), or 64,
        the FP64 path's solve; giving nref alone selects 32."""
        self.case = case
        self.precision = precision
        self.lib = load()
        prm = _lib.IpgParams(itot=case.itot, jtot=case.jtot, ktot=case.ktot, np=case.np,
                             nl=case.nl if case.np > 0 else 0, lx=case.lx, ly=case.ly, lz=case.lz,
                             visc=case.visc, pran=case.pran, beta=case.beta, radius=case.radius,
                             ratiorho=case.ratiorho, particle_rotation=int(case.particle_rotation),
                             bc_uvw_top=case.bc_uvw_top, bc_uvw_bot=case.bc_uvw_bot, bc_c_top=case.bc_c_top,
                             bc_c_bot=case.bc_c_bot, bc_c_top_val=case.bc_c_top_val,
                             bc_c_bot_val=case.bc_c_bot_val, precision=precision)
        self.ctx = C.c_void_p()
        check(self.lib.ipg_init(C.byref(prm), C.byref(self.ctx)), "ipg_init")
        if use_graphs:
            check(self.lib.ipg_use_graphs(self.ctx, 1), "ipg_use_graphs")
        if solve is None and nref is not None:
            solve = 32
        if solve is not None:
            check(self.lib.ipg_set_mixed_solve(self.ctx, solve), "ipg_set_mixed_solve")
        if nref is not None:
            check(self.lib.ipg_set_refinement(self.ctx, nref), "ipg_set_refinement")
        if not side_stream:
            check(self.lib.ipg_use_side_stream(self.ctx, 0), "ipg_use_side_stream")
        self.shape = (case.ktot + 2, case.jtot + 2, case.itot + 2)
        self.time = 0.0
        self.dt = 0.0
        self.istep = 0

    def close(self):
        if self.ctx:
            self.lib.ipg_finalize(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- transfers ---------------------------------------------------------
    def put(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == self.shape
        check(self.lib.ipg_field_put(self.ctx, F[name], _ptr(a)), f"put {name}")

    def get(self, name):
        a = np.empty(self.shape)
        check(self.lib.ipg_field_get(self.ctx, F[name], _ptr(a)), f"get {name}")
        return a

    def pput(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        assert a.shape == (self.case.np,)
        check(self.lib.ipg_particles_put(self.ctx, PQ[name], _ptr(a)), f"pput {name}")

    def pget(self, name):
        a = np.empty(self.case.np)
        check(self.lib.ipg_particles_get(self.ctx, PQ[name], _ptr(a)), f"pget {name}")
        return a

    def set_dt(self, dt):
        self.dt = float(dt)
        check(self.lib.ipg_set_dt(self.ctx, C.c_double(self.dt)), "set_dt")

    # ---- reference procedures ----------------------------------------------
    def rk(self, rkiter):
        check(self.lib.ipg_rk(self.ctx, rkiter), "rk")

    def rk1(self):
        self.rk(1)

    def rk2(self):
        self.rk(2)

    def rk3(self):
        self.rk(3)

    def ibm(self):
        check(self.lib.ipg_ibm(self.ctx), "ibm")

    def ibm_begin(self):
        check(self.lib.ipg_ibm_begin(self.ctx), "ibm_begin")

    def eulr2lagr(self, ibmiter):
        check(self.lib.ipg_eulr2lagr(self.ctx, ibmiter), "eulr2lagr")

    def lagrforces(self, ibmiter):
        check(self.lib.ipg_lagrforces(self.ctx, ibmiter), "lagrforces")

    def lagr2eulr(self, ibmiter):
        check(self.lib.ipg_lagr2eulr(self.ctx, ibmiter), "lagr2eulr")

    def bounduvw(self, which=0):
        check(self.lib.ipg_bounduvw(self.ctx, which), "bounduvw")

    def boundp(self, which=0):
        check(self.lib.ipg_boundp(self.ctx, which), "boundp")

    def boundscal(self):
        check(self.lib.ipg_boundscal(self.ctx), "boundscal")

    def fillps(self, rkiter):
        check(self.lib.ipg_fillps(self.ctx, rkiter), "fillps")

    def initsolver(self):
        check(self.lib.ipg_initsolver(self.ctx), "initsolver")

    def solver2d(self):
        check(self.lib.ipg_solver2d(self.ctx), "solver2d")

    def correc(self, rkiter):
        check(self.lib.ipg_correc(self.ctx, rkiter), "correc")

    def pnew_update(self):
        check(self.lib.ipg_pnew_update(self.ctx), "pnew_update")

    def intgr_over_sphere(self, typ):
        check(self.lib.ipg_intgr_over_sphere(self.ctx, typ), "intgr_over_sphere")

    def intgr_nwtn_eulr(self, rkiter):
        check(self.lib.ipg_intgr_nwtn_eulr(self.ctx, rkiter), "intgr_nwtn_eulr")

    def chkdt(self):
        d = C.c_double()
        check(self.lib.ipg_chkdt(self.ctx, C.byref(d)), "chkdt")
        return d.value

    def chkdiv(self):
        a, b = C.c_double(), C.c_double()
        check(self.lib.ipg_chkdiv(self.ctx, C.byref(a), C.byref(b)), "chkdiv")
        return a.value, b.value

    def ibm_forces(self):
        """updtintermediatevel's IBM_forcing.out values (src/forcing.f90:440-458)
        of the last IBM block: (3 passes, 3 components) sums over the spheres
        of fxltot, fyltot, fzltot."""
        a = np.zeros(9)
        check(self.lib.ipg_ibm_forces(self.ctx, a.ctypes.data_as(C.POINTER(C.c_double))), "ibm_forces")
        return a.reshape(3, 3)

    def post3d(self, name):
        """One post3d quantity (pre vez con vox voy voz sol) as a (ktot, jtot, itot) array."""
        q = ("pre", "vez", "con", "vox", "voy", "voz", "sol").index(name)
        a = np.empty((self.case.ktot, self.case.jtot, self.case.itot))
        check(self.lib.ipg_post3d(self.ctx, q, _ptr(a)), f"post3d {name}")
        return a

    def substep(self, rkiter):
        check(self.lib.ipg_substep(self.ctx, rkiter), "substep")

    def autotune(self):
        """Pick the RK stage's z-block length on this grid (ipg_autotune)."""
        kb = C.c_int()
        check(self.lib.ipg_autotune(self.ctx, C.byref(kb)), "autotune")
        return kb.value

    def tuning(self):
        """(RK z planes per workgroup, sphere-integral blocks) in use
        (ipg_tuning; 0: chosen per grid, one block per sphere)."""
        kb, sb = C.c_int(), C.c_int()
        check(self.lib.ipg_tuning(self.ctx, C.byref(kb), C.byref(sb)), "tuning")
        return kb.value, sb.value

    def sync(self):
        check(self.lib.ipg_sync(self.ctx), "sync")

    def substep_bytes(self):
        b = C.c_double()
        check(self.lib.ipg_substep_bytes(self.ctx, C.byref(b)), "substep_bytes")
        return b.value

    # ---- program interpt ---------------------------------------------------
    def initialize(self, fields=None, spheres=None, markers=None):
        """src/main.f90:98-227 with begin = 0: initial conditions, particles,
        boundary conditions, dt = 0.5 dtmax, solver and marker set-up, the
        initial sphere integrals."""
        case = self.case
        u, v, w, p, c = fields if fields is not None else initial_fields(case)
        for name, a in (("U", u), ("V", v), ("W", w), ("PNEW", p), ("C", c)):
            self.put(name, a)
        if case.np > 0:
            x, y, z = spheres if spheres is not None else place_spheres(case)
            x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in (x, y, z))
            check(self.lib.ipg_initparticles(self.ctx, _ptr(x), _ptr(y), _ptr(z)), "initparticles")
        self.time = 0.0
        self.bounduvw(0)
        self.boundscal()
        self.set_dt(0.5 * self.chkdt())
        self.initsolver()
        self.bounduvw(0)
        self.boundscal()
        self.boundp(0)
        self.divergence = self.chkdiv()
        if case.np > 0:
            theta, phi, radfp = markers if markers is not None else markers_for(case)
            theta = np.ascontiguousarray(theta, dtype=np.float64)
            phi = np.ascontiguousarray(phi, dtype=np.float64)
            check(self.lib.ipg_coordsfp(self.ctx, _ptr(theta), _ptr(phi), C.c_double(radfp)), "coordsfp")
            self.intgr_over_sphere(1)
            self.volume = self.pget("INTU").copy()
            self.intgr_over_sphere(2)
            self.intgr_over_sphere(3)
        return self

    def step(self):
        """One time step: src/main.f90:253-441 (three substeps, then every
        ioutchk steps chkdiv, chkdt and dt = 0.75 dtmax)."""
        self.istep += 1
        self.time += self.dt
        for rkiter in (1, 2, 3):
            self.substep(rkiter)
        if self.case.ioutchk > 0 and self.istep % self.case.ioutchk == 0:
            self.divergence = self.chkdiv()
            self.set_dt(0.75 * self.chkdt())
        else:
            self.sync()


class Batch:
    """A batch of Simulations advanced through each substep together
    (ipg_batch_*): one merged launch per kernel and stage for all cases, one
    captured graph per substep.  The simulations stay usable on their own."""

    def __init__(self, sims, use_graphs=True):
        self.sims = list(sims)
        self.lib = load()
        arr = (C.c_void_p * len(self.sims))(*[s.ctx for s in self.sims])
        self.handle = C.c_void_p()
        check(self.lib.ipg_batch_create(arr, len(self.sims), C.byref(self.handle)), "batch_create")
        if not use_graphs:
            check(self.lib.ipg_batch_use_graphs(self.handle, 0), "batch_use_graphs")

    def close(self):
        if self.handle:
            self.lib.ipg_batch_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def substep(self, rkiter):
        check(self.lib.ipg_batch_substep(self.handle, rkiter), "batch_substep")

    def set_groups(self, n):
        """Split the cases into n stream groups (ipg_batch_set_groups)."""
        check(self.lib.ipg_batch_set_groups(self.handle, n), "batch_set_groups")

    def groups(self):
        n = C.c_int()
        check(self.lib.ipg_batch_groups(self.handle, C.byref(n)), "batch_groups")
        return n.value

    def set_carry(self, on):
        """Carry each substep's Newton-Euler update beside the next RK stage
        (ipg_batch_set_carry)."""
        check(self.lib.ipg_batch_set_carry(self.handle, 1 if on else 0), "batch_set_carry")

    def carry(self):
        n = C.c_int()
        check(self.lib.ipg_batch_carry(self.handle, C.byref(n)), "batch_carry")
        return bool(n.value)

    def sync(self):
        check(self.lib.ipg_batch_sync(self.handle), "batch_sync")

    def autotune(self):
        """The batch's tuner (ipg_batch_autotune); returns the substeps it ran."""
        n = C.c_int()
        check(self.lib.ipg_batch_autotune(self.handle, C.byref(n)), "batch_autotune")
        return n.value

    def launches(self, rkiter):
        """(merged launches, persistent ones) of the next substep of stage rkiter."""
        n, p = C.c_int(), C.c_int()
        check(self.lib.ipg_batch_launches(self.handle, rkiter, C.byref(n), C.byref(p)), "batch_launches")
        return n.value, p.value

    def phases(self, rkiter, maxn=128):
        """One eager substep with per-merged-launch times: [(stage, ms)]."""
        ms = (C.c_double * maxn)()
        names = (C.c_char_p * maxn)()
        n = C.c_int()
        check(self.lib.ipg_batch_phases(self.handle, rkiter, maxn, ms, names, C.byref(n)), "batch_phases")
        return [(names[q].decode(), ms[q]) for q in range(n.value)]
