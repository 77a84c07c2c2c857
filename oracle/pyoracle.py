"""ctypes wrapper of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The interface mirrors
interparts_mpich_amd.driver.Simulation so a test can run the same case on
both and compare.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# IPG_ORACLE_LIB: another build of the oracle (tests/test_oracle_mutations.py
# plants mistakes in copies of it and checks that the pins fail)
LIB = os.environ.get("IPG_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")
REF = os.path.join(_HERE, "_ref", "libref.so")

FIELDS = ["U", "V", "W", "PNEW", "C", "DUDT", "DVDT", "DWDT", "DCDT", "DURK", "DVRK", "DWRK", "DCRK", "P",
          "DUDTF", "DVDTF", "DWDTF", "DUDTOLD", "DVDTOLD", "DWDTOLD"]
PARTS = ["X", "Y", "Z", "THETA", "PHI", "U", "V", "W", "OMX", "OMY", "OMZ", "OMTHETA", "INTU", "INTV", "INTW",
         "INTOMX", "INTOMY", "INTOMZ", "COLFX", "COLFY", "COLFZ", "COLTX", "COLTY", "COLTZ", "FXLTOT", "FYLTOT",
         "FZLTOT", "TORQX", "TORQY", "TORQZ", "TORQTHETA", "RKFX", "RKFY", "RKFZ", "RKDUDT", "RKDVDT", "RKDWDT"]
MARKERS = ["XFP", "YFP", "ZFP", "UL", "VL", "WL", "DUDTL", "DVDTL", "DWDTL", "FXL", "FYL", "FZL"]

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    vp, ip, d, dp = C.c_void_p, C.c_int, C.c_double, C.POINTER(C.c_double)
    sig = {
        "orc_create": ([ip, ip, ip, d, d, d, d, d, d, d, d, ip, ip, ip, ip, ip, ip, ip, d, d], vp),
        "orc_destroy": ([vp], None),
        "orc_field": ([vp, ip], C.POINTER(C.c_double)),
        "orc_particle": ([vp, ip], C.POINTER(C.c_double)),
        "orc_marker": ([vp, ip], C.POINTER(C.c_double)),
        "orc_nla": ([vp], C.POINTER(C.c_int)),
        "orc_nbflags": ([vp], C.POINTER(C.c_int)),
        "orc_set_dt": ([vp, d], None),
        "orc_get_dt": ([vp], d),
        "orc_scalar": ([vp, ip], d),
        "orc_kernel": ([d], d),
        "orc_kerneltest": ([], d),
        "orc_rfftf": ([ip, vp], None),
        "orc_rfftb": ([ip, vp], None),
        "orc_nl_formula": ([ip, d, d], ip),
        "orc_bounduvw": ([vp, ip], None),
        "orc_boundp": ([vp, ip], None),
        "orc_boundscal": ([vp], None),
        "orc_rk": ([vp, ip], None),
        "orc_eulr2lagr": ([vp], None),
        "orc_lagrforces": ([vp, ip], None),
        "orc_lagr2eulr": ([vp], None),
        "orc_ibm": ([vp], None),
        "orc_fillps": ([vp, ip], None),
        "orc_initsolver": ([vp], None),
        "orc_solver2d": ([vp], None),
        "orc_correc": ([vp, ip], None),
        "orc_pnew_update": ([vp], None),
        "orc_chkdiv": ([vp, dp, dp], None),
        "orc_ibm_forces": ([vp, vp], None),
        "orc_post3d": ([vp, C.c_int, vp], None),
        "orc_chkdt": ([vp, dp], None),
        "orc_update_nb_flags": ([vp], None),
        "orc_coordsfp": ([vp, vp, vp, d], None),
        "orc_intgr_over_sphere": ([vp, ip], None),
        "orc_intgr_nwtn_eulr": ([vp, ip], None),
        "orc_init_particles": ([vp, vp, vp, vp], None),
        "orc_substep": ([vp, ip], None),
        "orc_step": ([vp, ip, ip], None),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, case):
        self.case = case
        self.lib = load()
        self.s = self.lib.orc_create(case.itot, case.jtot, case.ktot, case.lx, case.ly, case.lz, case.visc, case.pran,
                                     case.beta, case.radius, case.ratiorho, case.np,
                                     case.nl if case.np > 0 else 0, int(case.particle_rotation), case.bc_uvw_top,
                                     case.bc_uvw_bot, case.bc_c_top, case.bc_c_bot, case.bc_c_top_val,
                                     case.bc_c_bot_val)
        self.shape = (case.ktot + 2, case.jtot + 2, case.itot + 2)
        self.fshape = (case.ktot + 4, case.jtot + 4, case.itot + 4)
        self.time = 0.0
        self.istep = 0

    def close(self):
        if self.s:
            self.lib.orc_destroy(self.s)
            self.s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self, name):
        fid = FIELDS.index(name)
        shp = self.fshape if name.endswith("F") else self.shape
        n = int(np.prod(shp))
        return np.ctypeslib.as_array(self.lib.orc_field(self.s, fid), shape=(n,)).reshape(shp)

    def get(self, name):
        return self.view(name).copy()

    def put(self, name, arr):
        self.view(name)[...] = arr

    def pview(self, name):
        return np.ctypeslib.as_array(self.lib.orc_particle(self.s, PARTS.index(name)), shape=(max(self.case.np, 1),))[:self.case.np]

    def pget(self, name):
        return self.pview(name).copy()

    def pput(self, name, arr):
        self.pview(name)[...] = arr

    def mview(self, name):
        n = max(self.case.np, 1) * max(self.case.nl, 1)
        return np.ctypeslib.as_array(self.lib.orc_marker(self.s, MARKERS.index(name)), shape=(n,))

    def nla(self):
        return np.ctypeslib.as_array(self.lib.orc_nla(self.s), shape=(max(self.case.np, 1),))[:self.case.np].copy()

    def nbflags(self):
        return np.ctypeslib.as_array(self.lib.orc_nbflags(self.s), shape=(max(self.case.np, 1) * 9,)).reshape(-1, 9)[:self.case.np].copy()

    def scalar(self, i):
        return self.lib.orc_scalar(self.s, i)

    def set_dt(self, dt):
        self.dt = float(dt)
        self.lib.orc_set_dt(self.s, self.dt)

    # reference procedures
    def rk(self, rkiter): self.lib.orc_rk(self.s, rkiter)
    def ibm(self): self.lib.orc_ibm(self.s)
    def bounduvw(self, which=0): self.lib.orc_bounduvw(self.s, which)
    def boundp(self, which=0): self.lib.orc_boundp(self.s, which)
    def boundscal(self): self.lib.orc_boundscal(self.s)
    def fillps(self, rkiter): self.lib.orc_fillps(self.s, rkiter)
    def initsolver(self): self.lib.orc_initsolver(self.s)
    def solver2d(self): self.lib.orc_solver2d(self.s)
    def correc(self, rkiter): self.lib.orc_correc(self.s, rkiter)
    def pnew_update(self): self.lib.orc_pnew_update(self.s)
    def intgr_over_sphere(self, t): self.lib.orc_intgr_over_sphere(self.s, t)
    def intgr_nwtn_eulr(self, rkiter): self.lib.orc_intgr_nwtn_eulr(self.s, rkiter)
    def substep(self, rkiter): self.lib.orc_substep(self.s, rkiter)

    def chkdt(self):
        d = C.c_double()
        self.lib.orc_chkdt(self.s, C.byref(d))
        return d.value

    def post3d(self, name):
        q = ("pre", "vez", "con", "vox", "voy", "voz", "sol").index(name)
        a = np.empty((self.case.ktot, self.case.jtot, self.case.itot))
        self.lib.orc_post3d(self.s, q, a.ctypes.data_as(C.c_void_p))
        return a

    def ibm_forces(self):
        """(3 passes, 3 components): forcextot.. of updtintermediatevel."""
        a = np.zeros(9)
        self.lib.orc_ibm_forces(self.s, a.ctypes.data_as(C.c_void_p))
        return a.reshape(3, 3)

    def chkdiv(self):
        a, b = C.c_double(), C.c_double()
        self.lib.orc_chkdiv(self.s, C.byref(a), C.byref(b))
        return a.value, b.value

    def initialize(self, fields, spheres=None, markers=None):
        """src/main.f90:98-227 with begin = 0 (same sequence as
        interparts_mpich_amd.driver.Simulation.initialize)."""
        case = self.case
        u, v, w, p, c = fields
        for name, a in (("U", u), ("V", v), ("W", w), ("PNEW", p), ("C", c)):
            self.put(name, a)
        if case.np > 0:
            x, y, z = (np.ascontiguousarray(a, dtype=np.float64) for a in spheres)
            self.lib.orc_init_particles(self.s, _ptr(x), _ptr(y), _ptr(z))
        self.bounduvw(0)
        self.boundscal()
        self.set_dt(0.5 * self.chkdt())
        self.initsolver()
        self.bounduvw(0)
        self.boundscal()
        self.boundp(0)
        self.divergence = self.chkdiv()
        if case.np > 0:
            theta, phi, radfp = markers
            theta = np.ascontiguousarray(theta, dtype=np.float64)
            phi = np.ascontiguousarray(phi, dtype=np.float64)
            self.lib.orc_coordsfp(self.s, _ptr(theta), _ptr(phi), radfp)
            self.intgr_over_sphere(1)
            self.volume = self.pget("INTU").copy()
            self.intgr_over_sphere(2)
            self.intgr_over_sphere(3)
        return self

    def step(self):
        self.istep += 1
        self.time += self.dt
        for rkiter in (1, 2, 3):
            self.substep(rkiter)
        if self.case.ioutchk > 0 and self.istep % self.case.ioutchk == 0:
            self.divergence = self.chkdiv()
            self.set_dt(0.75 * self.chkdt())


def ref_lib():
    """The reference's own kernel.f90 and fft.f compiled into oracle/_ref
    (None when the reference tree was absent at build time)."""
    if not os.path.exists(REF):
        return None
    lib = C.CDLL(REF)
    lib._QMmod_kernelPkernel.restype = C.c_double
    lib._QMmod_kernelPkernel.argtypes = [C.POINTER(C.c_double)]
    return lib
