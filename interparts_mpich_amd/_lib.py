"""ctypes binding of libipg.so (include/ipg.h).

The library is the product: every call goes to the hand-written gfx950 HIP
kernels.  There is no CPU fallback; if the shared library has not been built
the import of this module raises.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libipg.so")

# field selectors (include/ipg.h)
FIELDS = ["U", "V", "W", "PNEW", "C", "DUDT", "DVDT", "DWDT", "DCDT", "DURK", "DVRK", "DWRK", "DCRK", "P"]
F = {n: i for i, n in enumerate(FIELDS)}
# particle quantity selectors (include/ipg.h)
PARTS = ["X", "Y", "Z", "THETA", "PHI", "U", "V", "W", "OMX", "OMY", "OMZ", "OMTHETA", "INTU", "INTV", "INTW",
         "INTOMX", "INTOMY", "INTOMZ", "COLFX", "COLFY", "COLFZ", "COLTX", "COLTY", "COLTZ", "FXLTOT", "FYLTOT",
         "FZLTOT", "TORQX", "TORQY", "TORQZ", "TORQTHETA", "RKFX", "RKFY", "RKFZ", "RKDUDT", "RKDVDT", "RKDWDT"]
PQ = {n: i for i, n in enumerate(PARTS)}

# every symbol include/ipg.h declares
EXPORTS = ["ipg_init", "ipg_finalize", "ipg_last_error", "ipg_nl_formula", "ipg_kerneltest", "ipg_field_put",
           "ipg_field_get",
           "ipg_particles_put", "ipg_particles_get", "ipg_set_dt", "ipg_get_dt", "ipg_initsolver",
           "ipg_initparticles", "ipg_coordsfp", "ipg_rk", "ipg_ibm", "ipg_ibm_begin", "ipg_eulr2lagr",
           "ipg_lagrforces", "ipg_lagr2eulr", "ipg_ibm_forces", "ipg_uniform", "ipg_bounduvw", "ipg_boundp",
           "ipg_boundscal", "ipg_fillps", "ipg_solver2d", "ipg_correc", "ipg_pnew_update",
           "ipg_intgr_over_sphere", "ipg_intgr_nwtn_eulr", "ipg_substep", "ipg_chkdt", "ipg_chkdiv", "ipg_chkdiv_noalloc",
           "ipg_post3d",
           "ipg_sync", "ipg_use_graphs", "ipg_use_side_stream", "ipg_set_mixed_solve", "ipg_mixed_solve", "ipg_set_refinement", "ipg_autotune", "ipg_tuning",
           "ipg_substep_bytes",
           "ipg_substep_phases", "ipg_batch_create", "ipg_batch_destroy", "ipg_batch_substep", "ipg_batch_sync",
           "ipg_batch_use_graphs", "ipg_batch_set_groups", "ipg_batch_groups", "ipg_batch_set_carry", "ipg_batch_carry",
           "ipg_batch_autotune", "ipg_batch_launches", "ipg_batch_phases"]


class IpgParams(C.Structure):
    _fields_ = [("itot", C.c_int), ("jtot", C.c_int), ("ktot", C.c_int), ("np", C.c_int), ("nl", C.c_int),
                ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double), ("visc", C.c_double),
                ("pran", C.c_double), ("beta", C.c_double), ("radius", C.c_double), ("ratiorho", C.c_double),
                ("particle_rotation", C.c_int), ("bc_uvw_top", C.c_int), ("bc_uvw_bot", C.c_int),
                ("bc_c_top", C.c_int), ("bc_c_bot", C.c_int), ("bc_c_top_val", C.c_double),
                ("bc_c_bot_val", C.c_double), ("precision", C.c_int)]


class IpgError(RuntimeError):
    pass


_lib = None


def load():
    """Load libipg.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IpgError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                       f"(python -m interparts_mpich_amd.build)")
    lib = C.CDLL(LIB_PATH)
    vp, ip, dp, d = C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_double
    sig = {
        "ipg_init": ([C.POINTER(IpgParams), C.POINTER(vp)], ip),
        "ipg_finalize": ([vp], ip),
        "ipg_last_error": ([], C.c_char_p),
        "ipg_nl_formula": ([ip, d, d], ip),
        "ipg_kerneltest": ([dp], ip),
        "ipg_field_put": ([vp, ip, vp], ip),
        "ipg_field_get": ([vp, ip, vp], ip),
        "ipg_particles_put": ([vp, ip, vp], ip),
        "ipg_particles_get": ([vp, ip, vp], ip),
        "ipg_set_dt": ([vp, d], ip),
        "ipg_get_dt": ([vp, dp], ip),
        "ipg_initsolver": ([vp], ip),
        "ipg_initparticles": ([vp, vp, vp, vp], ip),
        "ipg_coordsfp": ([vp, vp, vp, d], ip),
        "ipg_rk": ([vp, ip], ip),
        "ipg_ibm": ([vp], ip),
        "ipg_ibm_begin": ([vp], ip),
        "ipg_eulr2lagr": ([vp, ip], ip),
        "ipg_lagrforces": ([vp, ip], ip),
        "ipg_lagr2eulr": ([vp, ip], ip),
        "ipg_ibm_forces": ([vp, dp], ip),
        "ipg_uniform": ([C.c_ulonglong, C.c_ulonglong, C.c_ulonglong, C.c_longlong, vp], ip),
        "ipg_bounduvw": ([vp, ip], ip),
        "ipg_boundp": ([vp, ip], ip),
        "ipg_boundscal": ([vp], ip),
        "ipg_fillps": ([vp, ip], ip),
        "ipg_solver2d": ([vp], ip),
        "ipg_correc": ([vp, ip], ip),
        "ipg_pnew_update": ([vp], ip),
        "ipg_intgr_over_sphere": ([vp, ip], ip),
        "ipg_intgr_nwtn_eulr": ([vp, ip], ip),
        "ipg_substep": ([vp, ip], ip),
        "ipg_chkdt": ([vp, dp], ip),
        "ipg_chkdiv": ([vp, dp, dp], ip),
        "ipg_chkdiv_noalloc": ([vp, vp, vp, ip, ip, ip, d, d, d, dp, dp], ip),
        "ipg_post3d": ([vp, ip, vp], ip),
        "ipg_sync": ([vp], ip),
        "ipg_use_graphs": ([vp, ip], ip),
        "ipg_use_side_stream": ([vp, ip], ip),
        "ipg_set_refinement": ([vp, ip], ip),
        "ipg_set_mixed_solve": ([vp, ip], ip),
        "ipg_mixed_solve": ([vp, C.POINTER(C.c_int)], ip),
        "ipg_autotune": ([vp, C.POINTER(C.c_int)], ip),
        "ipg_tuning": ([vp, C.POINTER(C.c_int), C.POINTER(C.c_int)], ip),
        "ipg_substep_bytes": ([vp, dp], ip),
        "ipg_substep_phases": ([vp, ip, ip, dp, C.POINTER(C.c_char_p), C.POINTER(ip)], ip),
        "ipg_batch_create": ([C.POINTER(vp), ip, C.POINTER(vp)], ip),
        "ipg_batch_destroy": ([vp], ip),
        "ipg_batch_substep": ([vp, ip], ip),
        "ipg_batch_sync": ([vp], ip),
        "ipg_batch_use_graphs": ([vp, ip], ip),
        "ipg_batch_set_groups": ([vp, ip], ip),
        "ipg_batch_groups": ([vp, C.POINTER(ip)], ip),
        "ipg_batch_set_carry": ([vp, ip], ip),
        "ipg_batch_carry": ([vp, C.POINTER(ip)], ip),
        "ipg_batch_autotune": ([vp, C.POINTER(ip)], ip),
        "ipg_batch_launches": ([vp, ip, C.POINTER(ip), C.POINTER(ip)], ip),
        "ipg_batch_phases": ([vp, ip, ip, dp, C.POINTER(C.c_char_p), C.POINTER(ip)], ip),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def check(rc, what=""):
    if rc != 0:
        msg = load().ipg_last_error().decode(errors="replace")
        raise IpgError(f"{what}: {msg} (code {rc})")
