"""CPU-side checks of the product library: libipg.so loads and exports every
symbol include/ipg.h declares (no device calls)."""
import ctypes
import os
import re

from interparts_mpich_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, "include", "ipg.h")).read()
    return sorted(set(re.findall(r"\b(ipg_[a-z0-9_]+)\s*\(", text)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_nl_formula_matches_host_model():
    from interparts_mpich_amd.inputs import nl_formula, CONFIGS
    lib = _lib.load()
    for case in CONFIGS.values():
        assert lib.ipg_nl_formula(case.itot, case.lx, case.radius) == nl_formula(case.itot, case.lx, case.radius)
    # the marker counts of the benchmark configurations (radius retracted by 0.3h)
    assert nl_formula(64, 1.0, 4 / 64) == 173
    assert nl_formula(256, 1.0, 6 / 256) == 409
    assert nl_formula(512, 1.0, 8 / 512) == 746


def test_uniform_matches_host_generator():
    """ipg_uniform (C) and inputs.uniform (numpy) draw the same doubles."""
    import numpy as np
    from interparts_mpich_amd.inputs import uniform
    lib = _lib.load()
    for seed, stream, first, count in ((1, 0, 0, 1000), (7, 101, 12345, 257), (2**40 + 3, 200, 2**33, 64)):
        out = np.empty(count)
        assert lib.ipg_uniform(seed, stream, first, count, out.ctypes.data_as(ctypes.c_void_p)) == 0
        ref = uniform(seed, stream, np.arange(first, first + count, dtype=np.uint64))
        assert np.array_equal(out, ref)
    assert lib.ipg_uniform(1, 0, 0, -1, None) != 0


def test_kerneltest_matches_reference():
    """ipg_kerneltest (the start-up print of src/main.f90:142-145, host only):
    the reference's kernel.f90 value, bit for bit (the golden vector written
    by the reference compiled here), and the oracle's."""
    import json
    from oracle.pyoracle import load as oracle_load
    lib = _lib.load()
    s = ctypes.c_double()
    assert lib.ipg_kerneltest(ctypes.byref(s)) == 0
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "ref_vectors.json")))["kerneltest"]
    assert s.value == gold
    assert s.value == oracle_load().orc_kerneltest()
    assert lib.ipg_kerneltest(None) != 0


def test_allpartdata_layout():
    """allpartdata%07d sizes (src/loadd.f90:123-142, src/param.f90:100): the
    record of type particle_restart, 29 reals, 10 history arrays of nqmax,
    qmax, idp and firstc(nqmax), and the skip the reference moves per
    particle, two reals short of it."""
    from interparts_mpich_amd.inputs import ALLPART_FIELDS, allpartdata_idp_slot, allpartdata_layout
    assert len(ALLPART_FIELDS) == 29
    for npart, nq in ((1, 3), (6, 8), (13, 15), (1000, 15)):
        q, nrec, nfile = allpartdata_layout(npart)
        assert q == nq and nrec == 29 + 10 * nq + 2 + nq and nfile == (27 + 10 * nq) + (2 + nq)
        assert allpartdata_idp_slot(npart) == 29 + 10 * nq + 1
