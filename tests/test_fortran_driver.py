"""The Fortran host driver (program interpt, interparts_mpich_amd/fortran/)
calling libipg.so through iso_c_binding: seeded start-up, the time loop of
src/main.f90 (fused substep or routine by routine), the dt policy, and the
fld%07d / partstate%07d files, checked against the FP64 oracle."""
import dataclasses
import os
import shutil
import subprocess

import numpy as np
import pytest

from interparts_mpich_amd.inputs import (CONFIGS, POST3D, fld_path, initial_fields, markers_for, place_spheres,
                                         read_fld, read_partstate, read_partvisu, read_post3d, write_case_dir,
                                         write_fld)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "interparts_mpich_amd", "fortran", "interpt")
FLUID = ("U", "V", "W", "PNEW", "C")
PART = ("X", "Y", "Z", "U", "V", "W", "OMX", "OMY", "OMZ", "INTU", "INTV", "INTW", "FXLTOT", "FYLTOT", "FZLTOT",
        "COLFX", "COLFY", "COLFZ")


def rel(a, b):
    return float(np.max(np.abs(a - b))) / max(float(np.max(np.abs(b))), 1e-300)


def run(d, timeout=180):
    print(f"[interpt] {d}", flush=True)
    r = subprocess.run([EXE], cwd=d, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout)
    print(r.stdout[-3000:], flush=True)
    return r


def needs_exe():
    if not os.path.exists(EXE):
        pytest.skip("Fortran driver not built (make -C interparts_mpich_amd/fortran)")


def test_fld_round_trip(tmp_path):
    case = CONFIGS["t24x16x32"]
    u, v, w, p, c = initial_fields(case)
    path = str(tmp_path / "fld0000007")
    write_fld(path, case, {"U": u, "V": v, "W": w, "PNEW": p, "C": c}, 0.25, 7, 1e-3)
    f, info = read_fld(path, case)
    assert info == (0.25, 7.0, 1e-3)
    assert np.array_equal(f["V"], v[1:-1, 1:-1, 1:-1]) and np.array_equal(f["U"], u[1:-1, 1:-1, 1:-1])
    assert os.path.getsize(path) == 8 * (5 * case.itot * case.jtot * case.ktot + 3)


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_driver_fails_loudly_without_a_gpu(tmp_path):
    """No CPU fallback: with no device, ipg_init fails and interpt stops with
    the library's message and a non-zero status."""
    needs_exe()
    d = write_case_dir(str(tmp_path), CONFIGS["t16"], nstep=1)
    r = run(d, timeout=60)
    assert r.returncode != 0
    assert "ipg_init" in r.stdout


@pytest.mark.gpu
def test_driver_start_up_matches_host_inputs(tmp_path):
    """nstep = 0: the fields and spheres the driver seeds itself are the ones
    inputs.py draws, and the start-up (boundaries, sphere integrals) matches
    the oracle."""
    from oracle.pyoracle import Oracle
    needs_exe()
    case = CONFIGS["t32"]
    d = write_case_dir(str(tmp_path), case, nstep=0)
    assert run(d).returncode == 0
    f, info = read_fld(fld_path(d, 0), case)
    fields, spheres = initial_fields(case), place_spheres(case)
    o = Oracle(case).initialize(fields, spheres, markers_for(case))
    for n in FLUID:
        assert np.array_equal(f[n], o.get(n)[1:-1, 1:-1, 1:-1]), n
    assert info[2] == o.dt
    q, _ = read_partstate(os.path.join(d, "partstate0000000"), case)
    for a, b in zip(("X", "Y", "Z"), spheres):
        assert np.array_equal(q[a], b), a
    for n in ("INTU", "INTV", "INTW", "INTOMX", "INTOMY", "INTOMZ"):
        assert rel(q[n], o.pget(n)) <= 1e-13, n


@pytest.mark.gpu
@pytest.mark.parametrize("fused", [True, False])
def test_driver_steps_match_oracle(tmp_path, fused):
    """Three steps with a dt update after every step (ioutchk = 1): the fused
    substep and the routine-by-routine loop (rk1.., eulr2lagr/complagrforces/
    lagr2eulr, bounduvw, fillps, solver2d, correc, intgr_nwtn_eulr) both match
    the oracle to <= 1e-11 relative.  The start-up print of kerneltest
    (src/main.f90:142-145) and the per-pass lines of updtintermediatevel
    (src/forcing.f90:430-468: meanforces, IBM_forcing.out) match it too, to
    the files' eight digits."""
    from oracle.pyoracle import Oracle, load as oracle_load
    needs_exe()
    case = dataclasses.replace(CONFIGS["t32"], ioutchk=1)
    d = write_case_dir(str(tmp_path), case, nstep=3, fused_substep=fused)
    r = run(d)
    assert r.returncode == 0
    kline = [l for l in r.stdout.splitlines() if "Integral over kernel" in l]
    assert len(kline) == 1 and abs(float(kline[0].split("=")[1]) - oracle_load().orc_kerneltest()) <= 1e-14
    o = Oracle(case).initialize(initial_fields(case), place_spheres(case), markers_for(case))
    want = []   # (istep, time, fx, fy, fz) per IBM pass, as the reference appends them
    for _ in range(3):
        # Oracle.step, with updtintermediatevel's values after each substep
        o.istep += 1
        o.time += o.dt
        for rkiter in (1, 2, 3):
            o.substep(rkiter)
            for tot in o.ibm_forces():
                want.append((o.istep, o.time, *tot))
        o.divergence = o.chkdiv()
        o.set_dt(0.75 * o.chkdt())
    want = np.array(want)
    got = np.loadtxt(os.path.join(d, "IBM_forcing.out"), ndmin=2)
    assert got.shape == (27, 5), got.shape
    assert np.array_equal(got[:, 0], want[:, 0])
    assert np.allclose(got[:, 1], want[:, 1], rtol=1e-7, atol=0)
    scale = float(np.max(np.abs(want[:, 2:])))
    print("IBM_forcing.out max |difference|", float(np.max(np.abs(got[:, 2:] - want[:, 2:]))), "of", scale)
    assert np.max(np.abs(got[:, 2:] - want[:, 2:])) <= 1e-7 * scale
    mean = np.loadtxt(os.path.join(d, "meanforces"), ndmin=2)
    assert mean.shape == (27, 5) and np.array_equal(mean[:, 0], want[:, 0]) and np.all(mean[:, 2:] == 0.0)
    f, info = read_fld(fld_path(d, 3), case)
    assert info[1] == 3.0
    for n in FLUID:
        r = rel(f[n], o.get(n)[1:-1, 1:-1, 1:-1])
        print(n, r)
        assert r <= 1e-11, n
    q, _ = read_partstate(os.path.join(d, "partstate0000003"), case)
    for n in PART:
        r = rel(q[n], o.pget(n))
        print(n, r)
        assert r <= 1e-11, n
    rows = np.loadtxt(os.path.join(d, "timestep.txt"), ndmin=2)
    assert rows.shape == (3, 3) and list(rows[:, 0]) == [1.0, 2.0, 3.0]


@pytest.mark.gpu
def test_driver_allpartdata(tmp_path):
    """allpartdata%07d (src/loadd.f90:121-692) as the reference writes it: np
    records of type particle_restart (31 + 11 nqmax reals) of which the file
    holds the first (29 + 11 nqmax) np reals (skip, lines 140-142), x .. intomz,
    colf.., colt.., fxltot.. of each sphere in their slots, idp where the file
    reaches it.  Read back with no partstate%07d present (begin = 2, no step),
    each sphere takes the leading reals of its record in the layout of
    particle_old, as the reference's reader does: positions, velocities and
    integrals unchanged, colfx..coltz from the slots flag1, flag2, colfx,
    colfy, colfz, coltx; the idp lost to the missing tail comes from the
    record's position."""
    from interparts_mpich_amd.inputs import ALLPART_FIELDS, allpartdata_idp_slot, allpartdata_layout, read_allpartdata
    needs_exe()
    case = CONFIGS["t32"]
    d = write_case_dir(str(tmp_path), case, nstep=2)
    assert run(d).returncode == 0
    nq, nrec, nfile = allpartdata_layout(case.np)
    path = os.path.join(d, "allpartdata0000002")
    assert os.path.getsize(path) == 8 * nfile * case.np
    rec = read_allpartdata(path, case)
    q, _ = read_partstate(os.path.join(d, "partstate0000002"), case)
    for m, name in enumerate(ALLPART_FIELDS):
        if name.startswith("FLAG"):
            assert np.all(rec[:, m] == 0.0)
        else:
            assert np.array_equal(rec[:, m], q[name]), name
    ids = rec[:, allpartdata_idp_slot(case.np)]
    kept = ~np.isnan(ids)
    assert np.array_equal(ids[kept], np.arange(1, case.np + 1)[kept]) and not kept[-1]
    # restart from the reference's file alone, with a distinct value in every
    # slot from flag1 to fzltot so that the slots the reader takes show
    os.remove(os.path.join(d, "partstate0000002"))
    rec[:, 18:29] = 1000.0 * np.arange(1, case.np + 1)[:, None] + np.arange(18, 29)[None, :]
    rec.reshape(-1)[:nfile * case.np].astype("<f8").tofile(path)
    write_case_dir(d, case, nstep=2, begin=2)
    assert run(d).returncode == 0
    q2, _ = read_partstate(os.path.join(d, "partstate0000002"), case)
    for name in ALLPART_FIELDS[:18]:
        assert np.array_equal(q2[name], q[name]), name
    for m, name in enumerate(("COLFX", "COLFY", "COLFZ", "COLTX", "COLTY", "COLTZ")):
        assert np.array_equal(q2[name], rec[:, 18 + m]), name
    for name in ("FXLTOT", "FYLTOT", "FZLTOT"):
        assert np.all(q2[name] == 0.0), name


@pytest.mark.gpu
def test_driver_restart_round_trip(tmp_path):
    """begin = 2, nstep = 0 reads fld0000002 / partstate0000002 and writes them
    back unchanged (src/loadd.f90 read and write paths)."""
    needs_exe()
    case = CONFIGS["t32"]
    d = write_case_dir(str(tmp_path), case, nstep=2)
    assert run(d).returncode == 0
    keep = str(tmp_path / "keep")
    os.makedirs(keep)
    for name in ("fld0000002", "partstate0000002"):
        shutil.copy(os.path.join(d, name), os.path.join(keep, name))
    write_case_dir(d, case, nstep=0, begin=2)
    assert run(d).returncode == 0
    a, ia = read_fld(fld_path(d, 2), case)
    b, ib = read_fld(os.path.join(keep, "fld0000002"), case)
    assert ia == ib
    for n in FLUID:
        assert np.array_equal(a[n], b[n]), n
    qa, _ = read_partstate(os.path.join(d, "partstate0000002"), case)
    qb, _ = read_partstate(os.path.join(keep, "partstate0000002"), case)
    for n in qa:
        assert np.array_equal(qa[n], qb[n]), n


@pytest.mark.gpu
def test_driver_output_files(tmp_path):
    """post1d / post3d output (src/main.f90:454-472): info_1dstat.out and
    partvisu%07d.bin (outputpart) every iout1d steps and name%07d_3d.bin every
    iout2d steps, against the oracle."""
    from oracle.pyoracle import Oracle
    needs_exe()
    case = dataclasses.replace(CONFIGS["t32"], ioutchk=1)
    d = write_case_dir(str(tmp_path), case, nstep=2, iout2d=2)
    assert run(d).returncode == 0
    o = Oracle(case).initialize(initial_fields(case), place_spheres(case), markers_for(case))
    for _ in range(2):
        o.step()
    f, _ = read_fld(fld_path(d, 2), case)
    assert np.array_equal(read_post3d(os.path.join(d, "pre0000002_3d.bin"), case), f["PNEW"])
    for n in POST3D:
        a = read_post3d(os.path.join(d, f"{n}0000002_3d.bin"), case)
        b = o.post3d(n)
        scale = max(float(np.max(np.abs(b))), 1e-300)
        assert float(np.max(np.abs(a - b))) <= 1e-11 * scale or not np.any(b), n
    fltot, accel = read_partvisu(os.path.join(d, "partvisu0000002.bin"), case)
    for k, ax in enumerate("XYZ"):
        b = o.pget("RKF" + ax)
        assert float(np.max(np.abs(fltot[:, k] - b))) <= 1e-10 * max(float(np.max(np.abs(b))), 1.0), ax
    for k, ax in enumerate("UVW"):
        b = o.pget("RKD" + ax + "DT")
        assert float(np.max(np.abs(accel[:, k] - b))) <= 1e-10 * max(float(np.max(np.abs(b))), 1.0), ax
    info = np.loadtxt(os.path.join(d, "info_1dstat.out"))
    assert info.shape == (8,) and info[0] == 2.0 and info[3] == case.ktot
    assert abs(info[2] - case.visc) <= 1e-7 * case.visc


@pytest.mark.gpu
@pytest.mark.parametrize("mixed_solve", [32, 64])
def test_driver_mixed_precision(tmp_path, mixed_solve):
    """The Fortran driver with precision = 1 in &RUNCONTROL (the mixed path:
    FP32 RK history; the FP32-stored Poisson solve with one FP64-residual
    refinement, or the FP64 solve): three steps of the 32^3 case with a dt
    update after every step, the fields within 1e-6 of the FP64 oracle's."""
    from oracle.pyoracle import Oracle
    needs_exe()
    case = dataclasses.replace(CONFIGS["t32"], ioutchk=1)
    d = write_case_dir(str(tmp_path), case, nstep=3, precision=1, mixed_solve=mixed_solve)
    r = run(d)
    assert r.returncode == 0
    o = Oracle(case).initialize(initial_fields(case), place_spheres(case), markers_for(case))
    for _ in range(3):
        o.step()
    f, info = read_fld(fld_path(d, 3), case)
    assert info[1] == 3.0
    for n in FLUID:
        a, b = f[n], o.get(n)[1:-1, 1:-1, 1:-1]
        if n == "PNEW":
            # the pressure level is the solve's free constant, which the
            # FP32-stored solve pins at p(kmax) = 0 for the mean mode
            a, b = a - a.mean(), b - b.mean()
        e = rel(a, b)
        print(f"mixed solve {mixed_solve}: {n} {e:.3e}", flush=True)
        assert e <= 1e-6, n
