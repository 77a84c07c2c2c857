"""Seeded case inputs: parameters, initial fields, sphere placement and the
marker file.

The reference seeds its initial noise, sphere placement and marker layout
with Fortran's random_number (src/init.f90:287-305,
src/initparticles.f90:36-47, src/poslfp/lfp.f:55-93), which is compiler
specific.  This module replaces it with a documented counter-based generator
so that every consumer of a case (the GPU path, the CPU oracle, the Fortran
driver) sees bit-identical inputs:

    u(seed, stream, n) = splitmix64(seed*K1 + stream*K2 + n) >> 11, times 2^-53

Initial velocity: the 'poi' profile of src/init.f90:436-450 (v = 6 z (1-z)
with z the cell-centre height over lz, u = w = p = 0 on all cells) plus, on
interior cells, 0.1*(2u-1) per component keyed by the global cell index.
Spheres: rejection sampling with centre gap >= D + h (minimum image in x, y)
and wall gap >= R + h.  Markers: a golden-angle spiral of nl points, written
in the lagrangianforcepoints2 text format of src/poslfp/lfp.f:326-337 and read
back through that format (src/coordsfp.f90:21-29), so all consumers use the
rounded values the reference would read.
"""
import math
import os
from dataclasses import dataclass, field, asdict

import numpy as np

_K1 = np.uint64(0x9E3779B97F4A7C15)
_K2 = np.uint64(0xD1B54A32D192ED03)


def _splitmix(x):
    with np.errstate(over="ignore"):
        x = x + _K1
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def uniform(seed, stream, idx):
    """Counter-based U[0,1) doubles for integer counters idx (array)."""
    idx = np.asarray(idx, dtype=np.uint64)
    with np.errstate(over="ignore"):
        x = np.uint64(seed) * _K1 + np.uint64(stream) * _K2 + idx
    return (_splitmix(x) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


@dataclass
class Case:
    """One case: the &PARAMETERS namelist (src/param.f90:66-86) plus the sizes
    src/gridsize.f90 fixes at compile time, the seed and the run length."""
    name: str = "case"
    itot: int = 64
    jtot: int = 64
    ktot: int = 64
    lx: float = 1.0
    ly: float = 1.0
    lz: float = 1.0
    np: int = 8
    diameter_h: float = 8.0        # sphere diameter in grid spacings
    ratiorho: float = 1.5
    re_bulk: float = 2000.0        # visc = 1 / re_bulk (U_bulk = 1, lz = 1)
    pran: float = 1.0
    beta: float = 0.0
    perturb: float = 0.1
    particle_rotation: bool = False
    bc_uvw_top: int = 0            # 0 BC_NOSLIP, 1 BC_FREESLIP
    bc_uvw_bot: int = 0
    bc_c_top: int = 0              # 0 BC_DIRICHLET, 1 BC_NEUMANN
    bc_c_bot: int = 0
    bc_c_top_val: float = 0.0
    bc_c_bot_val: float = 0.0
    seed: int = 1
    steps: int = 100
    ioutchk: int = 10
    nu: float = 0.0                # explicit kinematic viscosity (overrides 1/re_bulk when > 0)

    @property
    def visc(self):
        return self.nu if self.nu > 0.0 else 1.0 / self.re_bulk

    @property
    def radius(self):
        return 0.5 * self.diameter_h * (self.lx / self.itot)

    @property
    def nl(self):
        return nl_formula(self.itot, self.lx, self.radius)

    @property
    def cells(self):
        return self.itot * self.jtot * self.ktot

    def as_dict(self):
        return asdict(self)


def nl_formula(itot, lx, radius):
    """src/init.f90:127-133: nl = nint((pi/3)*(12*((radius-retrac)*dxi)**2+1))."""
    pi = math.acos(-1.0)
    dxi = itot / lx
    retrac = 0.3 / dxi
    v = (pi / 3.0) * (12.0 * (((radius - retrac) * dxi) * ((radius - retrac) * dxi)) + 1.0)
    return int(math.floor(v + 0.5)) if v >= 0 else -int(math.floor(-v + 0.5))


# The configurations the benchmark and the parity tests run (sizes of the
# reference method's channel: lx = ly = lz = 1, flow along y, walls in z).
CONFIGS = {
    "ch64_8sph": Case(name="ch64_8sph", itot=64, jtot=64, ktot=64, np=8, diameter_h=8.0, ratiorho=1.5, re_bulk=2000.0,
                  seed=1, steps=100),
    "ch256_1000sph": Case(name="ch256_1000sph", itot=256, jtot=256, ktot=256, np=1000, diameter_h=12.0, ratiorho=1.5,
                   re_bulk=5000.0, seed=2, steps=1000),
    "ch512_8000sph": Case(name="ch512_8000sph", itot=512, jtot=512, ktot=512, np=8000, diameter_h=16.0, ratiorho=2.0,
                   re_bulk=10000.0, seed=4, steps=200),
    # small parity cases
    "t16": Case(name="t16", itot=16, jtot=16, ktot=16, np=2, diameter_h=5.0, ratiorho=1.5, re_bulk=200.0, seed=7,
                steps=4, ioutchk=2),
    "t32": Case(name="t32", itot=32, jtot=32, ktot=32, np=6, diameter_h=8.0, ratiorho=1.5, re_bulk=1000.0, seed=3,
                steps=10, ioutchk=5),
    "t24x16x32": Case(name="t24x16x32", itot=24, jtot=16, ktot=32, np=3, diameter_h=6.0, ratiorho=1.5,
                      re_bulk=500.0, seed=5, steps=4, ioutchk=2),
}


def initial_fields(case):
    """unew, vnew, wnew, pnew, cnew as (ktot+2, jtot+2, itot+2) arrays, i.e.
    real(8) (0:i1,0:j1,0:k1) in Fortran order, before any boundary call."""
    nx, ny, nz = case.itot, case.jtot, case.ktot
    shape = (nz + 2, ny + 2, nx + 2)
    u = np.zeros(shape)
    v = np.zeros(shape)
    w = np.zeros(shape)
    p = np.zeros(shape)
    c = np.zeros(shape)
    dz = 1.0 / (nz / case.lz)
    k = np.arange(nz + 2, dtype=np.float64)
    coorz = (k - 0.5) * dz / case.lz
    prof = 6.0 * 1.0 * coorz * (1.0 - coorz)
    v[:, :, :] = prof[:, None, None]
    if case.perturb != 0.0:
        idx = np.arange(nx * ny * nz, dtype=np.uint64).reshape(nz, ny, nx)
        for comp, arr in enumerate((u, v, w)):
            r = uniform(case.seed, 100 + comp, idx)
            arr[1:-1, 1:-1, 1:-1] = arr[1:-1, 1:-1, 1:-1] + case.perturb * (2.0 * r - 1.0)
    return u, v, w, p, c


def place_spheres(case, max_tries=2_000_000):
    """Seeded rejection sampling: centre gap >= D + h (minimum image in the
    periodic x, y directions), wall gap >= R + h.  Returns x, y, z arrays."""
    n = case.np
    if n == 0:
        return np.zeros(0), np.zeros(0), np.zeros(0)
    h = case.lx / case.itot
    R = case.radius
    gap = 2.0 * R + h
    zlo, zhi = R + h, case.lz - R - h
    ncx = max(1, int(case.lx / gap))
    ncy = max(1, int(case.ly / gap))
    ncz = max(1, int(case.lz / gap))
    grid = {}
    xs, ys, zs = [], [], []
    counter = 0
    block = 4096
    while len(xs) < n:
        r = uniform(case.seed, 200, np.arange(counter, counter + 3 * block, dtype=np.uint64)).reshape(block, 3)
        counter += 3 * block
        if counter > 3 * max_tries:
            raise RuntimeError(f"could not place {n} spheres")
        for a, b, cc in r:
            x = case.lx * a
            y = case.ly * b
            z = zlo + (zhi - zlo) * cc
            ix, iy, iz = int(x / case.lx * ncx) % ncx, int(y / case.ly * ncy) % ncy, min(int(z / case.lz * ncz), ncz - 1)
            ok = True
            for ddx in (-1, 0, 1):
                for ddy in (-1, 0, 1):
                    for ddz in (-1, 0, 1):
                        key = ((ix + ddx) % ncx, (iy + ddy) % ncy, iz + ddz)
                        for q in grid.get(key, ()):
                            dx = x - xs[q]
                            dx -= case.lx * round(dx / case.lx)
                            dy = y - ys[q]
                            dy -= case.ly * round(dy / case.ly)
                            dzz = z - zs[q]
                            if dx * dx + dy * dy + dzz * dzz < gap * gap:
                                ok = False
                                break
                        if not ok:
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if ok:
                grid.setdefault((ix, iy, iz), []).append(len(xs))
                xs.append(x)
                ys.append(y)
                zs.append(z)
                if len(xs) == n:
                    break
    return np.array(xs), np.array(ys), np.array(zs)


def fortran_e16_8(v):
    """Fortran E16.8 edit descriptor: ' 0.12345678E+01' right-justified in 16."""
    if v == 0.0 or not math.isfinite(v):
        s = "0.00000000E+00"
        return s.rjust(16)
    sgn = "-" if v < 0 else ""
    a = abs(v)
    m, e = f"{a:.7E}".split("E")
    digits = m.replace(".", "")
    ex = int(e) + 1
    s = f"{sgn}0.{digits}E{'+' if ex >= 0 else '-'}{abs(ex):02d}"
    return s.rjust(16)


def marker_angles(nl):
    """Golden-angle spiral of nl points on the unit sphere: (theta, phi)."""
    l = np.arange(nl, dtype=np.float64)
    zc = 1.0 - (2.0 * l + 1.0) / nl
    theta = np.arccos(zc)
    phi = np.mod(l * (math.pi * (3.0 - math.sqrt(5.0))), 2.0 * math.pi)
    return theta, phi


def write_markers(path, case):
    """Write lagrangianforcepoints2 (3 header lines, nl lines of 6E16.8:
    x, y, z, theta, phi, radfp) as src/poslfp/lfp.f:326-337 does."""
    nl = case.nl
    radfp = case.radius - 0.3 / (case.itot / case.lx)
    theta, phi = marker_angles(nl)
    with open(path, "w") as f:
        f.write(' VARIABLES = "lfpx","lfpy","lfpz","theta","phi","radfp"\n')
        f.write(f' ZONE T="Zone1" I={nl:12d}, F=POINT\n')
        f.write(" \n")
        for t, p in zip(theta, phi):
            vals = (radfp * math.sin(t) * math.cos(p), radfp * math.sin(t) * math.sin(p), radfp * math.cos(t), t, p, radfp)
            f.write("".join(fortran_e16_8(x) for x in vals) + "\n")
    return path


def read_markers(path, nl):
    """src/coordsfp.f90:21-29: skip 3 lines, read nl records (6E16.8) and keep
    theta, phi and radfp."""
    theta = np.zeros(nl)
    phi = np.zeros(nl)
    radfp = 0.0
    with open(path) as f:
        for _ in range(3):
            f.readline()
        for l in range(nl):
            line = f.readline()
            vals = [float(line[16 * q:16 * (q + 1)]) for q in range(6)]
            theta[l], phi[l], radfp = vals[3], vals[4], vals[5]
    return theta, phi, radfp


def markers_for(case, workdir=None):
    """Marker angles and radius as every consumer reads them from the file."""
    import tempfile
    d = workdir or tempfile.mkdtemp(prefix="ipg_lfp_")
    path = os.path.join(d, "lagrangianforcepoints2")
    write_markers(path, case)
    return read_markers(path, case.nl)


def write_namelist(path, case, iout2d=100000):
    """parameters.inp in the &PARAMETERS namelist of src/param.f90:66-86."""
    bcu = {0: "BC_NOSLIP", 1: "BC_FREESLIP"}
    bcc = {0: "BC_DIRICHLET", 1: "BC_NEUMANN"}
    with open(path, "w") as f:
        f.write("&PARAMETERS\n")
        f.write(f"  lx = {case.lx!r}, ly = {case.ly!r}, lz = {case.lz!r}\n")
        f.write("  iniu = 'poi', inic = 'ZERC'\n")
        f.write(f"  bc_uvw_top_type = '{bcu[case.bc_uvw_top]}', bc_uvw_bot_type = '{bcu[case.bc_uvw_bot]}'\n")
        f.write(f"  bc_c_top_type = '{bcc[case.bc_c_top]}', bc_c_bot_type = '{bcc[case.bc_c_bot]}'\n")
        f.write(f"  bc_c_top_val = {case.bc_c_top_val!r}, bc_c_bot_val = {case.bc_c_bot_val!r}\n")
        f.write(f"  visc = {case.visc!r}, Pran = {case.pran!r}, beta = {case.beta!r}\n")
        f.write(f"  radius = {case.radius!r}, ratiorho = {case.ratiorho!r}\n")
        f.write(f"  particle_rotation = {'.true.' if case.particle_rotation else '.false.'}\n")
        f.write(f"  ioutchk = {case.ioutchk}, iout1d = {case.ioutchk}, iout2d = {iout2d}, ioutfld = 100000\n")
        f.write("  runmode = 0\n")
        f.write("/\n")
    return path


def write_run_nml(path, case, nstep, begin=0, use_graphs=True, fused_substep=True, t_end=1.0e30, precision=0,
                  mixed_solve=32):
    """run.nml of the Fortran driver: &GRIDSIZE (the compile-time sizes of
    src/gridsize.f90:6-26) and &RUNCONTROL (start step, step count, seed; the
    arithmetic: precision 0 FP64, 1 mixed with its Poisson solve mixed_solve)."""
    tf = {True: ".true.", False: ".false."}
    with open(path, "w") as f:
        f.write(f"&GRIDSIZE itot = {case.itot}, jtot = {case.jtot}, ktot = {case.ktot}, np = {case.np}, "
                f"nlmax = {case.nl if case.np > 0 else 0}, t_end = {t_end!r} /\n")
        f.write(f"&RUNCONTROL begin = {begin}, nstep = {nstep}, seed = {case.seed}, perturb = {case.perturb!r}, "
                f"use_graphs = {tf[use_graphs]}, fused_substep = {tf[fused_substep]}, precision = {precision}, "
                f"mixed_solve = {mixed_solve} /\n")
    return path


def write_case_dir(d, case, nstep, iout2d=100000, **kw):
    """parameters.inp, run.nml and (with spheres) lagrangianforcepoints2 in d."""
    os.makedirs(d, exist_ok=True)
    write_namelist(os.path.join(d, "parameters.inp"), case, iout2d=iout2d)
    write_run_nml(os.path.join(d, "run.nml"), case, nstep, **kw)
    if case.np > 0:
        write_markers(os.path.join(d, "lagrangianforcepoints2"), case)
    return d


def fld_path(d, nr):
    return os.path.join(d, f"fld{nr:07d}")


def read_fld(path, case):
    """fld%07d (src/loadd.f90:30-67): unew vnew wnew pnew cnew over the interior,
    i fastest, then (time, istep, dt).  Returns ({name: (ktot, jtot, itot)}, info)."""
    n = case.itot * case.jtot * case.ktot
    a = np.fromfile(path, dtype="<f8")
    if a.size != 5 * n + 3:
        raise ValueError(f"{path}: {a.size} values, expected {5 * n + 3}")
    shape = (case.ktot, case.jtot, case.itot)
    out = {name: a[q * n:(q + 1) * n].reshape(shape) for q, name in enumerate(("U", "V", "W", "PNEW", "C"))}
    return out, tuple(a[5 * n:])


def write_fld(path, case, fields, time, istep, dt):
    """Inverse of read_fld; fields maps U V W PNEW C to interior (or ghosted) arrays."""
    parts = []
    for name in ("U", "V", "W", "PNEW", "C"):
        x = np.asarray(fields[name], dtype=np.float64)
        if x.shape == (case.ktot + 2, case.jtot + 2, case.itot + 2):
            x = x[1:-1, 1:-1, 1:-1]
        parts.append(np.ascontiguousarray(x).ravel())
    parts.append(np.array([time, float(istep), dt]))
    np.concatenate(parts).astype("<f8").tofile(path)


def read_partstate(path, case):
    """partstate%07d of the Fortran driver: 31 quantities x np FP64, then
    (time, istep, dt)."""
    from ._lib import PARTS
    a = np.fromfile(path, dtype="<f8")
    n = case.np
    if a.size != len(PARTS) * n + 3:
        raise ValueError(f"{path}: {a.size} values")
    return {q: a[m * n:(m + 1) * n] for m, q in enumerate(PARTS)}, tuple(a[-3:])


# the 29 leading reals of a record of type particle_restart (src/loadd.f90:123-138)
ALLPART_FIELDS = ("X", "Y", "Z", "THETA", "PHI", "U", "V", "W", "OMX", "OMY", "OMZ", "OMTHETA", "INTU", "INTV",
                  "INTW", "INTOMX", "INTOMY", "INTOMZ", "FLAG1", "FLAG2", "COLFX", "COLFY", "COLFZ", "COLTX", "COLTY",
                  "COLTZ", "FXLTOT", "FYLTOT", "FZLTOT")


def allpartdata_layout(npart):
    """(nqmax, reals per record, reals per particle in the file) of
    allpartdata%07d: nqmax = min(np + 2, 15) (src/param.f90:100), a record of
    type particle_restart is 31 + 11 nqmax reals, and the reference moves
    skip = 29 + 11 nqmax per particle (src/loadd.f90:140-142)."""
    nq = min(npart + 2, 15)
    return nq, 31 + 11 * nq, 29 + 11 * nq


def read_allpartdata(path, case):
    """allpartdata%07d as an (np, record) array of the reference's record
    array; the missing tail (the last 2 np reals, never written) is NaN."""
    nq, nrec, nfile = allpartdata_layout(case.np)
    a = np.fromfile(path, dtype="<f8")
    if a.size != nfile * case.np:
        raise ValueError(f"{path}: {a.size} values, expected {nfile * case.np}")
    rec = np.full(nrec * case.np, np.nan)
    rec[:a.size] = a
    return rec.reshape(case.np, nrec)


def allpartdata_idp_slot(npart):
    """index of idp in a record (after the 29 reals, 10 nqmax history, qmax)"""
    nq = min(npart + 2, 15)
    return 29 + 10 * nq + 1


def ensemble_cases(n=64, seed=5):
    """The batched ensemble: n independent cases with their own seeds.  Grids
    from {64^3, 96x96x64, 128^3, 128x128x192}; D uniform in [8h, 12h]; Re_bulk
    uniform in [1000, 5000]; sphere count uniform in [8, 200], capped so that
    the solid fraction stays <= 10 % (seeded rejection sampling with a gap of
    D + h cannot pack more).  All draws come from stream 300 of `seed`."""
    grids = [(64, 64, 64), (96, 96, 64), (128, 128, 128), (128, 128, 192)]
    r = uniform(seed, 300, np.arange(5 * n, dtype=np.uint64)).reshape(n, 5)
    out = []
    for q in range(n):
        nx, ny, nz = grids[min(int(r[q, 0] * 4), 3)]
        d_h = 8.0 + 4.0 * r[q, 1]
        re = 1000.0 + 4000.0 * r[q, 2]
        npart = 8 + int(r[q, 3] * 193)
        vol_sphere = 4.0 / 3.0 * math.pi * (0.5 * d_h) ** 3
        cap = int(0.10 * nx * ny * nz / vol_sphere)
        out.append(Case(name=f"e{q:02d}_{nx}x{ny}x{nz}", itot=nx, jtot=ny, ktot=nz, lx=1.0, ly=ny / nx,
                        lz=nz / nx, np=max(1, min(npart, cap)), diameter_h=d_h, ratiorho=1.5, re_bulk=re,
                        seed=1000 + q, steps=10))
    return out


# the files post3d writes (src/output.f90:1015-1074), in its order
POST3D = ("pre", "vez", "con", "vox", "voy", "sol")


def read_post3d(path, case):
    """name%07d_3d.bin of post3d (src/output.f90:1099-1125): the interior,
    i fastest, as a (ktot, jtot, itot) array."""
    a = np.fromfile(path, dtype="<f8")
    return a.reshape(case.ktot, case.jtot, case.itot)


def read_partvisu(path, case):
    """partvisu%07d.bin of outputpart (src/output.f90:1127-1273): fltot and
    accel, each a (np, 3) array (x, y, z per sphere, spheres in id order)."""
    a = np.fromfile(path, dtype="<f8")
    assert a.size == 6 * case.np, (a.size, case.np)
    return a[:3 * case.np].reshape(case.np, 3), a[3 * case.np:].reshape(case.np, 3)
