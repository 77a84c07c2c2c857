"""Pins the FP64 CPU oracle (oracle/ipg_oracle.c) against the reference's own
outputs and against what the mathematics fixes, so that the oracle can serve
as the parity reference for the GPU path.

 * delta kernel and kerneltest: bitwise against the reference's kernel.f90
   (golden vectors from oracle/_ref, tests/golden/ref_vectors.json), plus the
   moment conditions of the Roma et al. kernel;
 * real FFT: against VFFTPACK (fft.f golden vectors) and numpy.fft;
 * chkdt: the dtmax known answer of src/terminate.txt:7, bitwise;
 * momentum / scalar operators: closed forms on cosine fields (discrete
   eigenfunctions of the central differences);
 * Poisson solve: a dense least-squares solve of the same discrete operator;
 * projection: divergence-free to round-off;
 * IBM: interpolation/spreading adjointness, partition of unity, force
   conservation;
 * sphere integrals: volume vs 4/3 pi R^3, uniform and rigid-rotation fields,
   periodic images;
 * Newton-Euler: free fall closed form; collisions: the spring force closed
   form with the one-rank multiplicity (9 copies for sphere-sphere, 1 for walls).
"""
import json
import math
import os

import numpy as np
import pytest

from interparts_mpich_amd.inputs import Case, CONFIGS, initial_fields, place_spheres, markers_for, nl_formula
from oracle import pyoracle
from oracle.pyoracle import Oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = json.load(open(os.path.join(HERE, "golden", "ref_vectors.json")))


def lib():
    return pyoracle.load()


# ---------------------------------------------------------------- kernel ----
def test_kernel_bitwise_against_reference_vectors():
    L = lib()
    for r, want in zip(GOLD["kernel_r"], GOLD["kernel"]):
        assert L.orc_kernel(r) == want, r


def test_kernel_against_compiled_reference_dense():
    ref = pyoracle.ref_lib()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    import ctypes as C
    L = lib()
    for r in np.linspace(-2.0, 2.0, 20001):
        assert L.orc_kernel(float(r)) == ref._QMmod_kernelPkernel(C.byref(C.c_double(float(r))))


def test_kernel_moments():
    """Roma et al. (1999) 3-point kernel: sum_j phi(r-j) = 1, sum_j (r-j) phi(r-j) = 0,
    sum_j phi(r-j)^2 = 1/2 for every shift r."""
    L = lib()
    for r in np.linspace(0.0, 1.0, 97):
        ph = np.array([L.orc_kernel(float(r - j)) for j in range(-3, 4)])
        d = np.array([r - j for j in range(-3, 4)])
        assert abs(ph.sum() - 1.0) < 1e-15
        assert abs((d * ph).sum()) < 1e-15
        assert abs((ph * ph).sum() - 0.5) < 1e-15


def test_kerneltest_matches_reference():
    assert lib().orc_kerneltest() == GOLD["kerneltest"]


# ------------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("n", [8, 12, 16, 24, 64, 96])
def test_rfft_against_vfftpack_and_numpy(n):
    L = lib()
    g = GOLD["vfftpack"][str(n)]
    x = np.array(g["x"])
    a = x.copy()
    L.orc_rfftf(n, a.ctypes.data)
    np.testing.assert_allclose(a, g["forward"], rtol=0, atol=2e-15 * np.abs(g["forward"]).max() * 8)
    F = np.fft.rfft(x) / math.sqrt(n)
    pk = np.empty(n)
    pk[0] = F[0].real
    pk[1:n - 1:2] = F[1:n // 2].real
    pk[2:n - 1:2] = F[1:n // 2].imag
    pk[n - 1] = F[n // 2].real
    np.testing.assert_allclose(a, pk, rtol=0, atol=1e-14)
    b = x.copy()
    L.orc_rfftb(n, b.ctypes.data)
    np.testing.assert_allclose(b, g["backward"], rtol=0, atol=2e-15 * np.abs(g["backward"]).max() * 8)
    L.orc_rfftb(n, a.ctypes.data)
    np.testing.assert_allclose(a, x, rtol=0, atol=1e-14)


# ----------------------------------------------------------------- chkdt ----
def test_chkdt_known_answer():
    """src/terminate.txt:2,5,7: dtmax = 0.37765502929687511 for the 1003 deck
    (tests/1003/{parameters.inp,gridsize.inp}) at rest."""
    case = Case(name="deck1003", itot=128, jtot=128, ktot=256, lx=0.27, ly=0.27, lz=0.54, np=0, nu=1.62e-6,
                perturb=0.0)
    o = Oracle(case)
    assert o.chkdt() == 0.37765502929687511


# ------------------------------------------------------- stencil closed forms
def _setup_shear(case, comp, direc, amp, base, adv):
    """Field comp ('U','V','W','C') = base + amp*cos(k xi) along direction
    direc (0,1,2); the other velocity components constant adv; ghosts filled
    analytically (no boundary call)."""
    o = Oracle(case)
    n = (case.itot, case.jtot, case.ktot)[direc]
    k = 2.0 * math.pi * 2.0 / n  # two periods over the direction
    shp = o.shape
    idx = np.indices(shp)[2 - direc].astype(np.float64)
    prof = base + amp * np.cos(k * idx)
    vel = {"U": np.full(shp, adv[0]), "V": np.full(shp, adv[1]), "W": np.full(shp, adv[2]), "C": np.zeros(shp)}
    vel[comp] = prof
    for name, a in vel.items():
        o.put(name, a)
    o.put("PNEW", np.zeros(shp))
    o.set_dt(1e-3)
    o.rk(1)
    return o, k, idx


@pytest.mark.parametrize("comp", ["U", "V", "W", "C"])
@pytest.mark.parametrize("direc", [0, 1, 2])
def test_advection_diffusion_closed_form(comp, direc):
    case = Case(name="stencil", itot=12, jtot=16, ktot=20, np=0, re_bulk=3.0, pran=0.7)
    h = [case.lx / case.itot, case.ly / case.jtot, case.lz / case.ktot][direc]
    hi = 1.0 / h
    adv = [0.3, -0.45, 0.6]
    amp, base = 0.2, 0.8
    self_adv = (comp == "UVW"[direc]) if comp != "C" else False
    if self_adv:
        adv[direc] = 0.0
    o, k, idx = _setup_shear(case, comp, direc, amp, base, adv)
    out = {"U": "DURK", "V": "DVRK", "W": "DWRK", "C": "DCRK"}[comp]
    got = o.get(out)[1:-1, 1:-1, 1:-1]
    x = idx[1:-1, 1:-1, 1:-1]
    s, c = np.sin(k * x), np.cos(k * x)
    diff = (case.visc if comp != "C" else case.visc / case.pran) * hi * hi * (-4.0 * amp * c * math.sin(k / 2) ** 2)
    if comp == "C":
        # scalad: flux 0.5*U*(c_i+1 + c_i); constant advecting velocity U_d
        want = adv[direc] * amp * hi * s * math.sin(k) + diff
    elif not self_adv:
        want = adv[direc] * amp * hi * s * math.sin(k) + diff
    else:
        # -dxi/4 [(q_i+1 + q_i)^2 - (q_i-1 + q_i)^2] for q = base + amp cos(k i)
        dq = -2.0 * s * math.sin(k)                     # c_i+1 - c_i-1
        sq = 2.0 * c * (1.0 + math.cos(k))              # c_i+1 + 2 c_i + c_i-1
        want = -hi * 0.25 * (4.0 * base * amp * dq + amp * amp * dq * sq) + diff
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12 * max(1.0, np.abs(want).max()))


def test_pressure_gradient_closed_form():
    case = Case(name="grad", itot=12, jtot=16, ktot=20, np=0)
    o = Oracle(case)
    shp = o.shape
    kk, jj, ii = np.indices(shp).astype(np.float64)
    kx, ky, kz = 2 * math.pi / 12, 4 * math.pi / 16, 2 * math.pi / 20
    p = np.cos(kx * ii) + np.sin(ky * jj) + np.cos(kz * kk)
    z = np.zeros(shp)
    for n in ("U", "V", "W", "C"):
        o.put(n, z)
    o.put("PNEW", p)
    dt = 0.01
    o.set_dt(dt)
    o.rk(1)
    f1 = dt * (32.0 / 60.0)
    I = (slice(1, -1),) * 3
    want_u = -12.0 * (np.cos(kx * (ii + 1)) - np.cos(kx * ii))
    want_v = -16.0 * (np.sin(ky * (jj + 1)) - np.sin(ky * jj))
    want_w = -20.0 * (np.cos(kz * (kk + 1)) - np.cos(kz * kk))
    np.testing.assert_allclose(o.get("DUDT")[I], f1 * want_u[I], atol=1e-13)
    np.testing.assert_allclose(o.get("DVDT")[I], f1 * want_v[I], atol=1e-13)
    np.testing.assert_allclose(o.get("DWDT")[I], f1 * want_w[I], atol=1e-13)


# ---------------------------------------------------------------- Poisson ---
def _dense_laplacian(nx, ny, nz, dxi, dyi, dzi):
    N = nx * ny * nz
    A = np.zeros((N, N))
    idx = lambda i, j, k: (i % nx) + nx * ((j % ny) + ny * k)
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                r = idx(i, j, k)
                for di, s in ((1, dxi * dxi), (-1, dxi * dxi)):
                    A[r, idx(i + di, j, k)] += s
                    A[r, r] -= s
                for dj, s in ((1, dyi * dyi), (-1, dyi * dyi)):
                    A[r, idx(i, j + dj, k)] += s
                    A[r, r] -= s
                for dk in (1, -1):  # Neumann walls: zero-gradient ghost
                    if 0 <= k + dk < nz:
                        A[r, idx(i, j, k + dk)] += dzi * dzi
                        A[r, r] -= dzi * dzi
    return A


@pytest.mark.parametrize("dims", [(4, 6, 8), (6, 4, 5), (8, 8, 6)])
def test_solver2d_against_dense_least_squares(dims):
    nx, ny, nz = dims
    case = Case(name="poisson", itot=nx, jtot=ny, ktot=nz, lx=1.0, ly=1.3, lz=0.7, np=0)
    o = Oracle(case)
    o.initsolver()
    rng = np.random.default_rng(sum(dims))
    rhs = rng.standard_normal((nz, ny, nx))
    rhs -= rhs.mean()
    P = np.zeros(o.shape)
    P[1:-1, 1:-1, 1:-1] = rhs
    o.put("P", P)
    o.solver2d()
    got = o.get("P")[1:-1, 1:-1, 1:-1].ravel()
    A = _dense_laplacian(nx, ny, nz, nx / case.lx, ny / case.ly, nz / case.lz)
    want = np.linalg.lstsq(A, rhs.ravel(), rcond=None)[0]
    np.testing.assert_allclose(got - got.mean(), want - want.mean(), rtol=0, atol=1e-11 * np.abs(want).max())


def test_projection_is_divergence_free():
    case = CONFIGS["t24x16x32"]
    o = Oracle(case).initialize(initial_fields(case), place_spheres(case), markers_for(case))
    tot0, mx0 = o.chkdiv()
    assert mx0 > 1e-2  # the seeded noise is not solenoidal
    o.step()
    tot, mx = o.chkdiv()
    assert mx < 1e-12 and abs(tot) < 1e-10


# -------------------------------------------------------------------- IBM ---
def _ibm_case():
    case = Case(name="ibm", itot=24, jtot=20, ktot=28, np=3, diameter_h=6.0, seed=11)
    xs = np.array([0.05, 0.5, 0.93])   # first and last cross the periodic x boundary
    ys = np.array([0.5, 0.02, 0.7])
    zs = np.array([0.5, 0.45, 0.55])   # away from the walls
    o = Oracle(case).initialize(initial_fields(case), (xs, ys, zs), markers_for(case))
    return case, o


def test_ibm_adjointness_and_conservation():
    case, o = _ibm_case()
    rng = np.random.default_rng(5)
    dt = o.dt
    nl, np_ = case.nl, case.np
    c0 = (o.scalar(0) / o.scalar(1)) * dt
    for comp, fname, fl, ul in (("u", "DUDTF", "FXL", "DUDTL"), ("v", "DVDTF", "FYL", "DVDTL"), ("w", "DWDTF", "FZL", "DWDTL")):
        u = rng.standard_normal((case.ktot, case.jtot, case.itot))
        for f in ("DUDTF", "DVDTF", "DWDTF"):
            o.view(f)[...] = 0.0
        o.view(fname)[2:-2, 2:-2, 2:-2] = u
        o.lib.orc_eulr2lagr(o.s)
        nla = o.nla()
        Iu = np.concatenate([o.mview(ul)[p * nl:p * nl + nla[p]] for p in range(np_)])
        F = rng.standard_normal(Iu.shape)
        fv = o.mview(fl)
        off = 0
        for p in range(np_):
            fv[p * nl:p * nl + nla[p]] = F[off:off + nla[p]]
            off += nla[p]
        for f in ("DUDTOLD", "DVDTOLD", "DWDTOLD"):
            o.view(f)[...] = 0.0
        o.lib.orc_lagr2eulr(o.s)
        SF = o.view(fname)[2:-2, 2:-2, 2:-2]
        lhs = float(np.sum(u * SF))
        rhs = c0 * float(np.sum(F * Iu))
        assert abs(lhs - rhs) <= 1e-12 * abs(rhs), (comp, lhs, rhs)
        # force conservation: sum of spread force = c0 * sum of marker forces
        assert abs(float(SF.sum()) - c0 * float(F.sum())) <= 1e-12 * np.abs(F).sum() * c0
        # partition of unity: a uniform field is interpolated exactly
        o.view(fname)[2:-2, 2:-2, 2:-2] = 1.0
        o.lib.orc_eulr2lagr(o.s)
        Iu1 = np.concatenate([o.mview(ul)[p * nl:p * nl + nla[p]] for p in range(np_)])
        np.testing.assert_allclose(Iu1, 1.0, rtol=0, atol=1e-14)


# ---------------------------------------------------- sphere integrals ------
def test_sphere_integrals():
    case = Case(name="sph", itot=32, jtot=32, ktot=32, np=2, diameter_h=10.0, perturb=0.0)
    h = 1.0 / 32
    R = case.radius
    xs = np.array([3.3 * h, 3.3 * h + 0.5])   # the first straddles x = 0; same sub-cell offset
    ys = np.array([0.41, 0.41])
    zs = np.array([0.52, 0.52])
    o = Oracle(case).initialize(initial_fields(case), (xs, ys, zs), markers_for(case))
    o.intgr_over_sphere(1)
    vols = {n: o.pget(n) for n in ("INTU", "INTV", "INTW")}  # staggered-cell volumes of each component
    vol = vols["INTU"]
    exact = 4.0 / 3.0 * math.pi * R ** 3
    assert abs(vol[0] - vol[1]) <= 1e-13 * vol[1]          # periodic images complete the sphere
    assert abs(vol[1] - exact) / exact < 0.05               # % error print, src/intgr_over_sphere.f90:607-622
    shp = o.shape
    U0, V0, W0 = 0.7, -0.2, 0.35
    o.put("U", np.full(shp, U0)); o.put("V", np.full(shp, V0)); o.put("W", np.full(shp, W0))
    o.intgr_over_sphere(2)
    for name, val in (("INTU", U0), ("INTV", V0), ("INTW", W0)):
        np.testing.assert_allclose(o.pget(name), val * vols[name], rtol=1e-13)
    # rigid rotation about the z axis through sphere 1's centre
    Om = 1.3
    kk, jj, ii = np.indices(shp).astype(np.float64)
    yu = (jj - 0.5) * h - ys[1]          # u lives at (i h, (j-1/2) h)
    xv = (ii - 0.5) * h - xs[1]          # v lives at ((i-1/2) h, j h)
    o.put("U", -Om * yu); o.put("V", Om * xv); o.put("W", np.zeros(shp))
    o.intgr_over_sphere(3)
    Iz = o.pget("INTOMZ")[1]
    want = Om * 8.0 * math.pi / 15.0 * R ** 5
    assert abs(Iz - want) / want < 0.1
    assert abs(o.pget("INTOMX")[1]) < 1e-3 * want and abs(o.pget("INTOMY")[1]) < 1e-3 * want


# ----------------------------------------------------------- Newton-Euler ---
def test_free_fall_closed_form():
    case = Case(name="ff", itot=16, jtot=16, ktot=32, np=1, diameter_h=4.0, ratiorho=1.5, perturb=0.0)
    z0 = 0.6
    o = Oracle(case).initialize([np.zeros((34, 18, 18))] * 5, (np.array([0.5]), np.array([0.5]), np.array([z0])),
                                markers_for(case))
    dt = 1e-3
    o.set_dt(dt)
    g = -9.81 * (1.0 - 1.0 / case.ratiorho)
    t = 0.0
    for step in range(4):
        for rk in (1, 2, 3):
            o.intgr_nwtn_eulr(rk)
        t += dt
        np.testing.assert_allclose(o.pget("W"), g * t, rtol=1e-12)
        np.testing.assert_allclose(o.pget("Z"), z0 + 0.5 * g * t * t, rtol=1e-12)


def test_substep_sums_closed_form():
    """sumrk3 (src/intgr_nwtn_eulr.f90:1586-1669) is called on both passes of
    the collision loop (iter = -1..itermax-1 = 0..1, lines 149-152, 672) and
    resets at rkiter = 1 on each: a sphere in fluid at rest with a fixed IBM
    force f sees, per call, du/dt = -f/(V rho_r) + rkab g (1 - 1/rho_r) and
    fltot = -f, summed with multiplicities 1, 2, 2 over the three substeps."""
    case = Case(name="ff", itot=16, jtot=16, ktot=32, np=1, diameter_h=4.0, ratiorho=1.5, perturb=0.0)
    o = Oracle(case).initialize([np.zeros((34, 18, 18))] * 5, (np.array([0.5]), np.array([0.5]), np.array([0.6])),
                                markers_for(case))
    o.set_dt(1e-3)
    f = np.array([2e-4])
    vr = o.scalar(2) * case.ratiorho
    g = -9.81 * (1.0 - 1.0 / case.ratiorho)
    rkab = (8.0 / 15.0, 2.0 / 15.0, 1.0 / 3.0)
    for step in range(2):
        for rk in (1, 2, 3):
            o.pput("FZLTOT", f)
            o.pput("FXLTOT", -f)
            o.intgr_nwtn_eulr(rk)
        mult = (1.0, 2.0, 2.0)
        np.testing.assert_allclose(o.pget("RKFZ"), -5.0 * f, rtol=1e-13)
        np.testing.assert_allclose(o.pget("RKFX"), 5.0 * f, rtol=1e-13)
        assert o.pget("RKFY")[0] == 0.0
        want_w = sum(m * (-f[0] / vr + a * g) for m, a in zip(mult, rkab))
        np.testing.assert_allclose(o.pget("RKDWDT"), want_w, rtol=1e-9)
        np.testing.assert_allclose(o.pget("RKDUDT"), sum(m * (f[0] / vr) for m in mult), rtol=1e-9)


def test_collision_spring_force_and_multiplicity():
    """Two spheres at rest overlapping by delta: the normal spring force
    -kn*delta*n is added once per neighbour-table copy, i.e. 9 times at one
    rank (src/intgr_nwtn_eulr.f90:358-499, the H1 multiplicity); a sphere
    pressed into the bottom wall gets kn_sw*delta once."""
    case = Case(name="col", itot=32, jtot=32, ktot=32, np=3, diameter_h=6.0, perturb=0.0)
    R = case.radius
    delta = 0.01 * R
    xs = np.array([0.40, 0.40 + 2 * R - delta, 0.8])
    ys = np.array([0.5, 0.5, 0.3])
    zs = np.array([0.5, 0.5, R - delta])
    o = Oracle(case).initialize([np.zeros((34, 34, 34))] * 5, (xs, ys, zs), markers_for(case))
    o.set_dt(1e-12)
    o.intgr_nwtn_eulr(1)
    kn_ss, kn_sw = o.scalar(4), o.scalar(6)
    cf = o.pget("COLFX")
    assert abs(cf[0] - (-9.0 * kn_ss * delta)) <= 1e-9 * kn_ss * delta
    assert abs(cf[1] - (9.0 * kn_ss * delta)) <= 1e-9 * kn_ss * delta
    assert abs(o.pget("COLFZ")[2] - kn_sw * delta) <= 1e-9 * kn_sw * delta


def test_lubrication_vanishes_at_onset_and_opposes_approach():
    case = Case(name="lub", itot=32, jtot=32, ktot=32, np=2, diameter_h=6.0, perturb=0.0)
    R = case.radius

    def force(eps, vx):
        xs = np.array([0.4, 0.4 + 2 * R * (1 + eps / 2) + 2 * R * eps / 2])
        o = Oracle(case).initialize([np.zeros((34, 34, 34))] * 5, (xs, np.array([0.5, 0.5]), np.array([0.5, 0.5])),
                                    markers_for(case))
        o.pput("U", np.array([vx, -vx]))
        o.set_dt(1e-14)
        o.intgr_nwtn_eulr(1)
        return o.pget("COLFX")[0]

    f_near = force(0.005, 0.1)      # approaching: force on sphere 0 points away (-x)
    assert f_near < 0.0
    f_onset = force(0.0249999, 0.1)
    assert abs(f_onset) < 1e-3 * abs(f_near)


# --------------------------------------------------------------- inputs -----
def test_marker_count_formula():
    assert nl_formula(64, 1.0, 4 / 64) == 173 == pyoracle.load().orc_nl_formula(64, 1.0, 4 / 64)
    assert nl_formula(256, 1.0, 6 / 256) == 409
    assert nl_formula(512, 1.0, 8 / 512) == 746


def test_post3d_closed_forms():
    """post3d quantities (src/output.f90:997-1030, src/post.f90:15-61) of the
    oracle on analytic fields: a linear shear u = a z has voy = a and vox =
    voz = 0; a solid-body rotation u = -W (y - y0), v = W (x - x0) has
    voz = 2 W; vez and con average w and c to the cell centre."""
    from interparts_mpich_amd.inputs import Case
    from oracle.pyoracle import Oracle
    case = Case(name="p3", itot=12, jtot=10, ktot=8, lx=1.2, ly=1.0, lz=0.8, np=0)
    o = Oracle(case)
    nx, ny, nz = case.itot, case.jtot, case.ktot
    dx, dy, dz = case.lx / nx, case.ly / ny, case.lz / nz
    k = np.arange(nz + 2)[:, None, None]
    j = np.arange(ny + 2)[None, :, None]
    i = np.arange(nx + 2)[None, None, :]
    shape = (nz + 2, ny + 2, nx + 2)
    a, W = 0.7, 1.3
    zc = (k - 0.5) * dz + 0 * i * j
    o.put("U", a * zc)
    o.put("V", np.zeros(shape))
    o.put("W", np.zeros(shape))
    for name, want in (("voy", a), ("vox", 0.0), ("voz", 0.0)):
        got = o.post3d(name)
        assert np.allclose(got, want, rtol=0, atol=1e-12), name
    # solid-body rotation: u lives on x faces at (i dx, (j-0.5) dy), v on y faces
    x0, y0 = 0.37, 0.41
    o.put("U", -W * (((j - 0.5) * dy) - y0) + 0 * i * k)
    o.put("V", W * ((i * dx) - x0) + 0 * j * k)
    assert np.allclose(o.post3d("voz"), 2 * W, rtol=0, atol=1e-12)
    rng = np.random.default_rng(3)
    w, c, p = rng.standard_normal(shape), rng.standard_normal(shape), rng.standard_normal(shape)
    o.put("W", w); o.put("C", c); o.put("PNEW", p)
    assert np.array_equal(o.post3d("vez"), 0.5 * (w[1:-1, 1:-1, 1:-1] + w[:-2, 1:-1, 1:-1]))
    assert np.array_equal(o.post3d("con"), 0.5 * (c[1:-1, 1:-1, 1:-1] + c[:-2, 1:-1, 1:-1]))
    assert np.array_equal(o.post3d("pre"), p[1:-1, 1:-1, 1:-1])
