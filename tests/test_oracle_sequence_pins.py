"""More pins of the FP64 CPU oracle, chosen so that a plausible slip in the
restatement fails one of them (the GPU path is held to the oracle bit for bit
or to 1e-12, so a wrong oracle would make it wrong too):

 * substep sequencing (src/main.f90:268-414, src/rk3.f90): on a discrete
   z-eigenmode u(z) with v = w = 0 there is no convection and no divergence,
   so p stays 0 and three substeps apply the RK3 stability polynomial
   1 + z + z^2/2 + z^3/6, z = nu lambda dt.  A wrong coefficient (alpha,
   gamma or zeta of src/param.f90:110-112), a dropped RHS_{k-1} term or a
   stage reading u in place of its predecessor's output fails it.  The scalar
   (src/scal.f90, Dirichlet walls) follows the same polynomial with kappa;
 * pressure-only substep: u = 0 and pnew = phi: the predictor's -alpha dt
   grad(phi) (src/rk3.f90:32-34), fillps' 1/(alpha dt) (src/fillps.f90:15-18),
   solver2d, correc (src/correc.f90) and pnew += p (src/main.f90:370) must
   cancel exactly: u = 0 and pnew constant after the substep.  A sign or a
   scale slip anywhere in the chain fails it;
 * ghost rules (src/bound.f90:11-171) stated as the boundary conditions they
   implement: zero velocity at a no-slip wall face, zero normal derivative at
   a free-slip one, w = 0 on both walls, dp/dz = 0, the Dirichlet value at the
   face, periodic images in x and y;
 * marker stencils (src/interp_spread.f90:76-130): a linear field on each
   staggered component is interpolated exactly at every marker away from the
   walls (the 3-point Roma kernel reproduces linear functions only when the
   stencil is centred as the nint/floor rules centre it, and only on the
   component's own staggered positions); spreading conserves the force's
   first moment (src/interp_spread.f90:329-355);
 * markers (src/coordsfp.f90:49-51): the
   positions x_p + radfp (sin t cos f, sin t sin f, cos t), the rigid-body
   velocity u_p + omega x (x_l - x_p) (src/intgr_nwtn_eulr.f90:1474-1479), the pass-0 force (U_l - u_l)/dt
   (src/forcing.f90:63-65) and the force and torque sums (src/forcing.f90:66-89);
 * the top-wall plane (src/interp_spread.f90:100, no upper clamp for the w
   stencil): spreading conserves the total force only with the k1 = kmax+1
   plane counted, and interpolation and spreading stay adjoint with that plane
   included, so the gather reads what the spread left there.
"""
import math

import numpy as np

from interparts_mpich_amd.inputs import Case, initial_fields, markers_for
from oracle.pyoracle import Oracle


def ghosted_zeros(case):
    return np.zeros((case.ktot + 2, case.jtot + 2, case.itot + 2))


def zero_fields(case):
    return [ghosted_zeros(case) for _ in range(5)]


# ------------------------------------------------------ RK3 sequencing ------
def test_three_substeps_apply_the_rk3_polynomial():
    case = Case(name="rkpoly", itot=8, jtot=6, ktot=24, np=0, re_bulk=50.0, pran=0.7, perturb=0.0)
    o = Oracle(case).initialize(zero_fields(case))
    nz, dzi = case.ktot, case.ktot / case.lz
    k = np.arange(nz + 2, dtype=np.float64)
    out = {}
    for m in (1, 3):
        # sin(m pi (k - 1/2) / nz): its ghosts are minus the wall cells, the
        # no-slip rule for u and the zero-value Dirichlet rule for c
        mode = np.sin(m * math.pi * (k - 0.5) / nz)
        lam = -4.0 * dzi * dzi * math.sin(m * math.pi / (2.0 * nz)) ** 2
        u = ghosted_zeros(case) + mode[:, None, None]
        for n in ("U", "C"):
            o.put(n, u)
        for n in ("V", "W", "PNEW"):
            o.put(n, ghosted_zeros(case))
        o.bounduvw(0)
        o.boundscal()
        o.boundp(0)
        dt = 0.4 / (case.visc * dzi * dzi)
        o.set_dt(dt)
        for rk in (1, 2, 3):
            o.substep(rk)
        for n, diff in (("U", case.visc), ("C", case.visc / case.pran)):
            z = diff * lam * dt
            want = (1.0 + z + z * z / 2.0 + z ** 3 / 6.0) * u[1:-1, 1:-1, 1:-1]
            got = o.get(n)[1:-1, 1:-1, 1:-1]
            err = float(np.max(np.abs(got - want)))
            out[(m, n)] = err
            assert err <= 1e-13, (m, n, err, z)
        for n in ("V", "W", "PNEW"):
            assert float(np.max(np.abs(o.get(n)))) == 0.0, n
    print(out)


# --------------------------------------------------- pressure only ----------
def test_pressure_only_substep_cancels():
    case = Case(name="ponly", itot=12, jtot=10, ktot=14, np=0, re_bulk=100.0, perturb=0.0)
    o = Oracle(case).initialize(zero_fields(case))
    rng = np.random.default_rng(17)
    phi = ghosted_zeros(case)
    phi[1:-1, 1:-1, 1:-1] = rng.standard_normal((case.ktot, case.jtot, case.itot))
    o.put("PNEW", phi)
    o.boundp(0)
    dt = 1e-3
    o.set_dt(dt)
    for rk in (1, 2, 3):
        p0 = o.get("PNEW")
        o.substep(rk)
        alpha = o.scalar(11 + rk)
        gscale = alpha * dt * (case.itot / case.lx) * 2.0 * float(np.max(np.abs(p0)))
        for n in ("U", "V", "W"):
            assert float(np.max(np.abs(o.get(n)))) <= 1e-12 * gscale, (rk, n)
        pn = o.get("PNEW")[1:-1, 1:-1, 1:-1]
        assert float(np.max(np.abs(pn - pn.mean()))) <= 1e-12 * float(np.max(np.abs(p0))), rk
        o.put("PNEW", phi)
        o.boundp(0)


# ------------------------------------------------------ ghost rules ---------
def test_ghost_rules_are_the_boundary_conditions():
    rng = np.random.default_rng(3)
    for top, bot, ctop, cbot in ((0, 0, 0, 0), (1, 0, 1, 0), (0, 1, 0, 1)):
        case = Case(name="bc", itot=6, jtot=5, ktot=7, np=0, perturb=0.0, bc_uvw_top=top, bc_uvw_bot=bot,
                    bc_c_top=ctop, bc_c_bot=cbot, bc_c_top_val=0.7, bc_c_bot_val=-0.3)
        o = Oracle(case)
        f = {n: rng.standard_normal(o.shape) for n in ("U", "V", "W", "PNEW", "C")}
        for n, a in f.items():
            o.put(n, a)
        o.bounduvw(0)
        o.boundp(0)
        o.boundscal()
        g = {n: o.get(n) for n in f}
        nx, ny, nz = case.itot, case.jtot, case.ktot
        for n in ("U", "V"):
            a = g[n]
            for face, kin, kgh, bc in ((0, 1, 0, bot), (1, nz, nz + 1, top)):
                if bc == 0:   # no-slip: the value interpolated to the wall face is zero
                    assert np.all(a[kin][1:-1, 1:-1] + a[kgh][1:-1, 1:-1] == 0.0), (n, face)
                else:         # free-slip: zero normal derivative at the face
                    assert np.array_equal(a[kin][1:-1, 1:-1], a[kgh][1:-1, 1:-1]), (n, face)
            # interior cells are untouched
            assert np.array_equal(a[1:-1, 1:-1, 1:-1], f[n][1:-1, 1:-1, 1:-1]), n
        w = g["W"]
        assert np.all(w[0] == 0.0) and np.all(w[nz] == 0.0)            # w on both wall faces
        p = g["PNEW"]
        assert np.array_equal(p[0], p[1]) and np.array_equal(p[nz + 1], p[nz])   # dp/dz = 0
        c = g["C"]
        for kin, kgh, bc, val in ((1, 0, cbot, -0.3), (nz, nz + 1, ctop, 0.7)):
            if bc == 0:   # Dirichlet: the face value is the mean of the two cells
                np.testing.assert_allclose(0.5 * (c[kin] + c[kgh])[1:-1, 1:-1], val, rtol=0, atol=1e-15)
            else:
                assert np.array_equal(c[kin][1:-1, 1:-1], c[kgh][1:-1, 1:-1])
        # periodic images in x and y over the whole plane, wall ghosts included
        for n, a in g.items():
            assert np.array_equal(a[:, :, 0], a[:, :, nx]) and np.array_equal(a[:, :, nx + 1], a[:, :, 1]), n
            assert np.array_equal(a[:, 0, :], a[:, ny, :]) and np.array_equal(a[:, ny + 1, :], a[:, 1, :]), n


# ------------------------------------------------ marker stencils ------------
def _marker_case():
    case = Case(name="mk", itot=20, jtot=18, ktot=22, np=1, diameter_h=6.0, seed=9, perturb=0.0)
    o = Oracle(case).initialize(zero_fields(case), (np.array([0.5]), np.array([0.5]), np.array([0.5])),
                                markers_for(case))
    return case, o


def _set_markers(o, case, xyz):
    n = xyz.shape[0]
    for q, name in enumerate(("XFP", "YFP", "ZFP")):
        o.mview(name)[:n] = xyz[:, q]
    o.lib.orc_nla(o.s)[0] = n


def test_gather_reproduces_linear_fields_on_the_staggered_grid():
    case, o = _marker_case()
    rng = np.random.default_rng(21)
    hs = (case.lx / case.itot, case.ly / case.jtot, case.lz / case.ktot)
    n = min(case.nl, 150)
    # markers anywhere at least 3 cells from the periodic edges and the walls:
    # every fractional cell offset occurs, on both sides of each half cell
    xyz = np.column_stack([rng.uniform(3 * h, L - 3 * h, n) for h, L in zip(hs, (case.lx, case.ly, case.lz))])
    _set_markers(o, case, xyz)
    kk, jj, ii = np.indices(o.fshape).astype(np.float64) - 1.0   # index of dudtf(i, j, k)
    a, b = 0.3, np.array([1.7, -0.9, 0.55])
    # u at (i h, (j-1/2) h, (k-1/2) h), v at ((i-1/2) h, j h, (k-1/2) h), w at ((i-1/2) h, (j-1/2) h, k h)
    for fname, lname, shift in (("DUDTF", "DUDTL", (0.0, 0.5, 0.5)), ("DVDTF", "DVDTL", (0.5, 0.0, 0.5)),
                                ("DWDTF", "DWDTL", (0.5, 0.5, 0.0))):
        X = [(ii - shift[0]) * hs[0], (jj - shift[1]) * hs[1], (kk - shift[2]) * hs[2]]
        for f in ("DUDTF", "DVDTF", "DWDTF"):
            o.view(f)[...] = 0.0
        o.view(fname)[...] = a + b[0] * X[0] + b[1] * X[1] + b[2] * X[2]
        o.lib.orc_eulr2lagr(o.s)
        want = a + xyz @ b
        got = o.mview(lname)[:n]
        err = float(np.max(np.abs(got - want)))
        assert err <= 1e-13, (fname, err)


def test_spread_conserves_the_first_moment():
    case, o = _marker_case()
    rng = np.random.default_rng(23)
    hs = (case.lx / case.itot, case.ly / case.jtot, case.lz / case.ktot)
    n = 40
    xyz = np.column_stack([rng.uniform(3 * h, L - 3 * h, n) for h, L in zip(hs, (case.lx, case.ly, case.lz))])
    _set_markers(o, case, xyz)
    F = rng.standard_normal((n, 3))
    for q, name in enumerate(("FXL", "FYL", "FZL")):
        o.mview(name)[:n] = F[:, q]
    for f in ("DUDTOLD", "DVDTOLD", "DWDTOLD"):
        o.view(f)[...] = 0.0
    o.lib.orc_lagr2eulr(o.s)
    c0 = (o.scalar(0) / o.scalar(1)) * o.dt
    kk, jj, ii = np.indices(o.fshape).astype(np.float64) - 1.0
    for q, (fname, shift) in enumerate((("DUDTF", (0.0, 0.5, 0.5)), ("DVDTF", (0.5, 0.0, 0.5)),
                                        ("DWDTF", (0.5, 0.5, 0.0)))):
        S = o.view(fname)
        X = [(ii - shift[0]) * hs[0], (jj - shift[1]) * hs[1], (kk - shift[2]) * hs[2]]
        assert abs(float(S.sum()) - c0 * F[:, q].sum()) <= 1e-12 * c0 * np.abs(F[:, q]).sum()
        for d in range(3):
            got = float(np.sum(S * X[d]))
            want = c0 * float(np.sum(F[:, q] * xyz[:, d]))
            assert abs(got - want) <= 1e-12 * c0 * float(np.sum(np.abs(F[:, q]))), (fname, d, got, want)


def test_marker_positions_rigid_velocity_and_forces():
    case = Case(name="rig", itot=24, jtot=24, ktot=24, np=1, diameter_h=8.0, seed=4, perturb=0.0)
    xp, yp, zp = 0.43, 0.57, 0.49
    theta, phi, radfp = markers_for(case)
    o = Oracle(case).initialize(zero_fields(case), (np.array([xp]), np.array([yp]), np.array([zp])),
                                (theta, phi, radfp))
    up, om = np.array([0.3, -0.2, 0.1]), np.array([1.1, -0.4, 0.7])
    for n, v in zip(("U", "V", "W", "OMX", "OMY", "OMZ"), np.concatenate([up, om])):
        o.pput(n, np.array([v]))
    # the marker set-up places the markers and their rigid-body velocities
    o.lib.orc_coordsfp(o.s, np.ascontiguousarray(theta).ctypes.data_as(__import__("ctypes").c_void_p),
                       np.ascontiguousarray(phi).ctypes.data_as(__import__("ctypes").c_void_p), radfp)
    nl = o.nla()[0]
    assert nl == case.nl
    r = radfp * np.column_stack([np.sin(theta) * np.cos(phi), np.sin(theta) * np.sin(phi), np.cos(theta)])
    pos = np.column_stack([o.mview(n)[:nl] for n in ("XFP", "YFP", "ZFP")])
    np.testing.assert_allclose(pos, np.array([xp, yp, zp]) + r, rtol=0, atol=1e-15)
    vel = np.column_stack([o.mview(n)[:nl] for n in ("UL", "VL", "WL")])
    np.testing.assert_allclose(vel, up + np.cross(om, pos - np.array([xp, yp, zp])), rtol=0, atol=1e-14)
    # pass 0 with a fluid at rest: F_l = (U_l - 0) / dt, and the sums
    for n in ("DUDTL", "DVDTL", "DWDTL"):
        o.mview(n)[:nl] = 0.0
    o.lib.orc_lagrforces(o.s, 0)
    F = np.column_stack([o.mview(n)[:nl] for n in ("FXL", "FYL", "FZL")])
    np.testing.assert_allclose(F, vel / o.dt, rtol=1e-15, atol=0)
    dVl = o.scalar(0)
    tot = np.array([o.pget(n)[0] for n in ("FXLTOT", "FYLTOT", "FZLTOT")])
    np.testing.assert_allclose(tot, dVl * F.sum(axis=0), rtol=1e-13)
    torq = np.array([o.pget(n)[0] for n in ("TORQX", "TORQY", "TORQZ")])
    np.testing.assert_allclose(torq, dVl * np.cross(pos - np.array([xp, yp, zp]), F).sum(axis=0), rtol=1e-12,
                               atol=1e-12 * np.abs(torq).max())


# ------------------------------------------------ top-wall k1 plane ----------
def test_top_wall_plane_takes_spread_and_is_gathered():
    case = Case(name="k1", itot=16, jtot=16, ktot=16, np=1, diameter_h=6.0, seed=8, perturb=0.0)
    h, R = case.lz / case.ktot, case.radius
    zp = case.lz - R - 0.05 * h      # top markers at lz - 0.35 h
    o = Oracle(case).initialize(zero_fields(case), (np.array([0.5]), np.array([0.5]), np.array([zp])),
                                markers_for(case))
    nl = o.nla()[0]
    rng = np.random.default_rng(31)
    F = rng.standard_normal(nl)
    o.mview("FZL")[:nl] = F
    for n in ("FXL", "FYL"):
        o.mview(n)[:nl] = 0.0
    for f in ("DUDTOLD", "DVDTOLD", "DWDTOLD"):
        o.view(f)[...] = 0.0
    o.lib.orc_lagr2eulr(o.s)
    c0 = (o.scalar(0) / o.scalar(1)) * o.dt
    S = o.view("DWDTF")
    k1 = case.ktot + 2                           # array index of k = kmax + 1
    inner = S[2:k1, 2:-2, 2:-2]
    top = S[k1, 2:-2, 2:-2]
    assert float(np.max(np.abs(top))) > 0.0
    total = c0 * F.sum()
    # total force: conserved with the k1 plane counted, not without it
    assert abs(float(inner.sum() + top.sum()) - total) <= 1e-12 * c0 * np.abs(F).sum()
    assert abs(float(inner.sum()) - total) > 1e-6 * c0 * np.abs(F).sum()
    # adjointness with a field that is non-zero on the k1 plane too
    w = np.zeros(o.fshape)
    w[2:k1 + 1, 2:-2, 2:-2] = rng.standard_normal((case.ktot + 1, case.jtot, case.itot))
    SF = np.array(S)
    for f in ("DUDTF", "DVDTF", "DWDTF"):
        o.view(f)[...] = 0.0
    o.view("DWDTF")[...] = w
    o.lib.orc_eulr2lagr(o.s)
    Iw = o.mview("DWDTL")[:nl].copy()
    lhs = float(np.sum(w[2:k1 + 1, 2:-2, 2:-2] * SF[2:k1 + 1, 2:-2, 2:-2]))
    rhs = c0 * float(np.sum(F * Iw))
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs), (lhs, rhs)
    # and the k1 plane matters to the gather: without it the two differ
    w2 = w.copy()
    w2[k1] = 0.0
    o.view("DWDTF")[...] = w2
    o.lib.orc_eulr2lagr(o.s)
    assert float(np.max(np.abs(o.mview("DWDTL")[:nl] - Iw))) > 1e-6 * float(np.max(np.abs(Iw)))
