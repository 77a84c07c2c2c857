/*
 * ipg.h -- C-ABI of the MI355X-native InterPartS RK3 substep (libipg.so).
 *
 * One context holds one case resident in HBM: the staggered fields with one
 * ghost layer, the IBM neighbour lists, the Poisson spectra, the particle and
 * marker state.  Every entry point returns 0 on success and a non-zero code on
 * failure (ipg_last_error() gives the message); the Fortran layer turns a
 * non-zero code into `error stop`, as the reference aborts inside the routine.
 *
 * The entry points replace the Fortran module procedures of
 * ajayhudge/InterPartS_MPICH named beside each declaration (file:line of the
 * reference).  The reference's procedures work on module globals; here the
 * globals live in the context, so the array arguments of the reference become
 * the context handle plus field selectors.  Host arrays handed over at
 * upload/download use the reference's layout: real(8) (0:i1,0:j1,0:k1),
 * i fastest.
 */
#ifndef IPG_H
#define IPG_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ipg_ctx ipg_ctx;

/* &PARAMETERS namelist (src/param.f90:66-86) plus the compile-time sizes of
 * src/gridsize.f90:6-26, which are runtime values here. */
typedef struct ipg_params {
  int itot, jtot, ktot;          /* grid (src/gridsize.f90:6-8) */
  int np;                        /* number of spheres (src/gridsize.f90:18) */
  int nl;                        /* markers per sphere; must equal nl of src/init.f90:129 */
  double lx, ly, lz;
  double visc, pran, beta, radius, ratiorho;
  int particle_rotation;         /* logical particle_rotation */
  int bc_uvw_top, bc_uvw_bot;    /* 0 'BC_NOSLIP', 1 'BC_FREESLIP' */
  int bc_c_top, bc_c_bot;        /* 0 'BC_DIRICHLET', 1 'BC_NEUMANN' */
  double bc_c_top_val, bc_c_bot_val;
  /* 0: FP64 path, every value and step FP64 in the reference's order.
   * 1: mixed precision: u, v, w, pnew, c, dudt..dcdt, the particles, the
   * sphere integrals and the IBM FP64; the RK history, the Poisson spectra
   * and the pressure correction's pieces stored FP32 (all stencil and
   * Poisson arithmetic FP64, with FP64-residual refinement of the solve).
   * The environment variable IPG_MIXED_IBM=32 at init makes the IBM compute
   * in FP32.  itot and jtot in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384,
   * 512}. */
  int precision;
} ipg_params;

/* field selectors for ipg_field_put / ipg_field_get */
enum {
  IPG_U = 0, IPG_V, IPG_W, IPG_PNEW, IPG_C,         /* unew vnew wnew pnew cnew (src/common.f90:4) */
  IPG_DUDT, IPG_DVDT, IPG_DWDT, IPG_DCDT,           /* dudt dvdt dwdt dcdt (src/common.f90:7) */
  IPG_DURK, IPG_DVRK, IPG_DWRK, IPG_DCRK,           /* durkold.. (src/main.f90:66) */
  IPG_P,                                            /* correction pressure p (src/main.f90:67) */
  IPG_NFIELDS
};

/* particle quantity selectors (type particle, src/common.f90:35-71) */
enum {
  IPG_PX = 0, IPG_PY, IPG_PZ, IPG_PTHETA, IPG_PPHI, IPG_PU, IPG_PV, IPG_PW, IPG_POMX, IPG_POMY, IPG_POMZ,
  IPG_POMTHETA, IPG_PINTU, IPG_PINTV, IPG_PINTW, IPG_PINTOMX, IPG_PINTOMY, IPG_PINTOMZ, IPG_PCOLFX,
  IPG_PCOLFY, IPG_PCOLFZ, IPG_PCOLTX, IPG_PCOLTY, IPG_PCOLTZ, IPG_PFXLTOT, IPG_PFYLTOT, IPG_PFZLTOT,
  IPG_PTORQX, IPG_PTORQY, IPG_PTORQZ, IPG_PTORQTHETA,
  /* rkp%fxltot..fzltot, rkp%dudt..dwdt: the sums over one step's substeps that
   * sumrk3 keeps for outputpart (src/intgr_nwtn_eulr.f90:1586-1669,
   * src/output.f90:1184-1189) */
  IPG_PRKFX, IPG_PRKFY, IPG_PRKFZ, IPG_PRKDUDT, IPG_PRKDVDT, IPG_PRKDWDT,
  IPG_NPART
};

/* ---- context -------------------------------------------------------- */
/* Replaces initmpi + input_parameters + derived_parameters
 * (src/initmpi.f90:12, src/init.f90:10-243) for one rank holding the domain. */
int ipg_init(const ipg_params *prm, ipg_ctx **ctx);
int ipg_finalize(ipg_ctx *ctx);
const char *ipg_last_error(void);
/* nl = nint((pi/3)(12((radius-0.3/dxi)dxi)^2+1)), src/init.f90:127-133 */
int ipg_nl_formula(int itot, double lx, double radius);
/* kerneltest (src/kernel.f90:43-63), the start-up check of src/main.f90:142-145:
 * the integral of kernel(x) kernel(y) kernel(z) over a 51^3 grid of spacing
 * 1/12, in the reference's summation order.  Host only; needs no context. */
int ipg_kerneltest(double *sum);

/* ---- state transfer (host <-> HBM) ----------------------------------- */
/* host: real(8) (0:i1,0:j1,0:k1) Fortran order, (itot+2)(jtot+2)(ktot+2) values */
int ipg_field_put(ipg_ctx *ctx, int field, const double *host);
int ipg_field_get(ipg_ctx *ctx, int field, double *host);
/* host: np values of one particle quantity */
int ipg_particles_put(ipg_ctx *ctx, int quantity, const double *host);
int ipg_particles_get(ipg_ctx *ctx, int quantity, double *host);
/* dt and time (src/common.f90:9) */
int ipg_set_dt(ipg_ctx *ctx, double dt);
int ipg_get_dt(ipg_ctx *ctx, double *dt);

/* ---- set-up ---------------------------------------------------------- */
/* src/initsolver.f90:10-67: tridiagonal coefficients and eigenvalues */
int ipg_initsolver(ipg_ctx *ctx);
/* src/initparticles.f90:159-241: positions given, velocities zero, neighbour flags */
int ipg_initparticles(ipg_ctx *ctx, const double *x, const double *y, const double *z);
/* src/coordsfp.f90:10-106: marker angles (thetarc, phirc, radfp columns of
 * lagrangianforcepoints2), dVlagr, dVeul */
int ipg_coordsfp(ipg_ctx *ctx, const double *thetarc, const double *phirc, double radfp);

/* ---- the RK3 substep, one entry per reference procedure -------------- */
int ipg_rk(ipg_ctx *ctx, int rkiter);                 /* rk1/rk2/rk3, src/rk3.f90:13,108,224 */
int ipg_ibm(ipg_ctx *ctx);                            /* IBM block, src/main.f90:286-338:
                                                         eulr2lagr src/interp_spread.f90:12,
                                                         complagrforces/updtlagrforces src/forcing.f90:9,189,
                                                         lagr2eulr src/interp_spread.f90:230 */
/* the IBM block split as the reference calls it (src/main.f90:299-335):
 * ipg_ibm_begin stands where the copy dudtf = dudt is (main.f90:301-303) and
 * builds the sphere neighbour lists of this substep; per pass
 * ibmiter = 0..2: ipg_eulr2lagr (interp_spread.f90:12; for ibmiter > 0 it
 * first rebuilds the previous pass's spread around each sphere, and it also
 * updates the marker forces), ipg_lagrforces (the per-sphere sums of
 * complagrforces forcing.f90:9 / updtlagrforces forcing.f90:189),
 * ipg_lagr2eulr (interp_spread.f90:230: for ibmiter 0 and 1 nothing is
 * launched, as the next ipg_eulr2lagr rebuilds that spread; for ibmiter 2 the
 * spread is added to dudt, replacing main.f90:332-334).  Calls must come in
 * this order within one substep. */
int ipg_ibm_begin(ipg_ctx *ctx);
int ipg_eulr2lagr(ipg_ctx *ctx, int ibmiter);
int ipg_lagrforces(ipg_ctx *ctx, int ibmiter);
int ipg_lagr2eulr(ipg_ctx *ctx, int ibmiter);
/* updtintermediatevel (src/forcing.f90:393-513, called after every IBM pass,
 * src/main.f90:315): the values it appends to IBM_forcing.out, forcextot,
 * forceytot, forceztot = the sums over the spheres of fxltot.. after each pass
 * of the last IBM block: tot[3 ibmiter + component], 9 values.  Its other
 * file, meanforces, holds the mean of forcex.., arrays the reference never
 * writes (zero).  Returns 3 when the device error word is set. */
int ipg_ibm_forces(ipg_ctx *ctx, double *tot);
int ipg_bounduvw(ipg_ctx *ctx, int which);            /* src/bound.f90:11; which 0 = unew.., 1 = dudt.. */
int ipg_boundp(ipg_ctx *ctx, int which);              /* src/bound.f90:82; which 0 = pnew, 1 = p */
int ipg_boundscal(ipg_ctx *ctx);                      /* src/bound.f90:109 on cnew */
int ipg_fillps(ipg_ctx *ctx, int rkiter);             /* src/fillps.f90:8 */
int ipg_solver2d(ipg_ctx *ctx);                       /* src/solver.f90:43 */
int ipg_correc(ipg_ctx *ctx, int rkiter);             /* src/correc.f90:11 */
int ipg_pnew_update(ipg_ctx *ctx);                    /* src/main.f90:368-376 */
int ipg_intgr_over_sphere(ipg_ctx *ctx, int type);    /* src/intgr_over_sphere.f90:12 */
int ipg_intgr_nwtn_eulr(ipg_ctx *ctx, int rkiter);    /* src/intgr_nwtn_eulr.f90:12 */
/* src/main.f90:268-414, whole substep.  Returns with the stream running.  The
 * closing Newton-Euler update (src/main.f90:380-405) is left pending and runs
 * beside the next substep's RK stage, or first thing in any other entry point
 * that reads or writes the particle state, dt or u, v, w (ipg_sync included);
 * with ipg_use_side_stream(ctx, 0) it closes the substep instead. */
int ipg_substep(ipg_ctx *ctx, int rkiter);
int ipg_chkdt(ipg_ctx *ctx, double *dtmax);           /* src/chkdt.f90:6 */
int ipg_chkdiv(ipg_ctx *ctx, double *divtot, double *divmax); /* src/chkdiv.f90:7 */
/* ipg_chkdiv's check (src/chkdiv.f90:23-39) on caller-owned device arrays u,
 * v, w in the reference's real(8) (0:i1,0:j1,0:k1) layout, i fastest, for a
 * grid of at most 512^3 cells (itot*jtot*ktot <= 134217728; larger grids
 * return 2).  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (serving loops, harnesses that forbid hipMalloc):
 * it allocates nothing, its partial sums living in a fixed 256 KiB array on
 * each device.  It runs on the null stream and returns when done; calls that
 * overlap on one GPU share that array and race. */
int ipg_chkdiv_noalloc(const double *u, const double *v, const double *w, int itot, int jtot, int ktot, double lx,
                       double ly, double lz, double *divtot, double *divmax);

/* ---- output ---------------------------------------------------------- */
/* post3d quantities (src/output.f90:997-1074, vorticity src/post.f90:15-61) at
 * the cell centres, interior only, i fastest (itot*jtot*ktot values).
 * IPG_POST_SOL is the file 'sol': 1 minus the pressure-cell phase indicator of
 * src/phase_indicator.f90:11-559 with itype = 1 (src/output.f90:1068-1074);
 * IPG_POST_VOZ is computed by the reference and not written. */
enum { IPG_POST_PRE = 0, IPG_POST_VEZ, IPG_POST_CON, IPG_POST_VOX, IPG_POST_VOY, IPG_POST_VOZ, IPG_POST_SOL };
int ipg_post3d(ipg_ctx *ctx, int quantity, double *host);

/* ---- seeded inputs (replaces random_number in src/init.f90:287-305,
 * src/initparticles.f90:36-47): U[0,1) = splitmix64(seed*K1 + stream*K2 + n)
 * >> 11 * 2^-53 for n = first .. first+count-1 ---------------------------- */
int ipg_uniform(unsigned long long seed, unsigned long long stream, unsigned long long first, long long count,
                double *out);

/* ---- runtime controls (no reference counterpart) --------------------- */
int ipg_sync(ipg_ctx *ctx);                           /* finish all pending work and wait for it */
int ipg_use_graphs(ipg_ctx *ctx, int on);             /* capture each substep as a hipGraph */
/* run a substep's Newton-Euler update beside the next substep's RK stage on a
 * second stream (default on); off: the update closes its substep.  Turn it off
 * when many contexts share the GPU: their streams already overlap. */
int ipg_use_side_stream(ipg_ctx *ctx, int on);
/* mixed precision: the Poisson solve, 32 (default) with its spectra stored
 * FP32 between the passes, a pressure correction in FP32 pieces and
 * FP64-residual refinement, or 64 the FP64 path's solve (IPG_MIXED_SOLVE=64
 * at init) */
int ipg_set_mixed_solve(ipg_ctx *ctx, int bits);
int ipg_mixed_solve(ipg_ctx *ctx, int *bits);         /* 64, 32, or 0 for an FP64 context */
/* mixed precision, FP32-stored solve: FP64-residual refinement sweeps, 0..2
 * (default 1, or IPG_MIXED_REFINE at init) */
int ipg_set_refinement(ipg_ctx *ctx, int nref);
/* time the RK stage for 8..64 z planes per workgroup and keep the fastest;
 * overwrites dudt..dcdt (recomputed by every stage): call between substeps */
int ipg_autotune(ipg_ctx *ctx, int *kb);
/* the launch shapes in use: RK z planes per workgroup (0: chosen per grid),
 * sphere-integral blocks (0: one per sphere; with the side stream on and more
 * spheres than CUs, one per CU striding over them, or IPG_SPH_GRID at init) */
int ipg_tuning(ipg_ctx *ctx, int *rk_kb, int *sphere_blocks);
/* HBM bytes one substep must move by the algorithmic model, for roofline
 * (after ipg_initsolver) */
int ipg_substep_bytes(ipg_ctx *ctx, double *bytes);
/* one eager substep with a HIP event after every kernel launch: per-launch
 * durations (ms) and kernel names, at most maxn, count in *n */
int ipg_substep_phases(ipg_ctx *ctx, int rkiter, int maxn, double *ms, const char **names, int *n);

/* ---- batched ensemble (no reference counterpart: the reference runs one
 * case per MPI job) ----------------------------------------------------
 * A batch advances independent cases -- contexts with their own grids, sphere
 * counts, parameters and precision -- through each substep together: every
 * kernel of a substep stage runs once for all cases (once per FFT length),
 * over a merged index space in which each case owns its blocks and reads its
 * own launch arguments; a substep of all cases is one captured hipGraph.  A
 * case's results are those of ipg_substep on the case alone (bit for bit
 * where ipg_substep is deterministic; the spreads' atomics order their sums
 * freely).  The cases stay usable on their own: entry points on a case are
 * ordered after the batch's work on it, and the batch's after theirs.  The
 * cases must start every batch substep at the same c/dcdt buffer parity (the
 * same number of substeps run, or an even difference), the Newton-Euler
 * update closes each substep, and itot, jtot must have compiled FFT plans
 * (16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512).  The batch refers to
 * its cases: destroy it before any of them is finalized. */
typedef struct ipg_batch ipg_batch;
int ipg_batch_create(ipg_ctx *const *cases, int ncases, ipg_batch **out);
int ipg_batch_destroy(ipg_batch *batch);
/* one RK3 substep (rkiter 1..3) of every case */
int ipg_batch_substep(ipg_batch *batch, int rkiter);
/* finish the batch's work and check every case's device error word */
int ipg_batch_sync(ipg_batch *batch);
/* replay each substep from a captured graph (default on) */
int ipg_batch_use_graphs(ipg_batch *batch, int on);
/* split the cases into ngroups groups of about equal cell counts, each group's
 * merged launches on a stream of its own, side by side in the substep's graph
 * (default 1; the tuner chooses); ipg_batch_groups reads the number back */
int ipg_batch_set_groups(ipg_batch *batch, int ngroups);
int ipg_batch_groups(ipg_batch *batch, int *ngroups);
/* on (default): each case's Newton-Euler update is left to the next batch
 * substep, which runs it on a side stream beside its RK stage, as ipg_substep
 * does with its side stream; off: the update closes its substep.  Either way
 * the next entry point on a case that needs the update runs it first.  The
 * tuner chooses; ipg_batch_carry reads the setting back */
int ipg_batch_set_carry(ipg_batch *batch, int on);
int ipg_batch_carry(ipg_batch *batch, int *on);
/* The batch's tuner: times whole RK3 steps with every case's RK-stage launch
 * shape at 16, 32 and 64 z planes per workgroup and keeps the fastest, then,
 * for every merged launch, one workgroup per block against persistent grids of
 * 1, 2 and 4 workgroups per CU striding over the blocks, and keeps the fastest
 * per launch, then 1, 2 and 4 stream groups and the carried updates on or
 * off, on graph replays of whole steps.
 * The timed steps advance every case (results do not depend on the launch
 * shapes): call it at a time-step boundary (after rkiter 3); *substeps gets
 * the number of substeps it ran (a multiple of 3). */
int ipg_batch_autotune(ipg_batch *batch, int *substeps);
/* merged launches per substep of stage rkiter (after its first substep or the
 * tuner) and how many of them run as persistent grids */
int ipg_batch_launches(ipg_batch *batch, int rkiter, int *nlaunch, int *npersistent);
/* one eager substep with a HIP event after every merged launch: per-launch
 * durations (ms) and stage names, at most maxn, count in *n */
int ipg_batch_phases(ipg_batch *batch, int rkiter, int maxn, double *ms, const char **names, int *n);

#ifdef __cplusplus
}
#endif
#endif
