// ipg_internal.h -- context layout, grid indexing and device helpers shared by
// the HIP translation units of libipg.so.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>
#include <string>
#include <type_traits>
#include "../../include/ipg.h"
#include "ipg_batch.h"

#define IPG_NB 9
#define IPG_NGRAPH 12
#define IPG_NMARK 96

// Grid description passed by value to every field kernel.
// Field layout in HBM: element (i,j,k), i in 0..nx+1, j in 0..ny+1,
// k in 0..nz+1, lives at base[(i-1) + sx*j + sxy*k].  sx is a multiple of 16
// doubles, so each row's first interior element i=1 is 128-byte aligned; the
// ghost i=0 is the last slot of the previous row's padding (base[-1] exists).
struct Grid {
  int nx, ny, nz;
  long sx, sxy;
  double lx, ly, lz, dxi, dyi, dzi, dx, dy, dz;
  double visc, kappa, beta;
  int bc_uvw_top, bc_uvw_bot, bc_c_top, bc_c_bot;
  double bc_c_top_val, bc_c_bot_val;
  __host__ __device__ inline long at(int i, int j, int k) const { return (long)(i - 1) + sx * j + sxy * k; }
};

// Constants of the particle model (src/init.f90:74-243, src/param.f90:95-120).
struct PartConst {
  int np, nl, rotation;
  double radius, ratiorho, volp, mominert, offset, radfp, dVlagr, dVeul;
  double gaccx, gaccy, gaccz;
  double kn_ss, etan_ss, kn_sw, etan_sw, coeff_f;
  double a11a_ini_pp, a11b_ini_pp, a11a_sat_pp, a11b_sat_pp, a11_ini_pw, a11_sat_pw;
  double rkab[3];
};

// Particle state, structure of arrays (np each), quantity order = IPG_P*.
struct PartState {
  double *q[IPG_NPART];
  int *nbf;        // np*9 neighbour flags (slot 0 unused)
};

// Poisson plan (src/initsolver.f90, src/solver.f90).
struct Poisson {
  int nmx;          // nx/2+1 complex x-modes
  long sp;          // spectrum row pitch (complex elements)
  double2 *spec;    // [nz][ny][sp] complex spectrum (float2 in the mixed path)
  double *dtri;     // [nz][ny][sp] Thomas multipliers d
  double *a, *b, *c;// [nz+1] tridiagonal coefficients (1-based)
  double *xrt, *yrt;// eigenvalues by x-mode m (nmx) and y-mode my (ny)
  double2 *twx, *twy; // e^{-2 pi i t/n} tables
  int radx[16], nradx, rady[16], nrady;
  double d_rows;    // rows of d stored and read back, over the (m, |my|) columns
};

// Mixed precision (ipg_params.precision = 1).  Kept FP64: u, v, w, pnew, c,
// the stage increments dudt..dcdt, the particles, the sphere integrals, the
// IBM (unless IPG_MIXED_IBM=32) and every arithmetic step of the stencil and
// the Poisson passes.  Stored FP32: the RK history durk..dcrk, the Poisson
// spectra between passes, the pressure correction as a sum of FP32 pieces
// (the solve and its FP64-residual refinements) and the residual.
constexpr int IPG_MAX_REFINE = 2;

struct ipg_ctx {
  ipg_params prm;
  Grid g;
  PartConst pc;
  hipStream_t stream;
  size_t nfield;                  // doubles per field allocation
  double *slab;                   // one allocation for all fields
  double *f[IPG_NFIELDS];         // field base pointers (see Grid)
  bool mixed;                     // precision 1
  bool ibm32;                     // the IBM computes in FP32 (mixed path with IPG_MIXED_IBM=32)
  bool f32[IPG_NFIELDS];          // field stored FP32 (its slot read as float, same indexing)
  bool solve32;                   // mixed path: the Poisson solve stores its spectra FP32 and refines with FP64
                                  // residuals (the default); else it is the FP64 path's solve (IPG_MIXED_SOLVE=64,
                                  // ipg_set_mixed_solve)
  int nref;                       // FP64-residual refinement sweeps of the FP32-stored Poisson solve (0..2)
  float *pbuf;                    // their allocation
  float *pp[IPG_MAX_REFINE + 2];  // the pressure correction's FP32 pieces 0..nref, then the refinement residual (field indexing)
  bool p_in_pieces;               // p is the sum of the pieces; the P field holds the right-hand side
  // IBM
  double *K1[3];                  // dwdtf(.,.,k1) planes (nx*ny), src/interp_spread.f90:100 (no upper clamp)
  bool ibm_lds_set;
  bool p_ghosts_stale;            // the fused substep did not fill p's ghosts
  bool xcol_stale;                // nor the x-ghost columns of u, v, w, pnew, c
  // launch shape: z planes per RK-stage workgroup (IPG_TUNE_RK_KB at init; 0 = chosen per grid)
  int tune_rk_kb;
  int sph_grid;                   // sphere-integral blocks: -1 automatic (ipg_sphere_blocks), 0 one per sphere,
                                  // n > 0: n blocks (IPG_SPH_GRID at init)
  int ncu;                        // compute units of the device
  bool poisson_split;             // y passes and Thomas in two x-mode ranges on two streams (IPG_POISSON_SPLIT=0: off)
  double *mk;                     // [3][nl] marker offsets (radfp sin cos ...)
  double *fl;                     // [3][np*nl] marker forces fxl,fyl,fzl (after passes 0 and 2)
  double *fl2;                    // [3][np*nl] marker forces after pass 1
  int *nbl, *nbc;                 // IBM neighbour images: [np][64] codes, [np] counts
  int *nml, *nmc;                 // IBM neighbour markers per (sphere, component): [3 np][4096], [3 np]
  int *ibm_err;                   // device error word: IBM lists or boxes (1, 2, 4), collision candidates (8)
  unsigned *ibm_flg;              // per (sphere, component) box: cells its markers' stencils touch
  double *ibm_psum;               // [3 passes][3 np][IBM_WAVES] per-wavefront sums of each pass's marker forces
  size_t ibm_flg_n;               // its length in words
  // particles
  PartState ps;
  double *op;                     // [24][np] state at substep start (src/intgr_nwtn_eulr.f90:77-118)
  double *motion2;                // [9][np] positions, velocities, angular velocities after the first collision pass
  double *red;                    // reduction scratch
  double *dscal;                  // device scalars: [0] dt
  double dt_host;
  Poisson ps2;
  bool solver_ready, markers_ready, use_graphs;
  // one graph per (rkiter, buffer parity, carried Newton-Euler update): the
  // fused substep swaps the roles of the c and dcdt buffers (cnew = dcdt,
  // src/main.f90:375), so the captured pointers alternate, and it may carry
  // the previous substep's particle update on the side stream
  hipGraphExec_t graph[IPG_NGRAPH];
  bool graph_ok[IPG_NGRAPH];
  int cpar;                       // 1 when c and dcdt have swapped buffers an odd number of times
  bool dcdt_stale;                // dcdt's buffer holds the previous c; its value is c's
  // The Newton-Euler update that ends a fused substep (src/main.f90:380-405)
  // reads u, v, w and the particle state only; the next substep's RK stage
  // touches neither, so the update is left pending (ne_pending = its rkiter)
  // and runs on the side stream beside that RK stage.  Every entry point that
  // reads or writes the particle state, dt or u, v, w runs it first.
  int ne_pending;
  bool side_on;                   // carry the update over to the side stream (ipg_use_side_stream)
  hipStream_t side;
  hipEvent_t fork_ev, join_ev;
  // kernel timing (ipg_substep_phases): a marker after every launch while
  // timing is on -- a HIP event on eager launches, or, in a substep replayed
  // from a captured graph, a one-lane kernel storing the device's wall clock
  // (events recorded during a capture are not replayed); begin markers (no
  // name) where a stream starts after waiting on the other
  bool timing, stamping;
  int nev;
  hipEvent_t ev[IPG_NMARK];
  const char *evname[IPG_NMARK];
  int evtag[IPG_NMARK];           // 0: the main stream, 1: the side stream
  unsigned long long *stamps;     // [IPG_NMARK] wall-clock marks
  hipStream_t main_stream;        // the context's stream (c->stream is swapped while the side stream is filled)
  // set while a batch records this context's substep: launches are recorded
  // for the batch's merged launches instead of being issued
  struct IpgRec *rec;
  // counts the changes that alter launch arguments or shapes (marker set-up,
  // solver set-up, the RK-stage launch shape, the refinement sweeps): a batch
  // re-records its merged launches when a case's count moves
  unsigned gen;
  // entry-point calls on the context (check_ctx), and the batch's completion
  // event its stream must wait for before its next work (batch_pending)
  unsigned long calls;
  hipEvent_t batch_ev;
  bool batch_pending;
};

// Every device allocation a context holds outside the Poisson plan (which
// ipg_poisson_free releases), each listed once: release() frees exactly
// these, so none is freed twice.
constexpr int IPG_MAX_ALLOCS = 32 + IPG_NPART;
inline int ipg_ctx_allocations(const ipg_ctx *c, void **out) {
  void *const one[] = {c->slab, c->mk, c->fl, c->fl2, c->nbl, c->nbc, c->nml, c->nmc, c->ibm_err, c->ibm_flg,
                       c->ibm_psum, c->op, c->motion2, c->red, c->dscal, c->ps.nbf, c->K1[0], c->K1[1], c->K1[2],
                       c->pbuf, c->stamps};
  static_assert(sizeof(one) / sizeof(one[0]) + IPG_NPART <= IPG_MAX_ALLOCS, "IPG_MAX_ALLOCS");
  int n = 0;
  for (void *q : one)
    if (q) out[n++] = q;
  for (int q = 0; q < IPG_NPART; q++)
    if (c->ps.q[q]) out[n++] = c->ps.q[q];
  return n;
}

// an FP32 field's base pointer: the same slot and indexing as f[q], float
// elements (element -1, the ghost i = 0, exists)
// The pressure correction comes as FP32 pieces (the mixed path's FP32-stored
// solve with its refinements) rather than as the FP64 P field.
inline bool solve_in_pieces(const ipg_ctx *c) { return c->mixed && c->solve32; }
inline float *f32ptr(ipg_ctx *c, int q) { return (float *)(c->slab + (size_t)q * c->nfield) + 16; }

// a wall-clock mark into stamps[slot] on the context's current stream
void ipg_stamp(ipg_ctx *c, int slot);
// record a timing marker named for the kernel just launched (name null: a
// begin marker)
inline void ipg_mark(ipg_ctx *c, const char *name) {
  if (!c->timing || c->nev >= IPG_NMARK - 1) return;
  if (c->stamping) ipg_stamp(c, c->nev + 1);
  else (void)hipEventRecord(c->ev[c->nev + 1], c->stream);
  c->evname[c->nev] = name;
  c->evtag[c->nev + 1] = c->stream == c->main_stream ? 0 : 1;
  c->nev++;
}

// A launch of this context recorded by a batch: the merged kernels of its body
// (k_batch<K> and its persistent form), its own grid, block and dynamic LDS,
// a copy of its argument struct, and the substep stage it belongs to.
int ipg_rec_add(struct IpgRec *r, BatchFns fns, dim3 grid, dim3 block, size_t lds, const void *args, size_t argsz,
                const char *stage, bool side);
// where the main stream of a recorded substep waits for its side stream
void ipg_rec_join(struct IpgRec *r);

// Launches kernel body K for context c: its single-context kernel `one` on the
// context's stream, marked for timing as `stage`, or, while a batch records
// the context, a record of the launch.
// `one` is the single-context kernel (taking the argument struct), or a
// callable that launches it on the stream it is given (for a kernel that
// keeps its arguments as separate parameters).
#if defined(__HIPCC__)
template <class K, class One>
int ipg_go(ipg_ctx *c, One one, dim3 grid, dim3 block, size_t lds, const typename K::Args &a, const char *stage) {
  if (c->rec) return ipg_rec_add(c->rec, batch_kernels<K>(), grid, block, lds, &a, sizeof(a), stage, c->stream != c->main_stream);
  if constexpr (std::is_invocable_v<One, hipStream_t>) one(c->stream);
  else hipLaunchKernelGGL(one, grid, block, lds, c->stream, a);
  ipg_mark(c, stage);
  return 0;
}
#endif

// a context's part in a batch (ipg_api.hip)
int ipg_batch_record(ipg_ctx *c, int rkiter, int par, bool carried, bool defer, struct IpgRec *rec);
bool ipg_batch_carries(const ipg_ctx *c, int rkiter);
int ipg_batch_enter(ipg_ctx *c, bool carried);
void ipg_batch_leave(ipg_ctx *c, int rkiter, bool deferred, hipEvent_t done);

// error plumbing
void ipg_set_error(const std::string &msg);
#define IPG_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) {                                                             \
      ipg_set_error(std::string(#call) + ": " + hipGetErrorString(e_));                 \
      return 1;                                                                         \
    }                                                                                   \
  } while (0)

// ---- host and device helpers ----------------------------------------------
// Fortran nint: round half away from zero.
__host__ __device__ __forceinline__ int nint_f(double x) { return (int)llround(x); }
__host__ __device__ __forceinline__ double sqd(double x) { return x * x; }

// src/kernel.f90:7-25, Roma et al. 3-point regularized delta times h.
// Both branches share one square root: the argument is selected first, so a
// wavefront whose lanes fall in both branches evaluates one sqrt, not two.
__host__ __device__ __forceinline__ double roma_kernel(double r) {
  const double ar = fabs(r);
  const bool outer = ar > 0.5;
  const double s = sqrt(-(3.0 * sqd(outer ? 1.0 - ar : r)) + 1.0);
  const double w = outer ? (1.0 / 6.0) * ((5.0 - 3.0 * ar) - s) : (1.0 / 3.0) * (1.0 + s);
  return ar > 1.5 ? 0.0 : w;
}
// The same kernel in FP32 (the mixed path's IBM), with the hardware square
// root (1 ulp) on the device rather than the correctly rounded sequence.
__host__ __device__ __forceinline__ float sqrt_f32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_sqrtf(x);
#else
  return sqrtf(x);
#endif
}
__host__ __device__ __forceinline__ float roma_kernel(float r) {
  const float ar = fabsf(r);
  const bool outer = ar > 0.5f;
  const float x = outer ? 1.0f - ar : r;
  const float s = sqrt_f32(-(3.0f * (x * x)) + 1.0f);
  const float w = outer ? (1.0f / 6.0f) * ((5.0f - 3.0f * ar) - s) : (1.0f / 3.0f) * (1.0f + s);
  return ar > 1.5f ? 0.0f : w;
}

// periodic wrap of an index in [-1, n+2] onto 1..n
__device__ __forceinline__ int wrapi(int i, int n) {
  int r = i - 1;
  r = r < 0 ? r + n : (r >= n ? r - n : r);
  return r + 1;
}

// launch helpers implemented per translation unit
// wrapx: read the x halo at its periodic image (the fused substep, whose
// ghost fill leaves the x-ghost columns out) rather than at the ghosts
int ipg_launch_rk(ipg_ctx *c, int rkiter, bool wrapx = false);
int ipg_launch_bound(ipg_ctx *c, int mask);
int ipg_launch_fillps(ipg_ctx *c, int rkiter);
int ipg_launch_correc(ipg_ctx *c, int rkiter);
int ipg_launch_pnew(ipg_ctx *c);
int ipg_launch_p_combine(ipg_ctx *c);
int ipg_launch_resid(ipg_ctx *c, int npieces);
int ipg_launch_chkdt(ipg_ctx *c, double *dtmax);
int ipg_launch_chkdiv(ipg_ctx *c, double *tot, double *mx);
int ipg_poisson_setup(ipg_ctx *c);
int ipg_poisson_free(ipg_ctx *c);
int ipg_launch_poisson(ipg_ctx *c);
int ipg_launch_poisson_ex(ipg_ctx *c, int fused_rkiter);
int ipg_launch_ibm(ipg_ctx *c);
int ipg_ibm_setup(ipg_ctx *c);
int ipg_launch_ibm_begin(ipg_ctx *c);
int ipg_launch_ibm_passes(ipg_ctx *c);
int ipg_launch_ibm_gather(ipg_ctx *c, int pass);
int ipg_launch_ibm_final(ipg_ctx *c);
int ipg_launch_ibm_totals(ipg_ctx *c, int pass);
int ipg_launch_ibm_forces(ipg_ctx *c, double *out9);
int ipg_ibm_check(ipg_ctx *c);
int ipg_launch_post3d(ipg_ctx *c, int q, double *host);
int ipg_launch_phase_sol(ipg_ctx *c, double *out);
int ipg_launch_sphere_integrals(ipg_ctx *c, int type);
int ipg_sphere_blocks(const ipg_ctx *c);
int ipg_launch_nwtn_eulr(ipg_ctx *c, int rkiter);
int ipg_launch_nbflags(ipg_ctx *c);

// bound masks for ipg_launch_bound
enum { BND_UVW_NEW = 1, BND_UVW_DT = 2, BND_PNEW = 4, BND_P = 8, BND_C = 16, BND_NO_XCOL = 32 };
