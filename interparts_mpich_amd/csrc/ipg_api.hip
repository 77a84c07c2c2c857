// ipg_api.hip -- the C-ABI of libipg.so (declared in include/ipg.h): context
// set-up with the reference's derived parameters, host <-> HBM transfers, and
// the substep sequence of src/main.f90:268-414, optionally captured as one
// hipGraph per RK stage.
#include "ipg_internal.h"
#include "ipg_ibm_geom.h"
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

static thread_local std::string g_err;
void ipg_set_error(const std::string &msg) { g_err = msg; }

extern "C" const char *ipg_last_error(void) { return g_err.c_str(); }

int ipg_launch_correc_fused(ipg_ctx *c, int rkiter);

namespace {

double sq(double x) { return x * x; }

// src/init.f90:74-243 derived_parameters and src/param.f90:95-120 constants.
// x**2. is written x*x and x**(-1.) 1/x (gfortran folds both), x**3. pow().
void derived(ipg_ctx *c) {
  const ipg_params &p = c->prm;
  Grid &g = c->g;
  PartConst &pc = c->pc;
  const double pi = std::acos(-1.0);
  g.nx = p.itot; g.ny = p.jtot; g.nz = p.ktot;
  g.lx = p.lx; g.ly = p.ly; g.lz = p.lz;
  g.dxi = p.itot / p.lx; g.dyi = p.jtot / p.ly; g.dzi = p.ktot / p.lz;
  g.dx = 1.0 / g.dxi; g.dy = 1.0 / g.dyi; g.dz = 1.0 / g.dzi;
  g.visc = p.visc; g.kappa = p.visc / p.pran; g.beta = p.beta;
  g.bc_uvw_top = p.bc_uvw_top; g.bc_uvw_bot = p.bc_uvw_bot; g.bc_c_top = p.bc_c_top; g.bc_c_bot = p.bc_c_bot;
  g.bc_c_top_val = p.bc_c_top_val; g.bc_c_bot_val = p.bc_c_bot_val;
  g.sx = ((long)p.itot + 2 + 15) / 16 * 16;
  g.sxy = g.sx * (p.jtot + 2);
  pc.np = p.np; pc.nl = p.nl; pc.rotation = p.particle_rotation;
  pc.radius = p.radius; pc.ratiorho = p.ratiorho;
  pc.offset = (std::sqrt(3.0 * (1.5 * 1.5))) / g.dxi + 0.01 / g.dxi;
  pc.volp = ((4.0 / 3.0) * pi) * std::pow(p.radius, 3.0);
  pc.mominert = ((2.0 / 5.0) * pc.volp) * sq(p.radius);
  pc.gaccx = -9.81 * std::cos(pi / 2.0);
  pc.gaccy = -9.81 * std::cos(pi / 2.0);
  pc.gaccz = -9.81 * std::sin(pi / 2.0);
  pc.rkab[0] = 32.0 / 60.0 + 0.0;
  pc.rkab[1] = 25.0 / 60.0 + (-17.0 / 60.0);
  pc.rkab[2] = 45.0 / 60.0 + (-25.0 / 60.0);
  const double Nstretch = 8.0, dt_estim = 0.05, en = 0.97;
  const double eps_ini_pp = 0.025, eps_sat_pp = 0.001, eps_ini_pw = 0.075, eps_sat_pw = 0.001;
  const double meffn_ss = p.ratiorho * pc.volp / 2.0, meffn_sw = p.ratiorho * pc.volp;
  pc.kn_ss = (sq(pi) + sq(std::fabs(std::log(en)))) * meffn_ss / sq(Nstretch * dt_estim);
  pc.etan_ss = -2.0 * (std::log(en)) * meffn_ss / (Nstretch * dt_estim);
  pc.kn_sw = (sq(pi) + sq(std::fabs(std::log(en)))) * meffn_sw / sq(Nstretch * dt_estim);
  pc.etan_sw = -2.0 * (std::log(en)) * meffn_sw / (Nstretch * dt_estim);
  pc.coeff_f = 6.0 * pi * p.visc * p.radius;
  double a11 = -(1.0 / 4.0 * (1.0 / eps_ini_pp)) + 9.0 / 40.0 * std::log(eps_ini_pp) + 3.0 / 112.0 * eps_ini_pp * std::log(eps_ini_pp);
  pc.a11a_ini_pp = a11 - .995; pc.a11b_ini_pp = -a11 + .350;
  a11 = -(1.0 / 4.0 * (1.0 / eps_sat_pp)) + 9.0 / 40.0 * std::log(eps_sat_pp) + 3.0 / 112.0 * eps_sat_pp * std::log(eps_sat_pp);
  pc.a11a_sat_pp = a11 - .995; pc.a11b_sat_pp = -a11 + .350;
  pc.a11_ini_pw = -(1.0 / eps_ini_pw) + 1.0 / 5.0 * std::log(eps_ini_pw) + 1.0 / 21.0 * eps_ini_pw * std::log(eps_ini_pw) - .9713;
  pc.a11_sat_pw = -(1.0 / eps_sat_pw) + 1.0 / 5.0 * std::log(eps_sat_pw) + 1.0 / 21.0 * eps_sat_pw * std::log(eps_sat_pw) - .9713;
  pc.dVeul = 1.0 / (g.dxi * g.dyi * g.dzi);
}

// Every entry point on a context passes here: work it enqueues on the
// context's stream follows the batch substeps already issued on the case
// (their completion event), and the call is counted, so that the next batch
// substep waits for the context's stream only when something was enqueued
// there since the last one.
bool check_ctx(ipg_ctx *c) {
  if (!c) { ipg_set_error("null context"); return false; }
  c->calls++;
  if (c->batch_pending) {
    c->batch_pending = false;
    const hipError_t e = hipStreamWaitEvent(c->stream, c->batch_ev, 0);
    if (e != hipSuccess) { ipg_set_error(std::string("batch wait: ") + hipGetErrorString(e)); return false; }
  }
  return true;
}

// Runs the Newton-Euler update a fused substep left pending, on the main
// stream, in the order the reference runs it.
int settle(ipg_ctx *c) {
  if (!c->ne_pending) return 0;
  const int rkiter = c->ne_pending;
  c->ne_pending = 0;
  return ipg_launch_nwtn_eulr(c, rkiter);
}

// check_ctx and settle: the entry points that read or write the particle
// state, dt or u, v, w
bool ready(ipg_ctx *c) { return check_ctx(c) && settle(c) == 0; }

// The side stream and its two events, created when a substep first carries
// an update over (never during a capture).  Streams are spread over the
// device's hardware queues in the order they are created; creating the side
// stream with the context would put the main streams of many contexts on
// every other queue.
int ensure_side(ipg_ctx *c) {
  if (c->side) return 0;
  IPG_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  IPG_HIP(hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming));
  IPG_HIP(hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming));
  return 0;
}

int release(ipg_ctx *c) {
  for (int r = 0; r < IPG_NGRAPH; r++)
    if (c->graph_ok[r]) { (void)hipGraphExecDestroy(c->graph[r]); c->graph_ok[r] = false; }
  if (c->fork_ev) (void)hipEventDestroy(c->fork_ev);
  if (c->join_ev) (void)hipEventDestroy(c->join_ev);
  if (c->side) (void)hipStreamDestroy(c->side);
  for (int q = 0; q < IPG_NMARK; q++)
    if (c->ev[q]) { (void)hipEventDestroy(c->ev[q]); c->ev[q] = nullptr; }
  ipg_poisson_free(c);
  void *alloc[IPG_MAX_ALLOCS];
  const int nalloc = ipg_ctx_allocations(c, alloc);
  for (int q = 0; q < nalloc; q++) (void)hipFree(alloc[q]);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  return 0;
}

// One substep, src/main.f90:268-414.  Every launch reads its inputs at the
// source where the reference fills ghosts first:
//  * bounduvw(dudt, dvdt, dwdt) (src/main.f90:340) feeds only fillps, which
//    the forward x transform computes with the boundary values at the source;
//  * boundp(p) (src/main.f90:361) is read at the source by correc; p's own
//    ghosts are filled only when p is read back (ipg_field_get).
// pend > 0: the previous substep's Newton-Euler update (its rkiter) and the
// IBM neighbour lists of the positions it gives, run on the side stream
// beside this substep's RK stage and joined before the IBM passes.  defer: leave this substep's update
// pending instead of launching it at the end.
int substep_launches(ipg_ctx *c, int rkiter, int pend, bool defer) {
  const bool np = c->pc.np > 0;
  // (a batch recording the launches keeps the fork and the join as marks of
  // its own: the records of the side stream, and where the main one joins it)
  if (pend) {
    // the update, then the IBM's neighbour lists of the positions it gives
    if (!c->rec) {
      IPG_HIP(hipEventRecord(c->fork_ev, c->stream));
      IPG_HIP(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    }
    std::swap(c->stream, c->side);
    ipg_mark(c, nullptr);
    const int rc = ipg_launch_nwtn_eulr(c, pend) || (np && ipg_launch_ibm_begin(c));
    std::swap(c->stream, c->side);
    if (rc) return 1;
    if (!c->rec) IPG_HIP(hipEventRecord(c->join_ev, c->side));
  }
  if (ipg_launch_rk(c, rkiter, true)) return 1;
  if (pend) {
    if (c->rec) {
      ipg_rec_join(c->rec);
    } else {
      IPG_HIP(hipStreamWaitEvent(c->stream, c->join_ev, 0));
      ipg_mark(c, nullptr);
    }
  }
  if (np && (pend ? ipg_launch_ibm_passes(c) : ipg_launch_ibm(c))) return 1;
  if (ipg_launch_poisson_ex(c, rkiter)) return 1;
  if (ipg_launch_correc_fused(c, rkiter)) return 1;
  // cnew = dcdt (src/main.f90:375): the buffers swap roles, no copy; dcdt's
  // value (equal to cnew) is materialised only if it is read
  std::swap(c->f[IPG_C], c->f[IPG_DCDT]);
  c->cpar ^= 1;
  if (ipg_launch_bound(c, BND_UVW_NEW | BND_PNEW | BND_C | BND_NO_XCOL)) return 1;
  if (np && !defer && ipg_launch_nwtn_eulr(c, rkiter)) return 1;
  // left to the reader: p's ghosts, the x-ghost columns of the fields, dcdt
  c->p_ghosts_stale = true;
  c->xcol_stale = true;
  c->dcdt_stale = true;
  return 0;
}

// dcdt after a fused substep equals cnew (src/main.f90:375): copy it into
// dcdt's buffer before dcdt is read, or before c changes
int fill_dcdt(ipg_ctx *c) {
  if (!c->dcdt_stale) return 0;
  IPG_HIP(hipMemcpyAsync(c->f[IPG_DCDT] - 16, c->f[IPG_C] - 16, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                         c->stream));
  c->dcdt_stale = false;
  return 0;
}

// The x-ghost columns of u, v, w, pnew, c that the fused substep leaves
// unfilled (its readers take the periodic image): filled before any other
// reader sees them.
int fill_xcol(ipg_ctx *c) {
  if (!c->xcol_stale) return 0;
  if (ipg_launch_bound(c, BND_UVW_NEW | BND_PNEW | BND_C)) return 1;
  c->xcol_stale = false;
  return 0;
}

// Captures the substep graph of key r = (rkiter - 1) + 3 parity + 6 carried
// and instantiates it, leaving the host-side state (the c/dcdt buffer roles
// and parity, the stale flags) as it was: capturing records the launches only.
int capture_graph(ipg_ctx *c, int r) {
  const int rkiter = r % 3 + 1, par = (r / 3) % 2, carried = r / 6;
  double *fc = c->f[IPG_C], *fd = c->f[IPG_DCDT];
  const int cpar = c->cpar;
  const bool pgs = c->p_ghosts_stale, xcs = c->xcol_stale, dcs = c->dcdt_stale, pip = c->p_in_pieces;
  if (c->cpar != par) { std::swap(c->f[IPG_C], c->f[IPG_DCDT]); c->cpar ^= 1; }
  const bool defer = c->side_on && c->pc.np > 0;
  hipGraph_t graph;
  IPG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeGlobal));
  const int rc = substep_launches(c, rkiter, carried ? (rkiter == 1 ? 3 : rkiter - 1) : 0, defer);
  const hipError_t e = hipStreamEndCapture(c->stream, &graph);
  c->f[IPG_C] = fc; c->f[IPG_DCDT] = fd; c->cpar = cpar;
  c->p_ghosts_stale = pgs; c->xcol_stale = xcs; c->dcdt_stale = dcs; c->p_in_pieces = pip;
  if (rc) return rc;
  if (e != hipSuccess) { ipg_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return 1; }
  IPG_HIP(hipGraphInstantiate(&c->graph[r], graph, nullptr, nullptr, 0));
  (void)hipGraphDestroy(graph);
  c->graph_ok[r] = true;
  return 0;
}

// Every graph the RK3 sequence can replay, captured together the first time
// one is needed, so that no later substep pays for a capture: 6 (rkiter,
// parity) keys, and 6 more carrying the previous update when it is deferred.
int capture_all_graphs(ipg_ctx *c) {
  const bool defer = c->side_on && c->pc.np > 0;
  if (defer && ensure_side(c)) return 1;
  for (int r = 0; r < (defer ? IPG_NGRAPH : 6); r++)
    if (!c->graph_ok[r] && capture_graph(c, r)) return 1;
  return 0;
}

// After a mixed-path substep p is the sum of its FP32 pieces and the P field
// holds the Poisson right-hand side: the sum is formed before p is read.
int fill_p(ipg_ctx *c) {
  if (!c->p_in_pieces) return 0;
  if (ipg_launch_p_combine(c)) return 1;
  c->p_in_pieces = false;
  c->p_ghosts_stale = true;
  return 0;
}

// Host <-> HBM copy of one field in the reference's (0:i1,0:j1,0:k1) layout;
// an FP32 field (the mixed path's RK history) is converted on the host.
int field_copy(ipg_ctx *c, int field, const double *in, double *out) {
  const Grid &g = c->g;
  const size_t w = (size_t)g.nx + 2, rows = (size_t)(g.ny + 2) * (g.nz + 2);
  if (!c->f32[field]) {
    if (in) IPG_HIP(hipMemcpy2DAsync(c->f[field] - 1, g.sx * sizeof(double), in, w * sizeof(double), w * sizeof(double), rows,
                                     hipMemcpyHostToDevice, c->stream));
    else IPG_HIP(hipMemcpy2DAsync(out, w * sizeof(double), c->f[field] - 1, g.sx * sizeof(double), w * sizeof(double), rows,
                                  hipMemcpyDeviceToHost, c->stream));
    IPG_HIP(hipStreamSynchronize(c->stream));
    return 0;
  }
  std::vector<float> h(w * rows);
  if (in) {
    for (size_t q = 0; q < h.size(); q++) h[q] = (float)in[q];
    IPG_HIP(hipMemcpy2DAsync(f32ptr(c, field) - 1, g.sx * sizeof(float), h.data(), w * sizeof(float), w * sizeof(float), rows,
                             hipMemcpyHostToDevice, c->stream));
    IPG_HIP(hipStreamSynchronize(c->stream));
  } else {
    IPG_HIP(hipMemcpy2DAsync(h.data(), w * sizeof(float), f32ptr(c, field) - 1, g.sx * sizeof(float), w * sizeof(float), rows,
                             hipMemcpyDeviceToHost, c->stream));
    IPG_HIP(hipStreamSynchronize(c->stream));
    for (size_t q = 0; q < h.size(); q++) out[q] = (double)h[q];
  }
  return 0;
}

bool ghosted_field(int field) {
  return field == IPG_U || field == IPG_V || field == IPG_W || field == IPG_PNEW || field == IPG_C;
}

}  // namespace

// ---- a context's part in a batch (ipg_ensemble.hip) -------------------------

// Records the launches of one substep of c as the merged launches of a batch
// will run it: the substep of RK stage rkiter that starts from buffer parity
// par.  With defer, a case with spheres leaves its Newton-Euler update to
// the next substep, which, when carried, runs the previous substep's update on
// a side stream beside its RK stage, as ipg_substep does; without, the update
// closes the substep.  Like a capture, recording leaves the host-side state
// as it was.
int ipg_batch_record(ipg_ctx *c, int rkiter, int par, bool carried, bool defer, IpgRec *rec) {
  double *fc = c->f[IPG_C], *fd = c->f[IPG_DCDT];
  const int cpar = c->cpar;
  const bool pgs = c->p_ghosts_stale, xcs = c->xcol_stale, dcs = c->dcdt_stale, pip = c->p_in_pieces;
  if (c->cpar != par) { std::swap(c->f[IPG_C], c->f[IPG_DCDT]); c->cpar ^= 1; }
  const bool np = c->pc.np > 0;
  c->rec = rec;
  const int rc = substep_launches(c, rkiter, carried && np ? (rkiter == 1 ? 3 : rkiter - 1) : 0, defer && np);
  c->rec = nullptr;
  c->f[IPG_C] = fc; c->f[IPG_DCDT] = fd; c->cpar = cpar;
  c->p_ghosts_stale = pgs; c->xcol_stale = xcs; c->dcdt_stale = dcs; c->p_in_pieces = pip;
  return rc;
}

// Whether c's pending Newton-Euler update is the one a merged substep of RK
// stage rkiter carries (the update of its RK3 predecessor).
bool ipg_batch_carries(const ipg_ctx *c, int rkiter) {
  return c->pc.np <= 0 || c->ne_pending == (rkiter == 1 ? 3 : rkiter - 1);
}

// Before a merged substep: c is set up, and, unless the merged substep
// carries it, an update left pending has run (on its stream, which then
// counts as used).
int ipg_batch_enter(ipg_ctx *c, bool carried) {
  if (!c) { ipg_set_error("null context"); return 1; }
  if (!c->solver_ready) { ipg_set_error("batch: ipg_initsolver has not been called on a case"); return 1; }
  if (c->pc.np > 0 && !c->markers_ready) { ipg_set_error("batch: ipg_coordsfp has not been called on a case"); return 1; }
  if (carried) {
    c->ne_pending = 0;  // the merged substep runs it
  } else {
    if (c->ne_pending && !check_ctx(c)) return 1;
    if (settle(c)) return 1;
  }
  // the substep rewrites dcdt's interior before swapping it in as c
  c->dcdt_stale = false;
  return 0;
}

// After a merged substep: the host-side state its launches leave, as after a
// graph replay of the context's own substep; done is recorded after the
// merged launches, and the context's next entry point waits for it.
void ipg_batch_leave(ipg_ctx *c, int rkiter, bool deferred, hipEvent_t done) {
  c->batch_ev = done;
  c->batch_pending = true;
  // a deferred update is left to the next merged substep (or to the next
  // entry point on the case that needs it)
  c->ne_pending = deferred && c->pc.np > 0 ? rkiter : 0;
  std::swap(c->f[IPG_C], c->f[IPG_DCDT]);
  c->cpar ^= 1;
  c->p_ghosts_stale = true;
  c->xcol_stale = true;
  c->dcdt_stale = true;
  c->p_in_pieces = solve_in_pieces(c);
}

extern "C" {

int ipg_nl_formula(int itot, double lx, double radius) {
  const double pi = std::acos(-1.0), dxi = itot / lx, retrac = 0.3 / dxi;
  return (int)std::lround((pi / 3.0) * (12.0 * sq((radius - retrac) * dxi) + 1.0));
}

// src/kernel.f90:43-63 (kerneltest): the kernel's integral over a 51^3 grid
// of spacing 1/12, summed in the reference's order (k, j, i loops, each term
// ((((1/12)(1/12))(1/12)) kernel(x)) kernel(y)) kernel(z)); close to 1.
// The kernel is the same function the IBM launches evaluate, run on the host.
int ipg_kerneltest(double *sum) {
  if (!sum) { ipg_set_error("kerneltest: null argument"); return 1; }
  double s = 0.0;
  for (int k = -25; k <= 25; k++) {
    const double cz = k / 12.0;
    for (int j = -25; j <= 25; j++) {
      const double cy = j / 12.0;
      for (int i = -25; i <= 25; i++) {
        const double cx = i / 12.0;
        s = s + ((((1.0 / 12.0) * (1.0 / 12.0)) * (1.0 / 12.0)) * roma_kernel(cx)) * roma_kernel(cy) * roma_kernel(cz);
      }
    }
  }
  *sum = s;
  return 0;
}

int ipg_init(const ipg_params *prm, ipg_ctx **out) {
  if (!prm || !out) { ipg_set_error("ipg_init: null argument"); return 1; }
  *out = nullptr;
  if (prm->itot < 4 || prm->jtot < 4 || prm->ktot < 4) { ipg_set_error("ipg_init: grid too small"); return 1; }
  if (prm->np < 0 || (prm->np > 0 && prm->nl <= 0)) { ipg_set_error("ipg_init: bad np/nl"); return 1; }
  if (prm->np > 0 && prm->nl != ipg_nl_formula(prm->itot, prm->lx, prm->radius)) {
    // src/main.f90:117-133: the run stops when nl /= nlmax
    ipg_set_error("ipg_init: nl does not match the marker count of src/init.f90:129 (nl /= nlmax)");
    return 2;
  }
  if (prm->precision != 0 && prm->precision != 1) { ipg_set_error("ipg_init: precision must be 0 (FP64) or 1 (mixed)"); return 1; }
  if (prm->np > 0) {
    // the sphere-integral box (src/intgr_over_sphere.f90:117-130) is at most
    // 2 R / h + 6 cells wide; its cells are indexed with ten bits per axis
    const double w = 2.0 * prm->radius * std::max(prm->itot / prm->lx, std::max(prm->jtot / prm->ly, prm->ktot / prm->lz)) + 6.0;
    if (!(w < 1024.0)) { ipg_set_error("ipg_init: spheres wider than 1017 grid cells"); return 1; }
  }
  ipg_ctx *c = new ipg_ctx();
  std::memset((void *)c, 0, sizeof(ipg_ctx));
  c->prm = *prm;
  derived(c);
  if (const char *e = std::getenv("IPG_TUNE_RK_KB")) c->tune_rk_kb = std::atoi(e);
  c->sph_grid = -1;
  c->poisson_split = true;
  if (const char *e = std::getenv("IPG_POISSON_SPLIT")) c->poisson_split = std::atoi(e) != 0;
  if (const char *e = std::getenv("IPG_SPH_GRID")) c->sph_grid = std::atoi(e);
  auto fail = [&](int code) { release(c); delete c; return code; };
  c->side_on = true;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { ipg_set_error("stream create failed"); return fail(1); }
  {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      c->ncu = 0;
  }
  c->main_stream = c->stream;
  const Grid &g = c->g;
  c->nfield = 16 + (size_t)g.sxy * (g.nz + 2) + 16;
  c->nfield = (c->nfield + 15) / 16 * 16;
  if (hipMalloc(&c->slab, c->nfield * IPG_NFIELDS * sizeof(double)) != hipSuccess) { ipg_set_error("out of device memory (fields)"); return fail(1); }
  if (hipMemset(c->slab, 0, c->nfield * IPG_NFIELDS * sizeof(double)) != hipSuccess) { ipg_set_error("memset failed"); return fail(1); }
  for (int q = 0; q < IPG_NFIELDS; q++) c->f[q] = c->slab + (size_t)q * c->nfield + 16;
  if (prm->precision == 1) {
    c->mixed = true;
    // the IBM stays FP64 unless IPG_MIXED_IBM=32: its FP32 form spreads with
    // FP32 LDS atomic adds, which stall where the FP64 ones do not (2.5x
    // slower passes at 256^3)
    if (const char *e = std::getenv("IPG_MIXED_IBM")) c->ibm32 = std::atoi(e) == 32;
    // the Poisson solve stores its spectra FP32 and refines with FP64
    // residuals unless IPG_MIXED_SOLVE=64 (the FP64 path's solve, which is
    // faster: the FP32-stored solve needs a second solve to keep max |div u|
    // under 1e-10, and the two take 0.86 against 0.43 ms at 256^3)
    c->solve32 = true;
    if (const char *e = std::getenv("IPG_MIXED_SOLVE")) c->solve32 = std::atoi(e) != 64;
    c->nref = 1;
    if (const char *e = std::getenv("IPG_MIXED_REFINE")) c->nref = std::max(0, std::min(IPG_MAX_REFINE, std::atoi(e)));
    // the RK history is stored FP32, each field in its own slot
    for (int q : {IPG_DURK, IPG_DVRK, IPG_DWRK, IPG_DCRK}) c->f32[q] = true;
    const size_t n32 = (size_t)(IPG_MAX_REFINE + 2) * c->nfield;
    if (hipMalloc(&c->pbuf, n32 * sizeof(float)) != hipSuccess) { ipg_set_error("out of device memory (pressure pieces)"); return fail(1); }
    if (hipMemset(c->pbuf, 0, n32 * sizeof(float)) != hipSuccess) { ipg_set_error("memset failed"); return fail(1); }
    for (int r = 0; r <= IPG_MAX_REFINE + 1; r++) c->pp[r] = c->pbuf + (size_t)r * c->nfield + 16;
  }
  const size_t ncell = (size_t)g.nx * g.ny * g.nz, nplane = (size_t)g.nx * g.ny;
  const int np = prm->np;
  if (np > 0) {
    for (int b = 0; b < 3; b++) {
      if (hipMalloc(&c->K1[b], nplane * sizeof(double)) != hipSuccess) { ipg_set_error("out of device memory"); return fail(1); }
      if (hipMemset(c->K1[b], 0, nplane * sizeof(double)) != hipSuccess) { ipg_set_error("memset failed"); return fail(1); }
    }
    const size_t M = (size_t)np * prm->nl;
    if (hipMalloc(&c->mk, 3 * (size_t)prm->nl * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->fl, 3 * M * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->fl2, 3 * M * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->nbl, (size_t)np * IBM_MAXNB * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->nbc, (size_t)np * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->nml, (size_t)np * 3 * IBM_MAXNM * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->nmc, (size_t)np * 3 * sizeof(int)) != hipSuccess ||
        hipMalloc(&c->ibm_err, sizeof(int)) != hipSuccess ||
        hipMalloc(&c->ibm_psum, 9 * (size_t)np * IBM_WAVES * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->op, 24 * (size_t)np * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->motion2, 9 * (size_t)np * sizeof(double)) != hipSuccess ||
        hipMalloc(&c->ps.nbf, (size_t)np * IPG_NB * sizeof(int)) != hipSuccess) {
      ipg_set_error("out of device memory (particles)"); return fail(1);
    }
    (void)hipMemset(c->fl, 0, 3 * M * sizeof(double));
    (void)hipMemset(c->fl2, 0, 3 * M * sizeof(double));
    (void)hipMemset(c->nbc, 0, (size_t)np * sizeof(int));
    (void)hipMemset(c->nmc, 0, (size_t)np * 3 * sizeof(int));
    (void)hipMemset(c->ibm_err, 0, sizeof(int));
    (void)hipMemset(c->ibm_psum, 0, 9 * (size_t)np * IBM_WAVES * sizeof(double));
    (void)hipMemset(c->ps.nbf, 0, (size_t)np * IPG_NB * sizeof(int));
    for (int q = 0; q < IPG_NPART; q++) {
      if (hipMalloc(&c->ps.q[q], (size_t)np * sizeof(double)) != hipSuccess) { ipg_set_error("out of device memory"); return fail(1); }
      (void)hipMemset(c->ps.q[q], 0, (size_t)np * sizeof(double));
    }
  }
  const size_t nred = 4096;  // reduction scratch of chkdt (1024) and chkdiv (2048)
  if (hipMalloc(&c->red, nred * sizeof(double)) != hipSuccess || hipMalloc(&c->dscal, 16 * sizeof(double)) != hipSuccess) {
    ipg_set_error("out of device memory (scratch)"); return fail(1);
  }
  (void)hipMemset(c->red, 0, nred * sizeof(double));
  (void)hipMemset(c->dscal, 0, 16 * sizeof(double));
  if (hipDeviceSynchronize() != hipSuccess) { ipg_set_error("device sync failed at init"); return fail(1); }
  *out = c;
  return 0;
}

int ipg_finalize(ipg_ctx *c) {
  if (!c) return 0;
  (void)check_ctx(c);  // after the batch substeps issued on the case
  (void)hipStreamSynchronize(c->stream);
  release(c);
  delete c;
  return 0;
}

int ipg_field_put(ipg_ctx *c, int field, const double *host) {
  if (!ready(c)) return 1;
  if (field < 0 || field >= IPG_NFIELDS || !host) { ipg_set_error("field_put: bad field"); return 1; }
  if (field == IPG_P) { c->p_ghosts_stale = false; c->p_in_pieces = false; }
  if (ghosted_field(field) && fill_xcol(c)) return 1;
  if (field == IPG_C && fill_dcdt(c)) return 1;
  if (field == IPG_DCDT) c->dcdt_stale = false;
  return field_copy(c, field, host, nullptr);
}

int ipg_field_get(ipg_ctx *c, int field, double *host) {
  if (!check_ctx(c)) return 1;
  if (field < 0 || field >= IPG_NFIELDS || !host) { ipg_set_error("field_get: bad field"); return 1; }
  if (field == IPG_P && fill_p(c)) return 1;
  if (field == IPG_P && c->p_ghosts_stale) {
    // the fused substep leaves p's ghosts to the reader: boundp(p) now
    if (ipg_launch_bound(c, BND_P)) return 1;
    c->p_ghosts_stale = false;
  }
  if (ghosted_field(field) && fill_xcol(c)) return 1;
  if (field == IPG_DCDT && fill_dcdt(c)) return 1;
  if (field_copy(c, field, nullptr, host)) return 1;
  return ipg_ibm_check(c);
}

int ipg_particles_put(ipg_ctx *c, int quantity, const double *host) {
  if (!ready(c)) return 1;
  if (quantity < 0 || quantity >= IPG_NPART || !host) { ipg_set_error("particles_put: bad quantity"); return 1; }
  if (c->pc.np == 0) return 0;
  IPG_HIP(hipMemcpyAsync(c->ps.q[quantity], host, c->pc.np * sizeof(double), hipMemcpyHostToDevice, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_particles_get(ipg_ctx *c, int quantity, double *host) {
  if (!ready(c)) return 1;
  if (quantity < 0 || quantity >= IPG_NPART || !host) { ipg_set_error("particles_get: bad quantity"); return 1; }
  if (c->pc.np == 0) return 0;
  IPG_HIP(hipMemcpyAsync(host, c->ps.q[quantity], c->pc.np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  return ipg_ibm_check(c);
}

int ipg_set_dt(ipg_ctx *c, double dt) {
  if (!ready(c)) return 1;
  c->dt_host = dt;
  IPG_HIP(hipMemcpyAsync(c->dscal, &c->dt_host, sizeof(double), hipMemcpyHostToDevice, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_get_dt(ipg_ctx *c, double *dt) {
  if (!check_ctx(c) || !dt) return 1;
  *dt = c->dt_host;
  return 0;
}

int ipg_initsolver(ipg_ctx *c) {
  if (!ready(c)) return 1;
  if (c->solver_ready) return 0;
  if (ipg_poisson_setup(c)) return 1;
  c->solver_ready = true;
  c->gen++;
  return 0;
}

int ipg_initparticles(ipg_ctx *c, const double *x, const double *y, const double *z) {
  if (!ready(c)) return 1;
  const int np = c->pc.np;
  if (np == 0) return 0;
  std::vector<double> zero(np, 0.0);
  for (int q = 0; q < IPG_NPART; q++)
    if (ipg_particles_put(c, q, zero.data())) return 1;
  for (int p = 0; p < np; p++) {
    // src/initparticles.f90:173-184: particle outside the domain aborts
    if (x[p] < 0.0 || x[p] > c->g.lx || y[p] < 0.0 || y[p] > c->g.ly || z[p] < c->pc.radius || z[p] > c->g.lz - c->pc.radius) {
      ipg_set_error("initparticles: particle outside the domain");
      return 2;
    }
  }
  if (ipg_particles_put(c, IPG_PX, x) || ipg_particles_put(c, IPG_PY, y) || ipg_particles_put(c, IPG_PZ, z)) return 1;
  if (ipg_launch_nbflags(c)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_coordsfp(ipg_ctx *c, const double *thetarc, const double *phirc, double radfp) {
  if (!ready(c)) return 1;
  const int nl = c->pc.nl;
  if (c->pc.np == 0) return 0;
  std::vector<double> mk(3 * (size_t)nl);
  for (int l = 0; l < nl; l++) {
    // src/coordsfp.f90:49-51 with theta*0. = phi*0. = 0
    mk[l] = radfp * std::sin(0.0 + thetarc[l]) * std::cos(0.0 + phirc[l]);
    mk[nl + l] = radfp * std::sin(0.0 + thetarc[l]) * std::sin(0.0 + phirc[l]);
    mk[2 * nl + l] = radfp * std::cos(0.0 + thetarc[l]);
  }
  IPG_HIP(hipMemcpy(c->mk, mk.data(), mk.size() * sizeof(double), hipMemcpyHostToDevice));
  const double pi = std::acos(-1.0), dxi = c->g.dxi;
  c->pc.radfp = radfp;
  // src/coordsfp.f90:84-89
  c->pc.dVlagr = ((4.0 / 3.0) * pi * std::pow(radfp + 0.5 / dxi, 3.0) - (4.0 / 3.0) * pi * std::pow(radfp - 0.5 / dxi, 3.0)) / (1.0 * nl);
  c->pc.dVeul = 1.0 / (c->g.dxi * c->g.dyi * c->g.dzi);
  if (ipg_ibm_setup(c)) return 1;
  c->markers_ready = true;
  return 0;
}

int ipg_rk(ipg_ctx *c, int rkiter) {
  if (!ready(c)) return 1;
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("rk: rkiter must be 1..3"); return 1; }
  if (fill_xcol(c) || fill_dcdt(c) || ipg_launch_rk(c, rkiter)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_ibm(ipg_ctx *c) {
  if (!ready(c)) return 1;
  if (ipg_launch_ibm(c)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return ipg_ibm_check(c);
}

int ipg_ibm_begin(ipg_ctx *c) {
  if (!ready(c)) return 1;
  if (ipg_launch_ibm_begin(c)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return ipg_ibm_check(c);
}

int ipg_eulr2lagr(ipg_ctx *c, int ibmiter) {
  if (!ready(c)) return 1;
  if (ibmiter < 0 || ibmiter > 2) { ipg_set_error("eulr2lagr: ibmiter must be 0..2"); return 1; }
  if (ipg_launch_ibm_gather(c, ibmiter)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_lagrforces(ipg_ctx *c, int ibmiter) {
  if (!ready(c)) return 1;
  if (ibmiter < 0 || ibmiter > 2) { ipg_set_error("lagrforces: ibmiter must be 0..2"); return 1; }
  if (ipg_launch_ibm_totals(c, ibmiter)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// updtintermediatevel (src/forcing.f90:393-513): what it writes to
// IBM_forcing.out, the totals over the spheres of fxltot, fyltot, fzltot after
// each IBM pass of the last IBM block (fused substep or split entries).
int ipg_ibm_forces(ipg_ctx *c, double *tot) {
  if (!check_ctx(c) || !tot) return 1;
  for (int q = 0; q < 9; q++) tot[q] = 0.0;
  if (c->pc.np <= 0) return 0;
  double *d = c->red;  // reduction scratch, free between the checks that use it
  if (ipg_launch_ibm_forces(c, d)) return 1;
  IPG_HIP(hipMemcpyAsync(tot, d, 9 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  return ipg_ibm_check(c) ? 3 : 0;
}

int ipg_lagr2eulr(ipg_ctx *c, int ibmiter) {
  if (!ready(c)) return 1;
  if (ibmiter < 0 || ibmiter > 2) { ipg_set_error("lagr2eulr: ibmiter must be 0..2"); return 1; }
  // the spreads of passes 0 and 1 are rebuilt by the next eulr2lagr
  if (ibmiter == 2 && ipg_launch_ibm_final(c)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// splitmix64 counter generator of the seeded inputs (identical to
// interparts_mpich_amd/inputs.py uniform())
int ipg_uniform(unsigned long long seed, unsigned long long stream, unsigned long long first, long long count,
                double *out) {
  if (!out || count < 0) { ipg_set_error("uniform: bad arguments"); return 1; }
  const unsigned long long K1 = 0x9E3779B97F4A7C15ULL, K2 = 0xD1B54A32D192ED03ULL;
  for (long long n = 0; n < count; n++) {
    unsigned long long x = seed * K1 + stream * K2 + (first + (unsigned long long)n);
    x = x + K1;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    x = x ^ (x >> 31);
    out[n] = (double)(x >> 11) * (1.0 / 9007199254740992.0);
  }
  return 0;
}

int ipg_bounduvw(ipg_ctx *c, int which) {
  if (!ready(c) || fill_xcol(c)) return 1;
  if (ipg_launch_bound(c, which == 0 ? BND_UVW_NEW : BND_UVW_DT)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_boundp(ipg_ctx *c, int which) {
  if (!ready(c) || fill_xcol(c) || (which != 0 && fill_p(c))) return 1;
  if (ipg_launch_bound(c, which == 0 ? BND_PNEW : BND_P)) return 1;
  if (which != 0) c->p_ghosts_stale = false;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_boundscal(ipg_ctx *c) {
  if (!ready(c) || fill_xcol(c)) return 1;
  if (ipg_launch_bound(c, BND_C)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_fillps(ipg_ctx *c, int rkiter) {
  if (!ready(c)) return 1;
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("fillps: rkiter must be 1..3"); return 1; }
  c->p_in_pieces = false;
  if (ipg_launch_fillps(c, rkiter)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_solver2d(ipg_ctx *c) {
  if (!ready(c) || fill_p(c)) return 1;
  if (!c->solver_ready) { ipg_set_error("solver2d: ipg_initsolver has not been called"); return 1; }
  if (ipg_launch_poisson(c)) return 1;
  // the FP32-stored solve leaves FP32 pieces: p = their sum, as the field
  if (solve_in_pieces(c) && fill_p(c)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_correc(ipg_ctx *c, int rkiter) {
  if (!ready(c) || fill_p(c)) return 1;
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("correc: rkiter must be 1..3"); return 1; }
  if (ipg_launch_correc(c, rkiter)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_pnew_update(ipg_ctx *c) {
  if (!ready(c) || fill_dcdt(c) || fill_p(c)) return 1;
  if (ipg_launch_pnew(c)) return 1;
  if (ipg_launch_bound(c, BND_UVW_NEW | BND_PNEW | BND_C)) return 1;
  c->xcol_stale = false;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_intgr_over_sphere(ipg_ctx *c, int type) {
  if (!ready(c)) return 1;
  if (type < 1 || type > 3) { ipg_set_error("intgr_over_sphere: type must be 1..3"); return 1; }
  if (ipg_launch_sphere_integrals(c, type)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int ipg_intgr_nwtn_eulr(ipg_ctx *c, int rkiter) {
  if (!ready(c)) return 1;
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("intgr_nwtn_eulr: rkiter must be 1..3"); return 1; }
  if (ipg_launch_nwtn_eulr(c, rkiter)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return ipg_ibm_check(c);
}

int ipg_substep(ipg_ctx *c, int rkiter) {
  if (!check_ctx(c)) return 1;
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("substep: rkiter must be 1..3"); return 1; }
  if (!c->solver_ready) { ipg_set_error("substep: ipg_initsolver has not been called"); return 1; }
  if (c->pc.np > 0 && !c->markers_ready) { ipg_set_error("substep: ipg_coordsfp has not been called"); return 1; }
  // the substep rewrites dcdt's interior (RK stage) before swapping it in as
  // c, whose ghosts it then fills: the stale dcdt value is never needed
  c->dcdt_stale = false;
  // the pending update of the previous substep rides along beside this RK
  // stage when it is the RK3 predecessor's (3 -> 1 -> 2 -> 3), the only case
  // the captured graphs hold; any other is run first, on its own
  int pend = c->ne_pending;
  if (pend && pend != (rkiter == 1 ? 3 : rkiter - 1)) {
    if (settle(c)) return 1;
    pend = 0;
  }
  c->ne_pending = 0;
  if (pend && ensure_side(c)) return 1;
  // with the side stream off the update closes the substep, as in the
  // reference (ipg_use_side_stream drops the graphs when it changes)
  const bool defer = c->side_on && c->pc.np > 0;
  if (c->use_graphs) {
    const int r = (rkiter - 1) + 3 * c->cpar + 6 * (pend ? 1 : 0);
    if (!c->graph_ok[r] && capture_all_graphs(c)) return 1;
    if (!c->graph_ok[r] && capture_graph(c, r)) return 1;
    // the replay swaps the roles of c and dcdt as the captured launches did
    std::swap(c->f[IPG_C], c->f[IPG_DCDT]);
    c->cpar ^= 1;
    IPG_HIP(hipGraphLaunch(c->graph[r], c->stream));
    c->p_ghosts_stale = true;
    c->xcol_stale = true;
    c->dcdt_stale = true;
    c->p_in_pieces = solve_in_pieces(c);
  } else if (substep_launches(c, rkiter, pend, defer)) {
    return 1;
  }
  c->ne_pending = defer ? rkiter : 0;
  return 0;
}

// The entries whose results a driver writes out (chkdt, chkdiv, post3d, and
// field_get / particles_get above) synchronise anyway; each also checks the
// device error word, so no output is produced from a substep that dropped
// IBM or collision contributions.
int ipg_chkdt(ipg_ctx *c, double *dtmax) {
  if (!check_ctx(c) || !dtmax || fill_xcol(c)) return 1;
  return ipg_launch_chkdt(c, dtmax) || ipg_ibm_check(c) ? 3 : 0;
}

int ipg_chkdiv(ipg_ctx *c, double *tot, double *mx) {
  if (!check_ctx(c) || !tot || !mx || fill_xcol(c)) return 1;
  return ipg_launch_chkdiv(c, tot, mx) || ipg_ibm_check(c) ? 3 : 0;
}

int ipg_post3d(ipg_ctx *c, int quantity, double *host) {
  if (!check_ctx(c) || !host) return 1;
  if (quantity < 0 || quantity > IPG_POST_SOL) { ipg_set_error("post3d: quantity must be 0..6"); return 1; }
  if (!c->solver_ready) { ipg_set_error("post3d: ipg_initsolver has not been called"); return 1; }
  // the phase indicator reads the particle positions and neighbour flags
  if (quantity == IPG_POST_SOL && !ready(c)) return 1;
  if (fill_xcol(c)) return 1;
  return ipg_launch_post3d(c, quantity, host) || ipg_ibm_check(c) ? 3 : 0;
}

int ipg_sync(ipg_ctx *c) {
  if (!ready(c)) return 1;
  IPG_HIP(hipStreamSynchronize(c->stream));
  return ipg_ibm_check(c);
}

// One substep run as ipg_substep runs it (the previous substep's pending
// update carried beside the RK stage; replayed from a captured graph when
// graphs are on) with a timing marker after every launch: per-launch
// durations (ms) and kernel names.  A launch is timed from the marker before
// it on its own stream.  In a replay the markers are wall-clock stamps,
// captured into a timing copy of the substep's graph.
int ipg_substep_phases(ipg_ctx *c, int rkiter, int maxn, double *ms, const char **names, int *n) {
  if (!check_ctx(c) || !ms || !names || !n) return 1;
  if (rkiter < 1 || rkiter > 3) { ipg_set_error("substep_phases: rkiter must be 1..3"); return 1; }
  if (!c->solver_ready) { ipg_set_error("substep_phases: ipg_initsolver has not been called"); return 1; }
  if (c->pc.np > 0 && !c->markers_ready) { ipg_set_error("substep_phases: ipg_coordsfp has not been called"); return 1; }
  for (int q = 0; q < IPG_NMARK; q++)
    if (!c->ev[q]) IPG_HIP(hipEventCreate(&c->ev[q]));
  if (!c->stamps) IPG_HIP(hipMalloc(&c->stamps, IPG_NMARK * sizeof(unsigned long long)));
  c->dcdt_stale = false;
  int pend = c->ne_pending;
  if (pend && pend != (rkiter == 1 ? 3 : rkiter - 1)) {
    if (settle(c)) return 1;
    pend = 0;
  }
  c->ne_pending = 0;
  if (pend && ensure_side(c)) return 1;
  const bool defer = c->side_on && c->pc.np > 0;
  IPG_HIP(hipStreamSynchronize(c->stream));
  c->timing = true;
  c->stamping = c->use_graphs;
  c->nev = 0;
  c->evtag[0] = 0;
  int rc = 0;
  if (c->use_graphs) {
    // a timing copy of this substep's graph, replayed once
    double *fc = c->f[IPG_C], *fd = c->f[IPG_DCDT];
    const int cpar = c->cpar;
    hipGraph_t graph;
    hipGraphExec_t exec;
    IPG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeGlobal));
    ipg_stamp(c, 0);
    rc = substep_launches(c, rkiter, pend, defer);
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->f[IPG_C] = fc; c->f[IPG_DCDT] = fd; c->cpar = cpar;
    c->timing = c->stamping = false;
    if (rc) return rc;
    if (e != hipSuccess) { ipg_set_error(std::string("graph capture: ") + hipGetErrorString(e)); return 1; }
    IPG_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    std::swap(c->f[IPG_C], c->f[IPG_DCDT]);
    c->cpar ^= 1;
    IPG_HIP(hipGraphLaunch(exec, c->stream));
    IPG_HIP(hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(exec);
    c->p_ghosts_stale = c->xcol_stale = c->dcdt_stale = true;
    c->p_in_pieces = solve_in_pieces(c);
  } else {
    IPG_HIP(hipEventRecord(c->ev[0], c->stream));
    rc = substep_launches(c, rkiter, pend, defer);
    c->timing = false;
    if (rc) return rc;
    IPG_HIP(hipStreamSynchronize(c->stream));
  }
  c->ne_pending = defer ? rkiter : 0;
  std::vector<unsigned long long> st;
  double rate_khz = 0.0;
  if (c->use_graphs) {
    st.resize(c->nev + 1);
    IPG_HIP(hipMemcpy(st.data(), c->stamps, st.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    int dev = 0, r = 0;
    IPG_HIP(hipGetDevice(&dev));
    IPG_HIP(hipDeviceGetAttribute(&r, hipDeviceAttributeWallClockRate, dev));
    rate_khz = r;
  }
  int k = 0;
  for (int q = 0; q < c->nev && k < maxn; q++) {
    if (!c->evname[q]) continue;  // a begin marker
    const int e = q + 1;
    int b = e - 1;
    while (b > 0 && c->evtag[b] != c->evtag[e]) b--;
    float t = 0.f;
    if (c->use_graphs) t = (float)((double)(st[e] - st[b]) / rate_khz);
    else IPG_HIP(hipEventElapsedTime(&t, c->ev[b], c->ev[e]));
    ms[k] = t;
    names[k] = c->evname[q];
    k++;
  }
  *n = k;
  return 0;
}

// Launch-shape tuning: times the RK stage for z-block lengths 8..64 planes
// per workgroup on this grid and keeps the fastest.  It runs stage 3, which
// overwrites dudt..dcdt only (every RK stage recomputes them), so it may be
// called before any substep; captured graphs are rebuilt afterwards.
int ipg_autotune(ipg_ctx *c, int *kb_out) {
  if (!ready(c)) return 1;
  for (int q = 0; q < 2; q++)
    if (!c->ev[q]) IPG_HIP(hipEventCreate(&c->ev[q]));
  IPG_HIP(hipStreamSynchronize(c->stream));
  const int saved = c->tune_rk_kb;
  int best = 0;
  float best_ms = 0.f;
  for (int kb = 8; kb <= 64 && kb <= c->g.nz; kb *= 2) {
    c->tune_rk_kb = kb;
    if (ipg_launch_rk(c, 3, true)) { c->tune_rk_kb = saved; return 1; }  // warm-up
    float t_min = 0.f;
    for (int rep = 0; rep < 3; rep++) {
      IPG_HIP(hipEventRecord(c->ev[0], c->stream));
      if (ipg_launch_rk(c, 3, true)) { c->tune_rk_kb = saved; return 1; }
      IPG_HIP(hipEventRecord(c->ev[1], c->stream));
      IPG_HIP(hipEventSynchronize(c->ev[1]));
      float t = 0.f;
      IPG_HIP(hipEventElapsedTime(&t, c->ev[0], c->ev[1]));
      if (rep == 0 || t < t_min) t_min = t;
    }
    if (best == 0 || t_min < best_ms) { best = kb; best_ms = t_min; }
  }
  c->tune_rk_kb = best > 0 ? best : saved;
  c->dcdt_stale = false;  // the timed stages rewrote dcdt
  c->gen++;
  for (int r = 0; r < IPG_NGRAPH; r++)
    if (c->graph_ok[r]) { (void)hipGraphExecDestroy(c->graph[r]); c->graph_ok[r] = false; }
  if (kb_out) *kb_out = c->tune_rk_kb;
  return 0;
}

int ipg_tuning(ipg_ctx *c, int *rk_kb, int *sphere_blocks) {
  if (!check_ctx(c)) return 1;
  if (rk_kb) *rk_kb = c->tune_rk_kb;
  if (sphere_blocks) *sphere_blocks = ipg_sphere_blocks(c) < c->pc.np ? ipg_sphere_blocks(c) : 0;
  return 0;
}

int ipg_use_graphs(ipg_ctx *c, int on) {
  if (!check_ctx(c)) return 1;
  c->use_graphs = on != 0;
  return 0;
}

// Many contexts sharing the GPU already overlap one another's streams; a
// second stream each, joined by events, only adds waits to the hardware
// queues they share.
int ipg_use_side_stream(ipg_ctx *c, int on) {
  if (!ready(c)) return 1;
  if (c->side_on != (on != 0))
    for (int r = 0; r < IPG_NGRAPH; r++)
      if (c->graph_ok[r]) { (void)hipGraphExecDestroy(c->graph[r]); c->graph_ok[r] = false; }
  c->side_on = on != 0;
  return 0;
}

int ipg_set_refinement(ipg_ctx *c, int nref) {
  if (!ready(c)) return 1;
  if (!c->mixed) { ipg_set_error("set_refinement: the context is not mixed precision"); return 1; }
  if (nref < 0 || nref > IPG_MAX_REFINE) { ipg_set_error("set_refinement: nref must be 0..2"); return 1; }
  if (fill_p(c)) return 1;
  if (nref != c->nref) {
    c->gen++;
    for (int r = 0; r < IPG_NGRAPH; r++)
      if (c->graph_ok[r]) { (void)hipGraphExecDestroy(c->graph[r]); c->graph_ok[r] = false; }
  }
  c->nref = nref;
  return 0;
}

int ipg_set_mixed_solve(ipg_ctx *c, int bits) {
  if (!ready(c)) return 1;
  if (!c->mixed) { ipg_set_error("set_mixed_solve: the context is not mixed precision"); return 1; }
  if (bits != 32 && bits != 64) { ipg_set_error("set_mixed_solve: bits must be 32 or 64"); return 1; }
  // p as the FP64 field before the solve that wrote it changes form
  if (fill_p(c)) return 1;
  if ((bits == 32) != c->solve32) {
    c->gen++;
    for (int r = 0; r < IPG_NGRAPH; r++)
      if (c->graph_ok[r]) { (void)hipGraphExecDestroy(c->graph[r]); c->graph_ok[r] = false; }
  }
  c->solve32 = bits == 32;
  return 0;
}

int ipg_mixed_solve(ipg_ctx *c, int *bits) {
  if (!check_ctx(c) || !bits) return 1;
  *bits = !c->mixed ? 0 : c->solve32 ? 32 : 64;
  return 0;
}

// Algorithmic HBM bytes of one substep, the roofline numerator.
int ipg_substep_bytes(ipg_ctx *c, double *bytes) {
  if (!check_ctx(c) || !bytes) return 1;
  if (!c->solver_ready) { ipg_set_error("substep_bytes: ipg_initsolver has not been called"); return 1; }
  // FP64 words per substep, by kernel: RK stage 43/3 N (13, 17, 13 by stage),
  // forward x FFT with fillps 3 N + Sw, two y passes 2 Sw each, Thomas (the
  // spectrum read, written, re-read and re-written in rows 1 .. nz-1-48, once
  // each way in the last 48 rows, kept in registers, and in row nz; d shared
  // by the y-modes +-ky, written and read up to the row where it stops
  // changing) Sw (4 (K - 48) + 2 * 48 + 2) / nz + 2 d_rows with K = nz-1
  // (4 Sw (nz-1)/nz + 2 Sw / nz when K <= 48), inverse x FFT Sw + N, correc
  // 9 N; Sw = 2 (nx/2+1) ny nz.
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  const double Sw = 2.0 * (c->g.nx / 2 + 1) * (double)c->g.ny * c->g.nz;
  const int K = c->g.nz - 1, T = K > 48 ? 48 : 0;
  const double thomas = Sw * (4.0 * (K - T) + 2.0 * T + 2.0) / c->g.nz + 2.0 * c->ps2.d_rows;
  const double fluid = (43.0 / 3.0 + 3.0 + 1.0 + 9.0) * N + (1.0 + 4.0 + 1.0) * Sw + thomas;
  // IBM: per pass and component the stencil shell (radfp +- 2.5 h) is read
  // once from dudt (+ G) and receives one atomic add; sphere integrals read
  // each sphere's box once per component
  const double np = c->pc.np, R = c->pc.radius, rf = c->pc.radfp, h = c->g.dx;
  const double shell = (4.0 / 3.0) * 3.14159265358979 * (std::pow(rf / h + 2.5, 3.0) - std::pow(std::max(rf / h - 2.5, 0.0), 3.0));
  const double ibm = np * 3.0 * 3.0 * 3.0 * shell;
  const double box = np * 3.0 * std::pow(2.0 * R / h + 5.0, 3.0);
  *bytes = 8.0 * (fluid + ibm + box);
  if (c->mixed && !c->solve32) {
    // mixed path with the FP64 solve: the FP64 figure less the RK history's
    // 4 B per word (4 writes at stage 1, 4 reads and 4 writes at stage 2, 4
    // reads at stage 3: 16/3 words per cell and stage)
    *bytes -= 4.0 * (16.0 / 3.0) * N;
  } else if (c->mixed) {
    // mixed path with the FP32-stored solve, bytes: RK stage 72 N FP64 (u v w c pnew read, dudt..dcdt
    // written) + the FP32 history (4 written, 4 read and written, 4 read by
    // stage) 64/3 N; the first forward x pass reads dudt dvdt dwdt and stores
    // the right-hand side (32 N) and the FP32 spectrum (4 Sw); refinement r
    // reads the right-hand side and r pieces and stores the FP32 residual,
    // which its forward x pass reads before storing the spectrum; per solve
    // two y passes (16 Sw), Thomas (its spectrum words at 4 B, d at 8 B) and
    // the inverse x pass (4 Sw + 4 N); correc reads dudt dvdt dwdt, pnew and
    // the nref+1 pieces, writes u v w pnew
    const int ns = c->nref + 1;
    double b = (72.0 + 64.0 / 3.0) * N + 32.0 * N + 4.0 * Sw;
    for (int r = 1; r <= c->nref; r++) b += (8.0 + 4.0 * r + 4.0) * N + 4.0 * N + 4.0 * Sw;
    b += ns * (16.0 * Sw + 4.0 * (thomas - 2.0 * c->ps2.d_rows) + 16.0 * c->ps2.d_rows + 4.0 * Sw + 4.0 * N);
    b += (64.0 + 4.0 * ns) * N;
    *bytes = b + 8.0 * (ibm + box);
  }
  return 0;
}

}  // extern "C"
