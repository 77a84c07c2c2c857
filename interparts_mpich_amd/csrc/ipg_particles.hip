// ipg_particles.hip -- FP64 particle update of one substep:
//  * rigid-body integrals over the spheres, src/intgr_over_sphere.f90:12-625
//    (one block per sphere, periodic images as in the reference's neighbour
//    slots);
//  * Newton-Euler with Breugem's inner-fluid correction and the soft-sphere
//    collision / lubrication forces, src/intgr_nwtn_eulr.f90:12-714 and
//    src/collisions.f90 (one thread per sphere);
//  * periodic wrap and neighbour flags, src/intgr_nwtn_eulr.f90:702-714,
//    1076-1111.
// All arithmetic is FP64 in the reference's operation order.
#include "ipg_internal.h"
#include <algorithm>

namespace {

constexpr int NOP = 24;  // saved quantities (src/intgr_nwtn_eulr.f90:80-103)
// op layout: x y z theta phi u v w omx omy omz omtheta intu intv intw intomx intomy intomz colfx colfy colfz coltx colty coltz
__constant__ int c_op_src[NOP] = {IPG_PX, IPG_PY, IPG_PZ, IPG_PTHETA, IPG_PPHI, IPG_PU, IPG_PV, IPG_PW,
                                  IPG_POMX, IPG_POMY, IPG_POMZ, IPG_POMTHETA, IPG_PINTU, IPG_PINTV, IPG_PINTW,
                                  IPG_PINTOMX, IPG_PINTOMY, IPG_PINTOMZ, IPG_PCOLFX, IPG_PCOLFY, IPG_PCOLFZ,
                                  IPG_PCOLTX, IPG_PCOLTY, IPG_PCOLTZ};
struct PQ { double *q[IPG_NPART]; };

// image centre of neighbour slot nbrecv (src/intgr_over_sphere.f90:83-109)
__device__ __forceinline__ void image_shift(const Grid &g, int nbrecv, double &x, double &y) {
  switch (nbrecv) {
    case 1: x = x + g.lx; break;
    case 2: x = x + g.lx; y = y - g.ly; break;
    case 3: y = y - g.ly; break;
    case 4: x = x - g.lx; y = y - g.ly; break;
    case 5: x = x - g.lx; break;
    case 6: x = x - g.lx; y = y + g.ly; break;
    case 7: y = y + g.ly; break;
    case 8: x = x + g.lx; y = y + g.ly; break;
    default: break;
  }
}

// types: 1 volume, 2 int u dV, 3 int r x u dV, 23 both 2 and 3 from one sweep
//
// One block per sphere.  Each (image, component) sweep over the bounding box
// of src/intgr_over_sphere.f90:117-130 runs in slices of SPH_SLICE cells and
// two phases per slice: every cell is classified with a few operations (no
// corner inside: skipped, as its weight sum1/sum2 is exactly 0; all corners
// inside the inner radius: weight 1), and the cut cells are queued in LDS; the
// queue is then worked through densely, so that the eight corner square roots
// are evaluated only by the lanes that need them.
constexpr int SPH_THREADS = 512, SPH_SLICE = 16 * SPH_THREADS;
// the body of k_sphere_int for a batch's merged launches
template <int TYPE>
__device__ __forceinline__ void sphere_int_body(const Grid &g, const PartConst &pc, const double *__restrict__ U,
                                                const double *__restrict__ V, const double *__restrict__ W, const PQ &P,
                                                const int *__restrict__ nbf, double *op_save, const VB vb) {
  const int p = vb.x;
  constexpr int type = TYPE;
#include "ipg_sphere_int_body.inc"
}

__global__ __launch_bounds__(SPH_THREADS) void k_sphere_int(Grid g, PartConst pc, const double *__restrict__ U,
                                                            const double *__restrict__ V, const double *__restrict__ W, PQ P,
                                                            const int *__restrict__ nbf, int type, double *op_save) {
  const int p = blockIdx.x;
#include "ipg_sphere_int_body.inc"
}
// the same over a grid smaller than the spheres, each block striding over
// them (the barrier keeps one sphere's LDS reads apart from the next one's
// writes)
__global__ __launch_bounds__(SPH_THREADS) void k_sphere_int_persist(Grid g, PartConst pc, const double *__restrict__ U,
                                                                    const double *__restrict__ V, const double *__restrict__ W,
                                                                    PQ P, const int *__restrict__ nbf, int type,
                                                                    double *op_save) {
  for (int p = blockIdx.x; p < pc.np; p += gridDim.x) {
#include "ipg_sphere_int_body.inc"
    __syncthreads();
  }
}

struct SphereIntArgs {
  Grid g; PartConst pc;
  const double *U, *V, *W;
  PQ P;
  const int *nbf;
  double *op_save;
};
// TYPE (1, 2, 3 or 23) at compile time: each kind of integral its own code
template <int TYPE>
struct SphereIntK {
  using Args = SphereIntArgs;
  static constexpr int THREADS = SPH_THREADS, WPE = 6;  // three blocks per CU, as k_sphere_int
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    sphere_int_body<TYPE>(a.g, a.pc, a.U, a.V, a.W, a.P, a.nbf, a.op_save, v);
  }
};
int launch_sphere_int(ipg_ctx *c, const SphereIntArgs &a, int type, const char *stage) {
  const dim3 blk(SPH_THREADS), grid(ipg_sphere_blocks(c));
  auto one = [&](hipStream_t s) {
    if ((int)grid.x < a.pc.np)
      hipLaunchKernelGGL(k_sphere_int_persist, grid, blk, 0, s, a.g, a.pc, a.U, a.V, a.W, a.P, a.nbf, type, a.op_save);
    else
      hipLaunchKernelGGL(k_sphere_int, grid, blk, 0, s, a.g, a.pc, a.U, a.V, a.W, a.P, a.nbf, type, a.op_save);
  };
  switch (type) {
    case 1: return ipg_go<SphereIntK<1>>(c, one, grid, blk, 0, a, stage);
    case 2: return ipg_go<SphereIntK<2>>(c, one, grid, blk, 0, a, stage);
    case 3: return ipg_go<SphereIntK<3>>(c, one, grid, blk, 0, a, stage);
    case 23: return ipg_go<SphereIntK<23>>(c, one, grid, blk, 0, a, stage);
  }
  ipg_set_error("intgr_over_sphere: type must be 1, 2, 3 or 23");
  return 1;
}

// src/collisions.f90:222-438, normal squeeze force only (lines 430-432)
__device__ __forceinline__ void lubrication(const PartConst &pc, double up, double vp, double wp, bool wall, double nx, double ny,
                                            double nz, double eps, double un, double vn, double wn, double &fx, double &fy,
                                            double &fz) {
  const double u1a = up * nx + vp * ny + wp * nz;
  const double u1b = un * nx + vn * ny + wn * nz;
  double f1;
  if (wall) {
    double a11;
    if (eps < 0.001) a11 = pc.a11_sat_pw;
    else a11 = -(1.0 / eps) + 1.0 / 5.0 * log(eps) + 1.0 / 21.0 * eps * log(eps) - .9713;
    f1 = (a11 * u1a);
    f1 = pc.coeff_f * (f1 - (pc.a11_ini_pw * u1a));
  } else {
    double a11a, a11b;
    if (eps < 0.001) { a11a = pc.a11a_sat_pp; a11b = pc.a11b_sat_pp; }
    else {
      const double a11 = -(1.0 / 4.0 * (1.0 / eps)) + 9.0 / 40.0 * log(eps) + 3.0 / 112.0 * eps * log(eps);
      a11a = a11 - .995; a11b = -a11 + .350;
    }
    f1 = (a11a * u1a + a11b * u1b);
    f1 = pc.coeff_f * (f1 - (pc.a11a_ini_pp * u1a + pc.a11b_ini_pp * u1b));
  }
  fx = fx + f1 * nx; fy = fy + f1 * ny; fz = fz + f1 * nz;
}

// the force one lubrication call adds (fx = fx + term keeps its rounding)
__device__ __forceinline__ void lubrication_term(const PartConst &pc, double up, double vp, double wp, bool wall, double nx,
                                                 double ny, double nz, double eps, double un, double vn, double wn,
                                                 double &tx, double &ty, double &tz) {
  double fx = 0.0, fy = 0.0, fz = 0.0;
  lubrication(pc, up, vp, wp, wall, nx, ny, nz, eps, un, vn, wn, fx, fy, fz);
  tx = fx; ty = fy; tz = fz;  // 0 + t = t exactly
}

// src/collisions.f90:10-220, normal spring-dashpot force only (lines 199-201)
__device__ __forceinline__ void collision(const PartConst &pc, double up, double vp, double wp, double omx, double omy, double omz,
                                          bool wall, double dist, double deltax, double deltay, double deltaz, double un,
                                          double vn, double wn, double omxn, double omyn, double omzn, double &fx,
                                          double &fy, double &fz) {
  const double nx = deltax / dist, ny = deltay / dist, nz = deltaz / dist;
  const double R = pc.radius;
  double deltan, kn, etan;
  if (!wall) { deltan = 2.0 * R - dist; kn = pc.kn_ss; etan = pc.etan_ss; }
  else { deltan = R - dist; kn = pc.kn_sw; etan = pc.etan_sw; }
  const double dun = (up + (omy * R * nz) - (omz * R * ny)) - (un + (omyn * R * (-nz)) - (omzn * R * (-ny)));
  const double dvn = (vp - (omx * R * nz) + (omz * R * nx)) - (vn - (omxn * R * (-nz)) + (omzn * R * (-nx)));
  const double dwn = (wp + (omx * R * ny) - (omy * R * nx)) - (wn + (omxn * R * (-ny)) - (omyn * R * (-nx)));
  const double vabs = dun * nx + dvn * ny + dwn * nz;
  fx = fx + (-(kn * (deltan * nx)) - etan * (vabs * nx));
  fy = fy + (-(kn * (deltan * ny)) - etan * (vabs * ny));
  fz = fz + (-(kn * (deltan * nz)) - etan * (vabs * nz));
}

// copy nb of particle q as seen at one rank (src/intgr_nwtn_eulr.f90:223-322)
__device__ __forceinline__ void nb_copy(const Grid &g, int nb, double &x, double &y) {
  switch (nb) {
    case 1: if (x < g.lx) x = x + g.lx; break;
    case 2: if (x < g.lx) x = x + g.lx; if (y > 0.0) y = y - g.ly; break;
    case 3: if (y > 0.0) y = y - g.ly; break;
    case 4: if (x > 0.0) x = x - g.lx; if (y > 0.0) y = y - g.ly; break;
    case 5: if (x > 0.0) x = x - g.lx; break;
    case 6: if (x > 0.0) x = x - g.lx; if (y < g.ly) y = y + g.ly; break;
    case 7: if (y < g.ly) y = y + g.ly; break;
    case 8: if (x < g.lx) x = x + g.lx; if (y < g.ly) y = y + g.ly; break;
    default: break;
  }
}

// Positions, velocities and angular velocities of the spheres: the state the
// collision pass reads and the position update writes.  The first of the two
// passes of src/intgr_nwtn_eulr.f90:149-680 writes into a second buffer, so
// that the pass reads one state throughout and the next pass reads the
// updated one.
struct Motion { double *x, *y, *z, *u, *v, *w, *ox, *oy, *oz; };

// src/intgr_nwtn_eulr.f90:611-667 for sphere p with collision force (fx, fy,
// fz): the new position, velocity, angular velocity into dst; colf, phi,
// theta, omtheta and the sumrk3 sums into the particle state.
__device__ __forceinline__ void ne_update_one(const PartConst &pc, const PQ &P, const double *__restrict__ op, double fx,
                                              double fy, double fz, double dt, int rkiter, int p, const Motion &dst) {
  const int np = pc.np;
  auto O = [&](int s) { return op[(long)s * np + p]; };
  double *const *q = P.q;
  const double rkab = pc.rkab[rkiter - 1];
  q[IPG_PCOLFX][p] = fx; q[IPG_PCOLFY][p] = fy; q[IPG_PCOLFZ][p] = fz;
  q[IPG_PCOLTX][p] = 0.0; q[IPG_PCOLTY][p] = 0.0; q[IPG_PCOLTZ][p] = 0.0;
  const double vr = pc.volp * pc.ratiorho, rr = pc.ratiorho;
  const double u = O(5) + (1.0 - 0.0) * ((-1.0) * dt * (q[IPG_PFXLTOT][p]) / vr + (q[IPG_PINTU][p] - O(12)) / vr) +
                   rkab * dt * pc.gaccx * (1.0 - (1.0 / rr)) + rkab * 0.5 * dt * (fx + O(18)) / vr;
  const double x = O(0) + rkab * 0.5 * dt * (u + O(5));
  const double v = O(6) + (1.0 - 0.0) * ((-1.0) * dt * (q[IPG_PFYLTOT][p]) / vr + (q[IPG_PINTV][p] - O(13)) / vr) +
                   rkab * dt * pc.gaccy * (1.0 - (1.0 / rr)) + rkab * 0.5 * dt * (fy + O(19)) / vr;
  const double y = O(1) + rkab * 0.5 * dt * (v + O(6));
  const double w = O(7) + (1.0 - 0.0) * ((-1.0) * dt * (q[IPG_PFZLTOT][p]) / vr + (q[IPG_PINTW][p] - O(14)) / vr) +
                   rkab * dt * pc.gaccz * (1.0 - (1.0 / rr)) + rkab * 0.5 * dt * (fz + O(20)) / vr;
  const double z = O(2) + rkab * 0.5 * dt * (w + O(7));
  dst.u[p] = u; dst.x[p] = x; dst.v[p] = v; dst.y[p] = y; dst.w[p] = w; dst.z[p] = z;
  double ox = 0.0, oy = 0.0, oz = 0.0;
  if (pc.rotation) {
    const double mr = pc.mominert * pc.ratiorho;
    ox = O(8) + (-1.0) * dt * q[IPG_PTORQX][p] / mr + (q[IPG_PINTOMX][p] - O(15)) / mr + rkab * 0.5 * dt * (0.0 + O(21)) / mr;
    oy = O(9) + (-1.0) * dt * q[IPG_PTORQY][p] / mr + (q[IPG_PINTOMY][p] - O(16)) / mr + rkab * 0.5 * dt * (0.0 + O(22)) / mr;
    oz = O(10) + (-1.0) * dt * q[IPG_PTORQZ][p] / mr + (q[IPG_PINTOMZ][p] - O(17)) / mr + rkab * 0.5 * dt * (0.0 + O(23)) / mr;
  }
  dst.ox[p] = ox; dst.oy[p] = oy; dst.oz[p] = oz;
  const double phi = O(4) + 0.5 * rkab * dt * (oz + O(10));
  q[IPG_PPHI][p] = phi;
  const double omt = (oy * cos(phi)) - (ox * sin(phi));
  q[IPG_POMTHETA][p] = omt;
  q[IPG_PTHETA][p] = O(3) + rkab * 0.5 * dt * (omt + O(11));
  // sumrk3 (src/intgr_nwtn_eulr.f90:1586-1669), called on both passes of the
  // collision loop: the rkiter = 1 reset happens on each pass.
  const bool first = rkiter == 1;
  q[IPG_PRKFX][p] = (first ? 0.0 : q[IPG_PRKFX][p]) + (-1.0) * q[IPG_PFXLTOT][p] + (q[IPG_PINTU][p] - O(12)) / dt;
  q[IPG_PRKFY][p] = (first ? 0.0 : q[IPG_PRKFY][p]) + (-1.0) * q[IPG_PFYLTOT][p] + (q[IPG_PINTV][p] - O(13)) / dt;
  q[IPG_PRKFZ][p] = (first ? 0.0 : q[IPG_PRKFZ][p]) + (-1.0) * q[IPG_PFZLTOT][p] + (q[IPG_PINTW][p] - O(14)) / dt;
  q[IPG_PRKDUDT][p] = (first ? 0.0 : q[IPG_PRKDUDT][p]) + (u - O(5)) / dt;
  q[IPG_PRKDVDT][p] = (first ? 0.0 : q[IPG_PRKDVDT][p]) + (v - O(6)) / dt;
  q[IPG_PRKDWDT][p] = (first ? 0.0 : q[IPG_PRKDWDT][p]) + (w - O(7)) / dt;
}

// src/intgr_nwtn_eulr.f90:702-714 periodic wrap (when wrap), then the
// neighbour flags of src/intgr_nwtn_eulr.f90:1076-1111 (= src/initparticles.f90:200-241)
// at one rank, for sphere p.
__device__ __forceinline__ void wrap_flags_one(const Grid &g, const PartConst &pc, const PQ &P, int *nbf, int wrap, int p) {
  double x = P.q[IPG_PX][p], y = P.q[IPG_PY][p];
  if (wrap) {
    if (x > g.lx) x = x - g.lx;
    if (x < 0.0) x = x + g.lx;
    if (y > g.ly) y = y - g.ly;
    if (y < 0.0) y = y + g.ly;
    P.q[IPG_PX][p] = x; P.q[IPG_PY][p] = y;
  }
  const double ro = pc.radius + pc.offset;
  const double left = 0.0 * g.lx / 1.0, right = 1.0 * g.lx / 1.0, front = 0.0 * g.ly / 1.0, back = 1.0 * g.ly / 1.0;
  int *f = nbf + p * IPG_NB;
  f[0] = 0;
  f[1] = x > (right - ro);
  f[2] = fabs(sqrt(sqd(right - x) + sqd(front - y))) < ro;
  f[3] = y < (front + ro);
  f[4] = fabs(sqrt(sqd(left - x) + sqd(front - y))) < ro;
  f[5] = x < (left + ro);
  f[6] = fabs(sqrt(sqd(left - x) + sqd(back - y))) < ro;
  f[7] = y > (back - ro);
  f[8] = fabs(sqrt(sqd(right - x) + sqd(back - y))) < ro;
}

constexpr int MAXCAND = 48;

// Collision model, src/intgr_nwtn_eulr.f90:338-606: for each sphere the
// neighbour tables nb = 0..8 (at one rank: the sphere set and its 8 periodic
// copies) are scanned in order; every copy of a partner in range adds its
// lubrication and contact force (the reference's 9x multiplicity at one rank).
// One wavefront per sphere screens all partners (with a margin), keeping the
// candidates in ascending order; lane 0 then runs the reference's ordered scan
// over the candidates only.  Sphere centres are staged through LDS in chunks
// (one latency per chunk rather than one per 64 partners), and the candidates'
// state is gathered into LDS by the wave's lanes before the serial scan.
constexpr int COL_WAVES = 4, COL_CHUNK = 1024;
// LDS written by some lanes of a wavefront is visible to its other lanes
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// One collision pass with the position update of src/intgr_nwtn_eulr.f90:
// 611-672 fused: lane 0 of sphere p's wavefront updates p as soon as its
// force is summed.  The pass reads the state src and writes the update to dst
// (a different buffer), so no wavefront reads a state another one is
// updating; the last pass also wraps the position and sets the neighbour
// flags (src/intgr_nwtn_eulr.f90:702-714, 1076-1111).
__device__ __forceinline__ void collide_body(const Grid &g, const PartConst &pc, const PQ &P, const Motion &src,
                                             const Motion &dst, double *__restrict__ op, const double *__restrict__ dtp,
                                             int rkiter, int last, int *nbf, int *err, const VB vb) {
  __shared__ double cx[COL_CHUNK], cy[COL_CHUNK], cz[COL_CHUNK];
  __shared__ int cand[COL_WAVES][MAXCAND];
  __shared__ double cd[COL_WAVES][MAXCAND][9];
  __shared__ double terms[COL_WAVES][64][6];
  __shared__ int tuse[COL_WAVES][64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int p = vb.x * COL_WAVES + wv;
  const int np = pc.np;
  const double *X = src.x, *Y = src.y, *Zq = src.z, *Uq = src.u, *Vq = src.v, *Wq = src.w;
  const bool live = p < np;
  const double xp = live ? X[p] : 0.0, yp = live ? Y[p] : 0.0, zp = live ? Zq[p] : 0.0;
  const double R = pc.radius, lx = g.lx, ly = g.ly;
  const double reach = 2.0 * R * (1.0 + 0.025) + 1e-9 * (lx + ly + R);
  int nc = 0;
  for (int c0 = 0; c0 < np; c0 += COL_CHUNK) {
    const int n = min(COL_CHUNK, np - c0);
    __syncthreads();
    for (int t = threadIdx.x; t < n; t += COL_WAVES * 64) { cx[t] = X[c0 + t]; cy[t] = Y[c0 + t]; cz[t] = Zq[c0 + t]; }
    __syncthreads();
    if (!live) continue;
    for (int q0 = 0; q0 < n; q0 += 64) {
      const int ql = q0 + lane, q = c0 + ql;
      bool hit = false;
      if (ql < n && q != p) {
        const double dz = cz[ql] - zp;
        double dx = cx[ql] - xp, dy = cy[ql] - yp;
        dx = dx - lx * rint(dx / lx);
        dy = dy - ly * rint(dy / ly);
        hit = dx * dx + dy * dy + dz * dz <= reach * reach;
      }
      unsigned long long mask = __ballot(hit);
      if (hit) {
        const int slot = nc + (int)__popcll(mask & ((1ull << lane) - 1ull));
        if (slot < MAXCAND) cand[wv][slot] = q;
      }
      nc += (int)__popcll(mask);
    }
  }
  if (!live) return;
  wave_lds_sync();
  if (nc > MAXCAND) {
    if (lane == 0) atomicOr(err, 8);
    nc = MAXCAND;
  }
  // the candidates' state, one lane per candidate
  for (int c = lane; c < nc; c += 64) {
    const int q = cand[wv][c];
    double *e = cd[wv][c];
    e[0] = X[q]; e[1] = Y[q]; e[2] = Zq[q]; e[3] = Uq[q]; e[4] = Vq[q]; e[5] = Wq[q];
    e[6] = src.ox[q]; e[7] = src.oy[q]; e[8] = src.oz[q];
  }
  wave_lds_sync();
  const double up = Uq[p], vp = Vq[p], wp = Wq[p];
  const double omx = src.ox[p], omy = src.oy[p], omz = src.oz[p];
  double fx = 0.0, fy = 0.0, fz = 0.0;
  // the terms of the ordered scan (nb outer, candidate inner) one per lane,
  // 64 at a time; lane 0 then adds them in that order, as the reference does
  const int nterm = IPG_NB * nc;
  for (int l0 = 0; l0 < nterm; l0 += 64) {
    const int l = l0 + lane;
    double lt[3] = {0.0, 0.0, 0.0}, ct[3] = {0.0, 0.0, 0.0};
    int use = 0;
    if (l < nterm) {
      const int nb = l / nc, c = l - nb * nc;
      const double *e = cd[wv][c];
      double xq = e[0], yq = e[1];
      nb_copy(g, nb, xq, yq);
      const double deltaz = e[2] - zp;
      const double ax0 = xq - xp, ax1 = xq - lx - xp, ax2 = xq + lx - xp;
      double deltax = ax0;
      if (fabs(ax1) < fabs(deltax)) deltax = ax1;
      if (fabs(ax2) < fabs(deltax)) deltax = ax2;
      const double ay0 = yq - yp, ay1 = yq - ly - yp, ay2 = yq + ly - yp;
      double deltay = ay0;
      if (fabs(ay1) < fabs(deltay)) deltay = ay1;
      if (fabs(ay2) < fabs(deltay)) deltay = ay2;
      const double dist = sqrt(sqd(deltax) + sqd(deltay) + sqd(deltaz));
      const double nx = deltax / dist, ny = deltay / dist, nz = deltaz / dist;
      const double eps = (dist - 2.0 * R) / R;
      if ((eps < 0.025) && (eps > 0.0)) {
        lubrication_term(pc, up, vp, wp, false, nx, ny, nz, eps, e[3], e[4], e[5], lt[0], lt[1], lt[2]);
        use |= 1;
      }
      if (dist < (2.0 * R)) {
        collision(pc, up, vp, wp, omx, omy, omz, false, dist, deltax, deltay, deltaz, e[3], e[4], e[5], e[6], e[7], e[8],
                  ct[0], ct[1], ct[2]);
        use |= 2;
      }
    }
    terms[wv][lane][0] = lt[0]; terms[wv][lane][1] = lt[1]; terms[wv][lane][2] = lt[2];
    terms[wv][lane][3] = ct[0]; terms[wv][lane][4] = ct[1]; terms[wv][lane][5] = ct[2];
    tuse[wv][lane] = use;
    wave_lds_sync();
    if (lane == 0) {
      const int n = min(64, nterm - l0);
      for (int k = 0; k < n; k++) {
        const int u = tuse[wv][k];
        const double *t = terms[wv][k];
        if (u & 1) { fx = fx + t[0]; fy = fy + t[1]; fz = fz + t[2]; }
        if (u & 2) { fx = fx + t[3]; fy = fy + t[4]; fz = fz + t[5]; }
      }
    }
    wave_lds_sync();
  }
  if (lane != 0) return;
  for (int wq = 0; wq < 2; wq++) {  // walls q = botw, topw (src/intgr_nwtn_eulr.f90:504-603)
    const double deltax = 0.0, deltay = 0.0, deltaz = (double)wq * g.lz - zp;
    const double dist = sqrt(sqd(deltax) + sqd(deltay) + sqd(deltaz));
    const double deltan = R - dist;
    const double nx = deltax / dist, ny = deltay / dist, nz = deltaz / dist;
    const double eps = -deltan / R;
    if ((eps < 0.075) && (eps > 0.0)) lubrication(pc, up, vp, wp, true, nx, ny, nz, eps, 0.0, 0.0, 0.0, fx, fy, fz);
    if (dist < R) collision(pc, up, vp, wp, omx, omy, omz, true, dist, deltax, deltay, deltaz, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, fx, fy, fz);
  }
  if (!last) {
    // the first pass saves the rest of the state at the start of
    // intgr_nwtn_eulr (src/intgr_nwtn_eulr.f90:77-118; the integrals, slots
    // 12..17, were saved by the sphere-integral launch) before updating it
    for (int s = 0; s < NOP; s++)
      if (s < 12 || s > 17) op[(long)s * np + p] = P.q[c_op_src[s]][p];
  }
  ne_update_one(pc, P, op, fx, fy, fz, *dtp, rkiter, p, dst);
  if (last) wrap_flags_one(g, pc, P, nbf, 1, p);
}

struct CollideArgs {
  Grid g; PartConst pc; PQ P; Motion src, dst;
  double *op; const double *dtp;
  int rkiter, last;
  int *nbf, *err;
};
struct CollideK {
  using Args = CollideArgs;
  static constexpr int THREADS = COL_WAVES * 64, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    collide_body(a.g, a.pc, a.P, a.src, a.dst, a.op, a.dtp, a.rkiter, a.last, a.nbf, a.err, v);
  }
};
__global__ __launch_bounds__(COL_WAVES * 64) void k_collide(CollideArgs) { run_single<CollideK>(); }

__global__ void k_wrap_flags(Grid g, PartConst pc, PQ P, int *nbf, int wrap) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= pc.np) return;
  wrap_flags_one(g, pc, P, nbf, wrap, p);
}

PQ make_pq(ipg_ctx *c) {
  PQ P;
  for (int q = 0; q < IPG_NPART; q++) P.q[q] = c->ps.q[q];
  return P;
}

// Phase indicator of the pressure cells, src/phase_indicator.f90:11-559 with
// itype = 1 as post3d calls it (src/output.f90:1068-1074): gamp(i,j,k) is the
// largest, over the spheres and the periodic images their neighbour flags
// name, of 1 when the cell's upper corner lies within radius - 2 dx of the
// centre (lines 437-438), else sum1/sum2 of the eight corners' signed
// distances (lines 440-457).  post3d passes one array for gamu, gamv, gamw,
// upart, vpart and wpart, which nothing reads, so only gamp is formed.  One
// block per sphere; the per-cell maximum is an atomic max on the bit
// patterns (the values are >= 0, whose bit patterns order as they do), so
// the result is independent of the order.  A cell whose nearest point lies
// beyond the radius has every signed distance negative and a ratio of
// exactly 0, which cannot raise the maximum: it is skipped.
__global__ __launch_bounds__(256) void k_phase_gamp(Grid g, PartConst pc, PQ P, const int *__restrict__ nbf,
                                                    unsigned long long *gamp) {
  const int p = blockIdx.x;
  const double R = pc.radius;
  const double dx = 1.0 / g.dxi, dy = 1.0 / g.dyi, dz = 1.0 / g.dzi;
  const double radin = R - 2.0 * dx, radin2 = radin * radin;
  for (int nb = 0; nb < IPG_NB; nb++) {
    if (nb > 0 && nbf[p * IPG_NB + nb] != 1) continue;
    // lines 248-258 (images set up at 83-119): on one rank every neighbour
    // is the rank itself, so the image of flag nb sits in slot nb + 4 (mod 8)
    int nbrecv = nb;
    if (nb > 0) { nbrecv = nb + 4; if (nbrecv > 8) nbrecv -= 8; }
    double X = P.q[IPG_PX][p], Y = P.q[IPG_PY][p];
    const double Z = P.q[IPG_PZ][p];
    image_shift(g, nbrecv, X, Y);
    const double cxc = X - 0.0, cyc = Y - 0.0, czc = Z;
    int ilow = nint_f((cxc - R) * g.dxi - 2.0), ihigh = nint_f((cxc + R) * g.dxi + 2.0);
    int jlow = nint_f((cyc - R) * g.dyi - 2.0), jhigh = nint_f((cyc + R) * g.dyi + 2.0);
    int klow = nint_f((czc - R) * g.dzi - 2.0), khigh = nint_f((czc + R) * g.dzi + 2.0);
    ilow = max(ilow, 1); jlow = max(jlow, 1); klow = max(klow, 1);
    ihigh = min(ihigh, g.nx); jhigh = min(jhigh, g.ny); khigh = min(khigh, g.nz);
    const int ni = ihigh - ilow + 1, nj = jhigh - jlow + 1, nk = khigh - klow + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) continue;
    const int nij = ni * nj, tot = nij * nk;
    for (int t = threadIdx.x; t < tot; t += blockDim.x) {
      const int kk = t / nij, r = t - kk * nij, jj = r / ni, ii = r - jj * ni;
      const int i = ilow + ii, j = jlow + jj, k = klow + kk;
      const double zmin = (double)(k - 1) * dz, zplus = (double)k * dz;
      const double ymin = 0.0 + (double)(j - 1) * dy, yplus = 0.0 + (double)j * dy;
      const double xmin = 0.0 + (double)(i - 1) * dx, xplus = 0.0 + (double)i * dx;
      const double ex = xmin > X ? xmin - X : (xplus < X ? X - xplus : 0.0);
      const double ey = ymin > Y ? ymin - Y : (yplus < Y ? Y - yplus : 0.0);
      const double ez = zmin > Z ? zmin - Z : (zplus < Z ? Z - zplus : 0.0);
      if (!(ex * ex + ey * ey + ez * ez <= R * R * (1.0 + 1e-9))) continue;
      const double dzp2 = sqd(zplus - Z), dyp2 = sqd(yplus - Y), dxp2 = sqd(xplus - X);
      const double dist2 = dxp2 + dyp2 + dzp2;
      double v;
      if (dist2 < radin2) {
        v = 1.0;
      } else {
        const double dxm2 = sqd(xmin - X), dym2 = sqd(ymin - Y), dzm2 = sqd(zmin - Z);
        double sg[8];
        sg[0] = R - sqrt(dxm2 + dym2 + dzm2); sg[1] = R - sqrt(dxm2 + dyp2 + dzm2);
        sg[2] = R - sqrt(dxp2 + dyp2 + dzm2); sg[3] = R - sqrt(dxp2 + dym2 + dzm2);
        sg[4] = R - sqrt(dxm2 + dym2 + dzp2); sg[5] = R - sqrt(dxm2 + dyp2 + dzp2);
        sg[6] = R - sqrt(dxp2 + dyp2 + dzp2); sg[7] = R - sqrt(dxp2 + dym2 + dzp2);
        double s1 = 0.0, s2 = 0.0;
#pragma unroll
        for (int cp = 0; cp < 8; cp++) { if (sg[cp] > 0.0) s1 = s1 + sg[cp]; s2 = s2 + fabs(sg[cp]); }
        v = s1 / s2;
      }
      if (v > 0.0)
        atomicMax(&gamp[(long)(i - 1) + (long)g.nx * ((long)(j - 1) + (long)g.ny * (k - 1))],
                  (unsigned long long)__double_as_longlong(v));
    }
  }
}

// sol = 1 - gamp (src/phase_indicator.f90:497), in place over the interior
__global__ void k_phase_sol(double *a, long n) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) a[t] = 1.0 - a[t];
}

}  // namespace

// The 'sol' quantity of post3d into out (interior, i fastest, nx*ny*nz).
int ipg_launch_phase_sol(ipg_ctx *c, double *out) {
  const Grid &g = c->g;
  const long n = (long)g.nx * g.ny * g.nz;
  IPG_HIP(hipMemsetAsync(out, 0, n * sizeof(double), c->stream));
  if (c->pc.np > 0)
    hipLaunchKernelGGL(k_phase_gamp, dim3(c->pc.np), dim3(256), 0, c->stream, g, c->pc, make_pq(c), c->ps.nbf,
                       (unsigned long long *)out);
  hipLaunchKernelGGL(k_phase_sol, dim3((unsigned)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, c->stream, out, n);
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_sphere_integrals(ipg_ctx *c, int type) {
  if (c->pc.np <= 0) return 0;
  const SphereIntArgs a{c->g, c->pc, c->f[IPG_U], c->f[IPG_V], c->f[IPG_W], make_pq(c), c->ps.nbf, nullptr};
  if (launch_sphere_int(c, a, type, "sphere_int")) return 1;
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_nbflags(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  hipLaunchKernelGGL(k_wrap_flags, dim3((c->pc.np + 255) / 256), dim3(256), 0, c->stream, c->g, c->pc, make_pq(c), c->ps.nbf, 0);
  IPG_HIP(hipGetLastError());
  return 0;
}

// Blocks of the sphere-integral launch.  Beside the RK stage (the update on
// the side stream) one block per CU striding over the spheres, once there are
// more spheres than CUs: each CU then keeps most of its wavefront slots for
// the RK stage, whose HBM stream a full grid of sphere blocks would thin (256^3
// with 1 000 spheres: 706 against 687 substeps/s; 512^3 with 8 000: 74.2
// against 69.7).  Otherwise, and in a batch's merged launches, one per sphere.
int ipg_sphere_blocks(const ipg_ctx *c) {
  const int np = c->pc.np;
  if (c->rec || c->sph_grid == 0) return np;
  if (c->sph_grid > 0) return std::min(c->sph_grid, np);
  return c->side_on && c->ncu > 0 && np > c->ncu ? c->ncu : np;
}

// intgr_nwtn_eulr (src/intgr_nwtn_eulr.f90:12-714) in three launches: the
// sphere integrals of types 2 and 3 (120-121) with the saved state op
// (77-118); the first collision pass with its position update (338-672),
// into the second motion buffer; the second pass back into the particle
// state, with the periodic wrap and the neighbour flags (702-714).
int ipg_launch_nwtn_eulr(ipg_ctx *c, int rkiter) {
  const int np = c->pc.np;
  if (np <= 0) return 0;
  const PQ P = make_pq(c);
  const SphereIntArgs sa{c->g, c->pc, c->f[IPG_U], c->f[IPG_V], c->f[IPG_W], P, c->ps.nbf, c->op};
  if (launch_sphere_int(c, sa, 23, "sphere_int")) return 1;
  double **q = c->ps.q;
  const Motion state{q[IPG_PX], q[IPG_PY], q[IPG_PZ], q[IPG_PU], q[IPG_PV], q[IPG_PW], q[IPG_POMX], q[IPG_POMY], q[IPG_POMZ]};
  double *m = c->motion2;
  const Motion second{m, m + np, m + 2L * np, m + 3L * np, m + 4L * np, m + 5L * np, m + 6L * np, m + 7L * np, m + 8L * np};
  int *err = c->ibm_err;  // device error word, bit 8: a candidate list overflowed
  const dim3 grid((np + COL_WAVES - 1) / COL_WAVES), blk(COL_WAVES * 64);
  if (ipg_go<CollideK>(c, k_collide, grid, blk, 0, CollideArgs{c->g, c->pc, P, state, second, c->op, c->dscal, rkiter, 0,
                                                             c->ps.nbf, err}, "collide_update"))
    return 1;
  if (ipg_go<CollideK>(c, k_collide, grid, blk, 0, CollideArgs{c->g, c->pc, P, second, state, c->op, c->dscal, rkiter, 1,
                                                             c->ps.nbf, err}, "collide_update"))
    return 1;
  IPG_HIP(hipGetLastError());
  return 0;
}
