// ipg_fields.hip -- field kernels of the RK3 substep: fused RK stage
// (momentum + scalar + pressure gradient + RK combination), boundary/ghost
// fill, Poisson right-hand side, projection, divergence and time-step checks.
// FP64, with the reference's operation order kept term by term (the library
// is compiled with -ffp-contract=off so no multiply-add is fused).
#include "ipg_internal.h"
#include <algorithm>
#include <mutex>
#include <vector>

namespace {

constexpr int BX = 64, BY = 4;

// One RK stage, src/rk3.f90:13-106 (STAGE 1), 108-222 (2), 224-329 (3), with
// momxad/momyad/momzad/momxp/momyp/momzp/momzbuoy (src/mom.f90) and scalad
// (src/scal.f90:11-60) evaluated in registers: one read of each input and one
// write of each output per cell.
//
// A workgroup owns a 64 x 8 (i, j) column tile and marches up KB planes in z.
// The five stencil inputs (u v w c p) of planes k and k+1, each with a
// one-cell halo in i and j, sit in a two-slot LDS ring; every plane is read
// from HBM once per workgroup (coalesced 64-wide rows), the next plane is
// prefetched into registers while the current one is computed, and the k-1
// values a cell needs stay in registers.  Tiles are numbered i fastest, then
// j, then z, and handed out so that each XCD works through one contiguous
// range of them: neighbouring tiles, which share halo rows, then run on the
// same XCD and meet in its L2.
constexpr int RTX = 64, RTY = 8, RNT = RTX * RTY;
constexpr int LW = RTX + 2, LH = RTY + 2, LPL = LW * LH;
constexpr int NLD = (LPL + RNT - 1) / RNT;

//
// RT is the storage type of the RK history durk..dcrk: double, or float in
// the mixed path, where the history is rounded to FP32 when stored and every
// term is still computed in FP64 (the kernel is bound by HBM traffic, not by
// its FP64 arithmetic).
template <int STAGE, typename RT>
__device__ __forceinline__ void rk_body(const Grid g, const double *__restrict__ U, const double *__restrict__ V,
                                        const double *__restrict__ W, const double *__restrict__ C,
                                        const double *__restrict__ P, double *__restrict__ dudt,
                                        double *__restrict__ dvdt, double *__restrict__ dwdt,
                                        double *__restrict__ dcdt, RT *__restrict__ durk,
                                        RT *__restrict__ dvrk, RT *__restrict__ dwrk,
                                        RT *__restrict__ dcrk, const double *__restrict__ dtp, int ntx,
                                        int nty, int kb, int wrapx, const VB vb) {
  __shared__ double sm[2][5][LPL];
  const int nblk = vb.gx, L = vb.x;
  const int t = (nblk % 8 == 0) ? (L % 8) * (nblk / 8) + L / 8 : L;
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int i0 = tx * RTX + 1, j0 = ty * RTY + 1;
  const int k0 = tz * kb + 1, k1 = min(k0 + kb - 1, g.nz);
  const int lx = (int)threadIdx.x % RTX, ly = (int)threadIdx.x / RTX;
  const int i = i0 + lx, j = j0 + ly;
  const bool active = i <= g.nx && j <= g.ny;
  const long sx = g.sx, sxy = g.sxy;
  const double *src[5] = {U, V, W, C, P};
  // this thread's share of a halo plane: element e = jj*LW + ii <-> cell (i0-1+ii, j0-1+jj)
  long goff[NLD];
  bool gok[NLD];
#pragma unroll
  for (int q = 0; q < NLD; q++) {
    const int e = (int)threadIdx.x + q * RNT, ii = e % LW, jj = e / LW;
    const int gi = i0 - 1 + ii, gj = j0 - 1 + jj;
    gok[q] = e < LPL && gi <= g.nx + 1 && gj <= g.ny + 1;
    // wrapx: the x halo is read at its periodic image, so the x-ghost
    // columns of the inputs need not be filled (k_bound's plane and row parts
    // fill the ghost rows and planes, x ghosts included); else at the ghosts
    const int gw = !wrapx ? gi : (gi == 0 ? g.nx : (gi == g.nx + 1 ? 1 : gi));
    goff[q] = (long)(gw - 1) + sx * gj;
  }
  double pre[5][NLD];
  auto fetch = [&](int k) {
#pragma unroll
    for (int f = 0; f < 5; f++)
#pragma unroll
      for (int q = 0; q < NLD; q++) pre[f][q] = gok[q] ? src[f][goff[q] + sxy * k] : 0.0;
  };
  auto stash = [&](int slot) {
#pragma unroll
    for (int f = 0; f < 5; f++)
#pragma unroll
      for (int q = 0; q < NLD; q++) {
        const int e = (int)threadIdx.x + q * RNT;
        if (e < LPL) sm[slot][f][e] = pre[f][q];
      }
  };
  const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, visc = g.visc, kappa = g.kappa;
  const double dt = *dtp;
  double f1, f2 = 0.0, f3 = 0.0;
  if (STAGE == 1) f1 = dt * (32.0 / 60.0);
  else if (STAGE == 2) { f1 = dt * (8.0 / 60.0); f2 = dt * (25.0 / 60.0); f3 = dt * (17.0 / 60.0); }
  else { f1 = dt * (20.0 / 60.0); f2 = dt * (45.0 / 60.0); f3 = dt * (25.0 / 60.0); }
  const double bq = 0.5 * g.beta * 9.81;
  long c = g.at(i, j, k0);
  // k-1 values of the first plane
  double Ukm = 0.0, Vkm = 0.0, Wkm = 0.0, Wip_km = 0.0, Wjp_km = 0.0, Ckm = 0.0;
  if (active) {
    Ukm = U[c - sxy]; Vkm = V[c - sxy]; Wkm = W[c - sxy];
    Wip_km = W[(wrapx && i == g.nx ? c - (g.nx - 1) : c + 1) - sxy];
    Wjp_km = W[c + sx - sxy]; Ckm = C[c - sxy];
  }
  fetch(k0);
  stash(0);
  fetch(k0 + 1);
  stash(1);
  __syncthreads();
  const int e0 = (ly + 1) * LW + lx + 1;  // this cell in a halo plane
  int a = 0;
  for (int k = k0; k <= k1; k++, c += sxy) {
    const bool more = k + 2 <= k1 + 1;
    if (more) fetch(k + 2);
    const double(*pk)[LPL] = sm[a];
    const double(*pn)[LPL] = sm[a ^ 1];
    if (active) {
    const double Uc = pk[0][e0], Vc = pk[1][e0], Wc = pk[2][e0], Cc = pk[3][e0], Pc = pk[4][e0];
    const double Ukp = pn[0][e0], Uim_kp = pn[0][e0 - 1], Vkp = pn[1][e0], Vjm_kp = pn[1][e0 - LW];
    const double Wkp = pn[2][e0], Ckp = pn[3][e0], Pkp = pn[4][e0];
    const double Uim = pk[0][e0 - 1], Uip = pk[0][e0 + 1], Ujp = pk[0][e0 + LW], Ujm = pk[0][e0 - LW];
    const double Uimjp = pk[0][e0 - 1 + LW];
    const double Vip = pk[1][e0 + 1], Vim = pk[1][e0 - 1], Vjp = pk[1][e0 + LW], Vjm = pk[1][e0 - LW];
    const double Vipjm = pk[1][e0 + 1 - LW];
    const double Wip = pk[2][e0 + 1], Wim = pk[2][e0 - 1], Wjp = pk[2][e0 + LW], Wjm = pk[2][e0 - LW];
    const double Cip = pk[3][e0 + 1], Cim = pk[3][e0 - 1], Cjp = pk[3][e0 + LW], Cjm = pk[3][e0 - LW];
    const double Pip = pk[4][e0 + 1], Pjp = pk[4][e0 + LW];
    // momxp/momyp/momzp (src/mom.f90:56-79,129-155,205-231)
    const double gpx = -(dxi * (Pip - Pc));
    const double gpy = -(dyi * (Pjp - Pc));
    const double gpz = -(dzi * (Pkp - Pc));
    const double du0 = Uc + f1 * gpx, dv0 = Vc + f1 * gpy, dw0 = Wc + f1 * gpz;
    double ax, ay, az, ac;
    {  // momxad, src/mom.f90:30-47
      double uuip = 0.25 * (Uip + Uc) * (Uip + Uc);
      double uuim = 0.25 * (Uim + Uc) * (Uim + Uc);
      double uvjp = 0.25 * (Ujp + Uc) * (Vip + Vc);
      double uvjm = 0.25 * (Ujm + Uc) * (Vipjm + Vjm);
      double uwkp = 0.25 * (Ukp + Uc) * (Wip + Wc);
      double uwkm = 0.25 * (Ukm + Uc) * (Wip_km + Wkm);
      double dudxp = (Uip - Uc) * dxi, dudxm = (Uc - Uim) * dxi;
      double dudyp = (Ujp - Uc) * dyi, dudym = (Uc - Ujm) * dyi;
      double dudzp = (Ukp - Uc) * dzi, dudzm = (Uc - Ukm) * dzi;
      ax = dxi * (-uuip + uuim) + visc * (dudxp - dudxm) * dxi + dyi * (-uvjp + uvjm) + visc * (dudyp - dudym) * dyi +
           dzi * (-uwkp + uwkm) + visc * (dudzp - dudzm) * dzi;
    }
    {  // momyad, src/mom.f90:103-120
      double uvip = 0.25 * (Uc + Ujp) * (Vc + Vip);
      double uvim = 0.25 * (Uim + Uimjp) * (Vc + Vim);
      double vvjp = 0.25 * (Vc + Vjp) * (Vc + Vjp);
      double vvjm = 0.25 * (Vc + Vjm) * (Vc + Vjm);
      double wvkp = 0.25 * (Wc + Wjp) * (Vkp + Vc);
      double wvkm = 0.25 * (Wkm + Wjp_km) * (Vkm + Vc);
      double dvdxp = (Vip - Vc) * dxi, dvdxm = (Vc - Vim) * dxi;
      double dvdyp = (Vjp - Vc) * dyi, dvdym = (Vc - Vjm) * dyi;
      double dvdzp = (Vkp - Vc) * dzi, dvdzm = (Vc - Vkm) * dzi;
      ay = dxi * (-uvip + uvim) + visc * (dvdxp - dvdxm) * dxi + dyi * (-vvjp + vvjm) + visc * (dvdyp - dvdym) * dyi +
           dzi * (-wvkp + wvkm) + visc * (dvdzp - dvdzm) * dzi;
    }
    {  // momzad, src/mom.f90:179-196; momzbuoy src/mom.f90:252 added first (src/rk3.f90:56-58)
      double uwip = 0.25 * (Wc + Wip) * (Uc + Ukp);
      double uwim = 0.25 * (Wc + Wim) * (Uim + Uim_kp);
      double vwjp = 0.25 * (Wc + Wjp) * (Vc + Vkp);
      double vwjm = 0.25 * (Wc + Wjm) * (Vjm + Vjm_kp);
      double wwkp = 0.25 * (Wc + Wkp) * (Wc + Wkp);
      double wwkm = 0.25 * (Wc + Wkm) * (Wc + Wkm);
      double dwdxp = (Wip - Wc) * dxi, dwdxm = (Wc - Wim) * dxi;
      double dwdyp = (Wjp - Wc) * dyi, dwdym = (Wc - Wjm) * dyi;
      double dwdzp = (Wkp - Wc) * dzi, dwdzm = (Wc - Wkm) * dzi;
      double adv = dxi * (-uwip + uwim) + visc * (dwdxp - dwdxm) * dxi + dyi * (-vwjp + vwjm) +
                   visc * (dwdyp - dwdym) * dyi + dzi * (-wwkp + wwkm) + visc * (dwdzp - dwdzm) * dzi;
      double buoy = bq * (Cc + Ckp);
      az = buoy + adv;
    }
    {  // scalad, src/scal.f90:35-53
      double ucip = 0.5 * Uc * (Cip + Cc);
      double ucim = 0.5 * Uim * (Cc + Cim);
      double vcjp = 0.5 * Vc * (Cjp + Cc);
      double vcjm = 0.5 * Vjm * (Cc + Cjm);
      double wckp = 0.5 * Wc * (Ckp + Cc);
      double wckm = 0.5 * Wkm * (Cc + Ckm);
      double dcdxp = (Cip - Cc) * dxi, dcdxm = (Cc - Cim) * dxi;
      double dcdyp = (Cjp - Cc) * dyi, dcdym = (Cc - Cjm) * dyi;
      double dcdzp = (Ckp - Cc) * dzi, dcdzm = (Cc - Ckm) * dzi;
      ac = dxi * (-ucip + ucim) + kappa * (dcdxp - dcdxm) * dxi + dyi * (-vcjp + vcjm) + kappa * (dcdyp - dcdym) * dyi +
           dzi * (-wckp + wckm) + kappa * (dcdzp - dcdzm) * dzi;
    }
    if (STAGE == 1) {
      dudt[c] = du0 + f1 * ax;
      dvdt[c] = dv0 + f1 * ay;
      dwdt[c] = dw0 + f1 * az;
      dcdt[c] = Cc + f1 * ac;
      durk[c] = (RT)ax; dvrk[c] = (RT)ay; dwrk[c] = (RT)az; dcrk[c] = (RT)ac;
    } else {
      const double ru = durk[c], rv = dvrk[c], rw = dwrk[c], rc = dcrk[c];
      dudt[c] = du0 + (f2 * ax - f3 * ru);
      dvdt[c] = dv0 + (f2 * ay - f3 * rv);
      dwdt[c] = dw0 + (f2 * az - f3 * rw);
      dcdt[c] = Cc + (f2 * ac - f3 * rc);
      if (STAGE == 2) { durk[c] = (RT)ax; dvrk[c] = (RT)ay; dwrk[c] = (RT)az; dcrk[c] = (RT)ac; }
    }
    // the k-1 values of the next plane
    Ukm = Uc; Vkm = Vc; Wkm = Wc; Wip_km = Wip; Wjp_km = Wjp; Ckm = Cc;
    }
    __syncthreads();          // every wave is done with plane k
    if (more) stash(a);       // plane k+2 takes its slot
    __syncthreads();
    a ^= 1;
  }
}

template <typename RT>
struct RkArgs {
  Grid g;
  const double *U, *V, *W, *C, *P;
  double *dudt, *dvdt, *dwdt, *dcdt;
  RT *durk, *dvrk, *dwrk, *dcrk;
  const double *dtp;
  int ntx, nty, kb, wrapx;
};
template <int STAGE, typename RT>
struct RkK {
  using Args = RkArgs<RT>;
  static constexpr int THREADS = RNT, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    rk_body<STAGE, RT>(a.g, a.U, a.V, a.W, a.C, a.P, a.dudt, a.dvdt, a.dwdt, a.dcdt, a.durk, a.dvrk, a.dwrk, a.dcrk,
                       a.dtp, a.ntx, a.nty, a.kb, a.wrapx, v);
  }
};
// (the single-context kernel keeps its arguments as separate parameters: as a
// struct they measured 2.5 % slower in the RK stage at 256^3)
template <int STAGE, typename RT>
__global__ __launch_bounds__(RNT) void k_rk(const Grid g, const double *__restrict__ U, const double *__restrict__ V,
                                            const double *__restrict__ W, const double *__restrict__ C,
                                            const double *__restrict__ P, double *__restrict__ dudt,
                                            double *__restrict__ dvdt, double *__restrict__ dwdt,
                                            double *__restrict__ dcdt, RT *__restrict__ durk, RT *__restrict__ dvrk,
                                            RT *__restrict__ dwrk, RT *__restrict__ dcrk,
                                            const double *__restrict__ dtp, int ntx, int nty, int kb, int wrapx) {
  rk_body<STAGE, RT>(g, U, V, W, C, P, dudt, dvdt, dwdt, dcdt, durk, dvrk, dwrk, dcrk, dtp, ntx, nty, kb, wrapx, vb_hw());
}
template <int STAGE, typename RT>
void launch_rk(const RkArgs<RT> &a, dim3 grid, hipStream_t s) {
  hipLaunchKernelGGL((k_rk<STAGE, RT>), grid, dim3(RNT), 0, s, a.g, a.U, a.V, a.W, a.C, a.P, a.dudt, a.dvdt, a.dwdt,
                     a.dcdt, a.durk, a.dvrk, a.dwrk, a.dcrk, a.dtp, a.ntx, a.nty, a.kb, a.wrapx);
}

// Ghost fill for up to five fields in one launch: src/bound.f90:11-80
// (bounduvw), 82-107 (boundp), 109-171 (boundscal) followed by updthalos
// (173-221) on one rank.  Each ghost takes its final value directly: the z
// rule applied to the periodic image (i',j') of its column, which is what the
// reference's sequence z-rule -> x-halo -> y-halo produces.
enum { KIND_UV = 0, KIND_W = 1, KIND_P = 2, KIND_C = 3 };
// xcol = 0 leaves out the x-ghost columns (i = 0, nx+1 for j in 1..ny, k in
// 1..nz), one scattered value per row: the fused substep's readers take the
// periodic image instead, and the library fills them before anything else
// reads them.
struct BoundJob { double *f[5]; int kind[5]; int nf; int xcol; };

constexpr int BND_THREADS = 256;
__device__ __forceinline__ void bound_body(const Grid &g, const BoundJob &job, const VB vb) {
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int nplane = (nx + 2) * (ny + 2), nrow = (nx + 2) * nz, ncol = job.xcol ? ny * nz : 0, nwall = nx * ny;
  const int total = 2 * nplane + 2 * nrow + 2 * ncol + nwall;
  for (int t = vb.x * BND_THREADS + threadIdx.x; t < total; t += vb.gx * BND_THREADS) {
    int i, j, k; bool wallplane = false;
    int r = t;
    if (r < 2 * nplane) { k = r < nplane ? 0 : nz + 1; r = r < nplane ? r : r - nplane; i = r % (nx + 2); j = r / (nx + 2); }
    else if ((r -= 2 * nplane) < 2 * nrow) {
      j = r < nrow ? 0 : ny + 1; r = r < nrow ? r : r - nrow; i = r % (nx + 2); k = r / (nx + 2) + 1;
    } else if ((r -= 2 * nrow) < 2 * ncol) {
      i = r < ncol ? 0 : nx + 1; r = r < ncol ? r : r - ncol; j = r % ny + 1; k = r / ny + 1;
    } else {
      r -= 2 * ncol; i = r % nx + 1; j = r / nx + 1; k = nz; wallplane = true;
    }
    const int ii = i == 0 ? (int)nx : (i == nx + 1 ? 1 : i);
    const int jj = j == 0 ? (int)ny : (j == ny + 1 ? 1 : j);
    const long dst = g.at(i, j, k);
    for (int q = 0; q < job.nf; q++) {
      double *f = job.f[q];
      const int kind = job.kind[q];
      if (wallplane) { if (kind == KIND_W) f[dst] = 0.0; continue; }
      double val;
      if (kind == KIND_W) {
        if (k == 0 || k == nz) val = 0.0;
        else val = f[g.at(ii, jj, k)];
      } else if (k == nz + 1) {
        const double in = f[g.at(ii, jj, (int)nz)];
        if (kind == KIND_UV) val = g.bc_uvw_top == 0 ? -in : in;
        else if (kind == KIND_P) val = in;
        else val = g.bc_c_top == 0 ? 2.0 * g.bc_c_top_val - in : in;
      } else if (k == 0) {
        const double in = f[g.at(ii, jj, 1)];
        if (kind == KIND_UV) val = g.bc_uvw_bot == 0 ? -in : in;
        else if (kind == KIND_P) val = in;
        else val = g.bc_c_bot == 0 ? 2.0 * g.bc_c_bot_val - in : in;
      } else {
        val = f[g.at(ii, jj, k)];
      }
      f[dst] = val;
    }
  }
}

struct BoundArgs { Grid g; BoundJob job; };
struct BoundK {
  using Args = BoundArgs;
  static constexpr int THREADS = BND_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) { bound_body(a.g, a.job, v); }
};
__global__ __launch_bounds__(BND_THREADS) void k_bound(BoundArgs) { run_single<BoundK>(); }

// src/fillps.f90:8-53
__global__ __launch_bounds__(BX *BY) void k_fillps(Grid g, const double *__restrict__ dudt, const double *__restrict__ dvdt,
                                                   const double *__restrict__ dwdt, double *__restrict__ p,
                                                   const double *__restrict__ dtp, double rkab) {
  const int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const double dti = 1.0 / (rkab * *dtp);
  const double dtidxi = dti * g.dxi, dtidyi = dti * g.dyi, dtidzi = dti * g.dzi;
  const long c = g.at(i, j, k);
  p[c] = (dwdt[c] - dwdt[c - g.sxy]) * dtidzi + (dvdt[c] - dvdt[c - g.sx]) * dtidyi + (dudt[c] - dudt[c - 1]) * dtidxi;
}

// src/correc.f90:11-42.  p is read at i+1, j+1, k+1 through the values
// boundp (src/bound.f90:82-107) gives its ghosts -- the periodic image in x
// and y, p(kmax) above the top wall -- so p's ghosts need not be filled.
// FUSED adds src/main.f90:370 (pnew += p) on the interior; cnew = dcdt (375)
// is a swap of the two buffers' roles made by the caller, and the ghosts of
// u, v, w, pnew, cnew are filled by k_bound after.
// (Writing those ghosts from the threads of their interior images was
// measured: the scattered x-ghost writes and the divergent boundary lanes
// cost as much as the separate ghost-fill launch.)
template <bool FUSED>
__device__ __forceinline__ void correc_body(const Grid g, const double *__restrict__ dudt, const double *__restrict__ dvdt,
                                            const double *__restrict__ dwdt, const double *__restrict__ p,
                                            double *__restrict__ u, double *__restrict__ v, double *__restrict__ w,
                                            double *__restrict__ pnew, const double *__restrict__ dtp, double rkab,
                                            const VB vb) {
  const int i = vb.x * BX + threadIdx.x + 1, j = vb.y * BY + threadIdx.y + 1, k = vb.z + 1;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
  const double factor = rkab * *dtp;
  const double fi = factor * g.dxi, fj = factor * g.dyi, fk = factor * g.dzi;
  const long c = g.at(i, j, k);
  const double pc = p[c];
  const double pi1 = p[i == nx ? c - (nx - 1) : c + 1];
  const double pj1 = p[j == ny ? c - (long)(ny - 1) * g.sx : c + g.sx];
  const double pk1 = k == nz ? pc : p[c + g.sxy];
  u[c] = dudt[c] - fi * (pi1 - pc);
  v[c] = dvdt[c] - fj * (pj1 - pc);
  w[c] = dwdt[c] - fk * (pk1 - pc);
  if (FUSED) pnew[c] = pnew[c] + pc;
}

struct CorrecArgs {
  Grid g;
  const double *dudt, *dvdt, *dwdt, *p;
  double *u, *v, *w, *pnew;
  const double *dtp;
  double rkab;
};
template <bool FUSED>
struct CorrecK {
  using Args = CorrecArgs;
  static constexpr int THREADS = BX * BY, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    correc_body<FUSED>(a.g, a.dudt, a.dvdt, a.dwdt, a.p, a.u, a.v, a.w, a.pnew, a.dtp, a.rkab, v);
  }
};
template <bool FUSED>
__global__ __launch_bounds__(BX *BY) void k_correc(CorrecArgs) { run_single<CorrecK<FUSED>>(); }

// The mixed path's pressure correction is the sum of NP FP32 pieces (the
// FP32-stored solve and its FP64-residual refinements), each exact in FP64:
// p = (p0 + p1) + p2 in FP64 wherever p is read.
struct Pieces { const float *q[IPG_MAX_REFINE + 1]; };
template <int NP>
__device__ __forceinline__ double psum(const Pieces &P, long c) {
  double s = (double)P.q[0][c];
#pragma unroll
  for (int r = 1; r < NP; r++) s += (double)P.q[r][c];
  return s;
}

// k_correc<true> with p read as the sum of its pieces (src/correc.f90:11-42,
// src/main.f90:370), the same periodic and top-wall source values.
template <int NP>
__device__ __forceinline__ void correc_pieces_body(const Grid g, const double *__restrict__ dudt,
                                                   const double *__restrict__ dvdt, const double *__restrict__ dwdt,
                                                   const Pieces P, double *__restrict__ u, double *__restrict__ v,
                                                   double *__restrict__ w, double *__restrict__ pnew,
                                                   const double *__restrict__ dtp, double rkab, const VB vb) {
  const int i = vb.x * BX + threadIdx.x + 1, j = vb.y * BY + threadIdx.y + 1, k = vb.z + 1;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
  const double factor = rkab * *dtp;
  const double fi = factor * g.dxi, fj = factor * g.dyi, fk = factor * g.dzi;
  const long c = g.at(i, j, k);
  const double pc = psum<NP>(P, c);
  const double pi1 = psum<NP>(P, i == nx ? c - (nx - 1) : c + 1);
  const double pj1 = psum<NP>(P, j == ny ? c - (long)(ny - 1) * g.sx : c + g.sx);
  const double pk1 = k == nz ? pc : psum<NP>(P, c + g.sxy);
  u[c] = dudt[c] - fi * (pi1 - pc);
  v[c] = dvdt[c] - fj * (pj1 - pc);
  w[c] = dwdt[c] - fk * (pk1 - pc);
  pnew[c] = pnew[c] + pc;
}

struct CorrecPiecesArgs {
  Grid g;
  const double *dudt, *dvdt, *dwdt;
  Pieces P;
  double *u, *v, *w, *pnew;
  const double *dtp;
  double rkab;
};
template <int NP>
struct CorrecPiecesK {
  using Args = CorrecPiecesArgs;
  static constexpr int THREADS = BX * BY, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    correc_pieces_body<NP>(a.g, a.dudt, a.dvdt, a.dwdt, a.P, a.u, a.v, a.w, a.pnew, a.dtp, a.rkab, v);
  }
};
template <int NP>
__global__ __launch_bounds__(BX *BY) void k_correc_pieces(CorrecPiecesArgs) { run_single<CorrecPiecesK<NP>>(); }

// The mixed path's refinement residual r = rhs - L p, p the sum of its NP
// pieces so far, in FP64, stored FP32.  L is the operator the Poisson solve
// inverts: the periodic 3-point second differences in x and y, and in z the
// Neumann rows of src/initsolver.f90:17-30 (p(0) = p(1), p(kmax+1) =
// p(kmax)); the pieces' ghosts are never filled, so the neighbours are taken
// at the source.
// Each thread marches RZ planes up its column, keeping p below, at and above
// the current plane in registers, so that every plane of the pieces is read
// from HBM once; the in-plane neighbours come from the caches.
constexpr int RZ = 16;
template <int NP>
__device__ __forceinline__ void resid_body(const Grid g, const Pieces P, const double *__restrict__ rhs,
                                           float *__restrict__ r, const VB vb) {
  const int i = vb.x * BX + threadIdx.x + 1, j = vb.y * BY + threadIdx.y + 1;
  const int k0 = vb.z * RZ + 1, k1 = min(k0 + RZ - 1, g.nz);
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (i > nx || j > ny) return;
  const long ce = i == nx ? -(long)(nx - 1) : 1, cw = i == 1 ? (long)(nx - 1) : -1;
  const long cn = j == ny ? -(long)(ny - 1) * g.sx : g.sx, cs = j == 1 ? (long)(ny - 1) * g.sx : -g.sx;
  long c = g.at(i, j, k0);
  double pc = psum<NP>(P, c);
  double pd = k0 == 1 ? pc : psum<NP>(P, c - g.sxy);
  for (int k = k0; k <= k1; k++, c += g.sxy) {
    const double pu = k == nz ? pc : psum<NP>(P, c + g.sxy);
    const double pe = psum<NP>(P, c + ce), pw = psum<NP>(P, c + cw);
    const double pn = psum<NP>(P, c + cn), ps = psum<NP>(P, c + cs);
    const double lap = ((pe - pc) - (pc - pw)) * (g.dxi * g.dxi) + ((pn - pc) - (pc - ps)) * (g.dyi * g.dyi) +
                       ((pu - pc) - (pc - pd)) * (g.dzi * g.dzi);
    r[c] = (float)(rhs[c] - lap);
    pd = pc;
    pc = pu;
  }
}

struct ResidArgs { Grid g; Pieces P; const double *rhs; float *r; };
template <int NP>
struct ResidK {
  using Args = ResidArgs;
  static constexpr int THREADS = BX * BY, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) { resid_body<NP>(a.g, a.P, a.rhs, a.r, v); }
};
template <int NP>
__global__ __launch_bounds__(BX *BY) void k_resid(ResidArgs) { run_single<ResidK<NP>>(); }

// p = the sum of its pieces on the interior (before p is read as a field)
template <int NP>
__global__ __launch_bounds__(BX *BY) void k_p_combine(Grid g, Pieces P, double *__restrict__ p) {
  const int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long c = g.at(i, j, k);
  p[c] = psum<NP>(P, c);
}

__global__ __launch_bounds__(BX *BY) void k_pnew(Grid g, const double *__restrict__ p, double *__restrict__ pnew,
                                                 const double *__restrict__ dcdt, double *__restrict__ cnew) {
  const int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long c = g.at(i, j, k);
  pnew[c] = pnew[c] + p[c];
  cnew[c] = dcdt[c];
}

// ---- reductions -------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// src/chkdt.f90:59-86 (u), 121-148 (v), 182-209 (w): per cell the three
// components' min(dt1, dt2); the minimum over the domain is order-free.
__device__ __forceinline__ double chkdt_dt2(double u2, double v2, double w2, double dxi, double dyi, double dzi) {
  double velo = sqrt(u2 + v2 + w2);
  if (fabs(1.0 * velo) > 1.e-12) return sqrt(3.0) / (sqrt(u2) * dxi + sqrt(v2) * dyi + sqrt(w2) * dzi);
  return 99999.0;
}

// src/chkdt.f90:6-271 as a march in z: a workgroup takes a 64 x 4 (i, j)
// tile through kb planes, one column per lane.  A cell's u and v values at
// k+1 and its w values at k-1 are carried in registers between planes, and
// its i and j neighbours are in the cache lines the tile's other lanes load,
// so each field crosses HBM about once (a grid-stride form re-read the
// planes either side: 3.7 times the bytes).  At each cell the three dt
// limits of the u, v and w points, each from the velocity interpolated to
// that point, in the reference's order of terms.
__global__ __launch_bounds__(256) void k_chkdt_march(Grid g, const double *__restrict__ U, const double *__restrict__ V,
                                                      const double *__restrict__ W, double *__restrict__ part, int kb) {
  __shared__ double sm[4];
  const int i = blockIdx.x * 64 + (int)(threadIdx.x & 63) + 1, j = blockIdx.y * 4 + (int)(threadIdx.x >> 6) + 1;
  const int k0 = blockIdx.z * kb + 1, k1 = min(k0 + kb - 1, g.nz);
  const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi;
  const double dt1 = 1.65 / (4.0 * g.visc * (dxi * dxi + dyi * dyi + dzi * dzi));
  double m = 899999.0;
  if (i <= g.nx && j <= g.ny && k0 <= k1) {
    const long sx = g.sx, sxy = g.sxy;
    long c = g.at(i, j, k0);
    // plane k: u(i,j), u(i-1,j), v(i,j), v(i,j-1); plane k-1: w(i,j), w(i+1,j), w(i,j+1)
    double Uc = U[c], Um = U[c - 1], Vc = V[c], Vm = V[c - sx];
    double Wc1 = W[c - sxy], Wp1 = W[c + 1 - sxy], Wj1 = W[c + sx - sxy];
    for (int k = k0; k <= k1; k++, c += sxy) {
      const double Ujp = U[c + sx], Umjp = U[c - 1 + sx], Vip = V[c + 1], Vipjm = V[c + 1 - sx];
      const double Wc = W[c], Wp = W[c + 1], Wj = W[c + sx];
      const double Uk1 = U[c + sxy], Umk1 = U[c - 1 + sxy], Vk1 = V[c + sxy], Vmk1 = V[c - sx + sxy];
      double u2 = sqd(Uc);
      double v2 = sqd(0.25 * (Vc + Vm + Vip + Vipjm));
      double w2 = sqd(0.25 * (Wc + Wc1 + Wp + Wp1));
      m = fmin(m, fmin(dt1, chkdt_dt2(u2, v2, w2, dxi, dyi, dzi)));
      u2 = sqd(0.25 * (Uc + Ujp + Umjp + Um));
      v2 = sqd(Vc);
      w2 = sqd(0.25 * (Wc + Wj + Wj1 + Wc1));
      m = fmin(m, fmin(dt1, chkdt_dt2(u2, v2, w2, dxi, dyi, dzi)));
      u2 = sqd(0.25 * (Uc + Um + Umk1 + Uk1));
      v2 = sqd(0.25 * (Vc + Vm + Vmk1 + Vk1));
      w2 = sqd(Wc);
      m = fmin(m, fmin(dt1, chkdt_dt2(u2, v2, w2, dxi, dyi, dzi)));
      Uc = Uk1; Um = Umk1; Vc = Vk1; Vm = Vmk1;
      Wc1 = Wc; Wp1 = Wp; Wj1 = Wj;
    }
  }
  m = wave_min(m);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = sm[0];
    for (int q = 1; q < 4; q++) r = fmin(r, sm[q]);
    part[blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z)] = r;
  }
}

// src/chkdiv.f90:23-39 as the same z march (k_chkdiv below is the
// grid-stride form, used on caller-owned arrays): w(k-1) carried, the i-1 and
// j-1 neighbours from the tile's cache lines.
__global__ __launch_bounds__(256) void k_chkdiv_march(Grid g, const double *__restrict__ U, const double *__restrict__ V,
                                                       const double *__restrict__ W, double *__restrict__ part, int kb) {
  __shared__ double s1[4], s2[4];
  const int i = blockIdx.x * 64 + (int)(threadIdx.x & 63) + 1, j = blockIdx.y * 4 + (int)(threadIdx.x >> 6) + 1;
  const int k0 = blockIdx.z * kb + 1, k1 = min(k0 + kb - 1, g.nz);
  double tot = 0.0, mx = 0.0;
  if (i <= g.nx && j <= g.ny && k0 <= k1) {
    long c = g.at(i, j, k0);
    double Wm = W[c - g.sxy];
    for (int k = k0; k <= k1; k++, c += g.sxy) {
      const double Wc = W[c], Vc = V[c], Vm = V[c - g.sx], Uc = U[c], Um = U[c - 1];
      const double div = (Wc - Wm) * g.dzi + (Vc - Vm) * g.dyi + (Uc - Um) * g.dxi;
      tot += div;
      mx = fmax(mx, fabs(div));
      Wm = Wc;
    }
  }
  tot = wave_sum(tot); mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) { s1[threadIdx.x >> 6] = tot; s2[threadIdx.x >> 6] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    part[2 * b] = ((s1[0] + s1[1]) + (s1[2] + s1[3]));
    part[2 * b + 1] = fmax(fmax(s2[0], s2[1]), fmax(s2[2], s2[3]));
  }
}

// src/chkdiv.f90:23-39
__global__ __launch_bounds__(256) void k_chkdiv(Grid g, const double *__restrict__ U, const double *__restrict__ V,
                                                 const double *__restrict__ W, double *__restrict__ part) {
  __shared__ double s1[4], s2[4];
  const long n = (long)g.nx * g.ny * g.nz;
  double tot = 0.0, mx = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    int i = (int)(t % g.nx) + 1, j = (int)((t / g.nx) % g.ny) + 1, k = (int)(t / ((long)g.nx * g.ny)) + 1;
    long c = g.at(i, j, k);
    double div = (W[c] - W[c - g.sxy]) * g.dzi + (V[c] - V[c - g.sx]) * g.dyi + (U[c] - U[c - 1]) * g.dxi;
    tot += div;
    mx = fmax(mx, fabs(div));
  }
  tot = wave_sum(tot); mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) { s1[threadIdx.x >> 6] = tot; s2[threadIdx.x >> 6] = mx; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); q++) { a += s1[q]; b = fmax(b, s2[q]); }
    part[2 * blockIdx.x] = a; part[2 * blockIdx.x + 1] = b;
  }
}

// Cell-centred output quantities of post3d (src/output.f90:997-1030): pnew;
// w and c averaged to the cell centre; the three vorticity components of
// src/post.f90:15-61, term by term in the reference's order.  Interior only,
// i fastest, into out[(i-1) + nx((j-1) + ny(k-1))].
__global__ __launch_bounds__(BX *BY) void k_post3d(Grid g, int q, const double *__restrict__ U, const double *__restrict__ V,
                                                   const double *__restrict__ W, const double *__restrict__ P,
                                                   const double *__restrict__ C, double *__restrict__ out) {
  const int i = blockIdx.x * BX + threadIdx.x + 1, j = blockIdx.y * BY + threadIdx.y + 1, k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long c = g.at(i, j, k), sx = g.sx, sz = g.sxy;
  const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi;
  double r;
  switch (q) {
    case 0: r = P[c]; break;
    case 1: r = 0.5 * (W[c] + W[c - sz]); break;
    case 2: r = 0.5 * (C[c] + C[c - sz]); break;
    case 3:
      r = 0.25 * ((W[c + sx] - W[c]) * dyi - (V[c + sz] - V[c]) * dzi + (W[c + sx - sz] - W[c - sz]) * dyi -
                  (V[c] - V[c - sz]) * dzi + (W[c] - W[c - sx]) * dyi - (V[c - sx + sz] - V[c - sx]) * dzi +
                  (W[c - sz] - W[c - sx - sz]) * dyi - (V[c - sx] - V[c - sx - sz]) * dzi);
      break;
    case 4:
      r = 0.25 * ((U[c + sz] - U[c]) * dzi - (W[c + 1] - W[c]) * dxi + (U[c] - U[c - sz]) * dzi -
                  (W[c + 1 - sz] - W[c - sz]) * dxi + (U[c - 1 + sz] - U[c - 1]) * dzi - (W[c] - W[c - 1]) * dxi +
                  (U[c - 1] - U[c - 1 - sz]) * dzi - (W[c - sz] - W[c - 1 - sz]) * dxi);
      break;
    default:
      r = 0.25 * ((V[c + 1] - V[c]) * dxi - (U[c + sx] - U[c]) * dyi + (V[c + 1 - sx] - V[c - sx]) * dxi -
                  (U[c] - U[c - sx]) * dyi + (V[c] - V[c - 1]) * dxi - (U[c - 1 + sx] - U[c - 1]) * dyi +
                  (V[c - sx] - V[c - 1 - sx]) * dxi - (U[c - 1] - U[c - 1 - sx]) * dyi);
      break;
  }
  out[(long)(i - 1) + (long)g.nx * ((long)(j - 1) + (long)g.ny * (k - 1))] = r;
}

// one lane stores the device's wall clock (constant rate) into t[slot]
__global__ void k_stamp(unsigned long long *t, int slot) {
  if (threadIdx.x == 0) t[slot] = wall_clock64();
}

}  // namespace

void ipg_stamp(ipg_ctx *c, int slot) { hipLaunchKernelGGL(k_stamp, dim3(1), dim3(64), 0, c->stream, c->stamps, slot); }

static dim3 cell_grid(const Grid &g) { return dim3((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, g.nz); }

int ipg_launch_rk(ipg_ctx *c, int rkiter, bool wrapx) {
  const Grid &g = c->g;
  double **f = c->f;
  const int ntx = (g.nx + RTX - 1) / RTX, nty = (g.ny + RTY - 1) / RTY;
  // z planes per workgroup: enough workgroups (>= 2048) to fill 256 CUs a few
  // times over, but long columns so the LDS ring pays for its prologue
  int kb = c->tune_rk_kb;
  if (kb <= 0) {
    kb = 64;
    while (kb > 4 && (long)ntx * nty * ((g.nz + kb - 1) / kb) < 2048) kb /= 2;
  }
  const int ntz = (g.nz + kb - 1) / kb;
  const dim3 grid(ntx * nty * ntz), blk(RNT);
  const int wx = wrapx ? 1 : 0;
  int rc;
  if (c->mixed) {
    float *r[4] = {f32ptr(c, IPG_DURK), f32ptr(c, IPG_DVRK), f32ptr(c, IPG_DWRK), f32ptr(c, IPG_DCRK)};
    const RkArgs<float> a{g, f[IPG_U], f[IPG_V], f[IPG_W], f[IPG_C], f[IPG_PNEW], f[IPG_DUDT], f[IPG_DVDT], f[IPG_DWDT],
                          f[IPG_DCDT], r[0], r[1], r[2], r[3], c->dscal, ntx, nty, kb, wx};
    if (rkiter == 1) rc = ipg_go<RkK<1, float>>(c, [&](hipStream_t s) { launch_rk<1>(a, grid, s); }, grid, blk, 0, a, "rk_stage");
    else if (rkiter == 2) rc = ipg_go<RkK<2, float>>(c, [&](hipStream_t s) { launch_rk<2>(a, grid, s); }, grid, blk, 0, a, "rk_stage");
    else rc = ipg_go<RkK<3, float>>(c, [&](hipStream_t s) { launch_rk<3>(a, grid, s); }, grid, blk, 0, a, "rk_stage");
  } else {
    const RkArgs<double> a{g, f[IPG_U], f[IPG_V], f[IPG_W], f[IPG_C], f[IPG_PNEW], f[IPG_DUDT], f[IPG_DVDT], f[IPG_DWDT],
                           f[IPG_DCDT], f[IPG_DURK], f[IPG_DVRK], f[IPG_DWRK], f[IPG_DCRK], c->dscal, ntx, nty, kb, wx};
    if (rkiter == 1) rc = ipg_go<RkK<1, double>>(c, [&](hipStream_t s) { launch_rk<1>(a, grid, s); }, grid, blk, 0, a, "rk_stage");
    else if (rkiter == 2) rc = ipg_go<RkK<2, double>>(c, [&](hipStream_t s) { launch_rk<2>(a, grid, s); }, grid, blk, 0, a, "rk_stage");
    else rc = ipg_go<RkK<3, double>>(c, [&](hipStream_t s) { launch_rk<3>(a, grid, s); }, grid, blk, 0, a, "rk_stage");
  }
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_bound(ipg_ctx *c, int mask) {
  BoundJob job{};
  job.xcol = (mask & BND_NO_XCOL) ? 0 : 1;
  auto add = [&](double *f, int kind) { job.f[job.nf] = f; job.kind[job.nf] = kind; job.nf++; };
  if (mask & BND_UVW_NEW) { add(c->f[IPG_U], KIND_UV); add(c->f[IPG_V], KIND_UV); add(c->f[IPG_W], KIND_W); }
  if (mask & BND_PNEW) add(c->f[IPG_PNEW], KIND_P);
  if (mask & BND_C) add(c->f[IPG_C], KIND_C);
  if (mask & BND_UVW_DT) {
    if (job.nf > 2) { ipg_set_error("bound: too many fields"); return 1; }
    add(c->f[IPG_DUDT], KIND_UV); add(c->f[IPG_DVDT], KIND_UV); add(c->f[IPG_DWDT], KIND_W);
  }
  if (mask & BND_P) {
    if (job.nf > 4) { ipg_set_error("bound: too many fields"); return 1; }
    add(c->f[IPG_P], KIND_P);
  }
  if (job.nf == 0) return 0;
  const Grid &g = c->g;
  const long total = 2L * (g.nx + 2) * (g.ny + 2) + 2L * (g.nx + 2) * g.nz + (job.xcol ? 2L * g.ny * g.nz : 0L) + (long)g.nx * g.ny;
  if (total > 2000000000L) { ipg_set_error("bound: grid too large"); return 1; }
  int blocks = (int)((total + 255) / 256);
  if (blocks > 4096) blocks = 4096;
  if (ipg_go<BoundK>(c, k_bound, dim3(blocks), dim3(BND_THREADS), 0, BoundArgs{g, job}, "bound")) return 1;
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_fillps(ipg_ctx *c, int rkiter) {
  const Grid &g = c->g;
  hipLaunchKernelGGL(k_fillps, cell_grid(g), dim3(BX, BY), 0, c->stream, g, c->f[IPG_DUDT], c->f[IPG_DVDT], c->f[IPG_DWDT],
                     c->f[IPG_P], c->dscal, c->pc.rkab[rkiter - 1]);
  IPG_HIP(hipGetLastError());
  ipg_mark(c, "fillps");
  return 0;
}

int ipg_launch_correc(ipg_ctx *c, int rkiter) {
  const Grid &g = c->g;
  const CorrecArgs a{g, c->f[IPG_DUDT], c->f[IPG_DVDT], c->f[IPG_DWDT], c->f[IPG_P], c->f[IPG_U], c->f[IPG_V],
                     c->f[IPG_W], c->f[IPG_PNEW], c->dscal, c->pc.rkab[rkiter - 1]};
  if (ipg_go<CorrecK<false>>(c, k_correc<false>, cell_grid(g), dim3(BX, BY), 0, a, "correc")) return 1;
  IPG_HIP(hipGetLastError());
  return 0;
}

static Pieces pieces_of(ipg_ctx *c) {
  Pieces P{};
  for (int r = 0; r <= IPG_MAX_REFINE; r++) P.q[r] = c->pp[r];
  return P;
}

int ipg_launch_p_combine(ipg_ctx *c) {
  const Grid &g = c->g;
  const Pieces P = pieces_of(c);
  if (c->nref == 0) hipLaunchKernelGGL(k_p_combine<1>, cell_grid(g), dim3(BX, BY), 0, c->stream, g, P, c->f[IPG_P]);
  else if (c->nref == 1) hipLaunchKernelGGL(k_p_combine<2>, cell_grid(g), dim3(BX, BY), 0, c->stream, g, P, c->f[IPG_P]);
  else hipLaunchKernelGGL(k_p_combine<3>, cell_grid(g), dim3(BX, BY), 0, c->stream, g, P, c->f[IPG_P]);
  IPG_HIP(hipGetLastError());
  ipg_mark(c, "p_combine");
  return 0;
}

int ipg_launch_resid(ipg_ctx *c, int npieces) {
  const Grid &g = c->g;
  const Pieces P = pieces_of(c);
  float *r = c->pp[IPG_MAX_REFINE + 1];
  const dim3 grid((g.nx + BX - 1) / BX, (g.ny + BY - 1) / BY, (g.nz + RZ - 1) / RZ);
  const ResidArgs a{g, P, c->f[IPG_P], r};
  const int rc = npieces == 1 ? ipg_go<ResidK<1>>(c, k_resid<1>, grid, dim3(BX, BY), 0, a, "resid")
                              : ipg_go<ResidK<2>>(c, k_resid<2>, grid, dim3(BX, BY), 0, a, "resid");
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_correc_fused(ipg_ctx *c, int rkiter) {
  const Grid &g = c->g;
  int rc;
  if (solve_in_pieces(c)) {
    const CorrecPiecesArgs a{g, c->f[IPG_DUDT], c->f[IPG_DVDT], c->f[IPG_DWDT], pieces_of(c), c->f[IPG_U], c->f[IPG_V],
                             c->f[IPG_W], c->f[IPG_PNEW], c->dscal, c->pc.rkab[rkiter - 1]};
    if (c->nref == 0) rc = ipg_go<CorrecPiecesK<1>>(c, k_correc_pieces<1>, cell_grid(g), dim3(BX, BY), 0, a, "correc");
    else if (c->nref == 1) rc = ipg_go<CorrecPiecesK<2>>(c, k_correc_pieces<2>, cell_grid(g), dim3(BX, BY), 0, a, "correc");
    else rc = ipg_go<CorrecPiecesK<3>>(c, k_correc_pieces<3>, cell_grid(g), dim3(BX, BY), 0, a, "correc");
  } else {
    const CorrecArgs a{g, c->f[IPG_DUDT], c->f[IPG_DVDT], c->f[IPG_DWDT], c->f[IPG_P], c->f[IPG_U], c->f[IPG_V],
                       c->f[IPG_W], c->f[IPG_PNEW], c->dscal, c->pc.rkab[rkiter - 1]};
    rc = ipg_go<CorrecK<true>>(c, k_correc<true>, cell_grid(g), dim3(BX, BY), 0, a, "correc");
  }
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_pnew(ipg_ctx *c) {
  const Grid &g = c->g;
  hipLaunchKernelGGL(k_pnew, cell_grid(g), dim3(BX, BY), 0, c->stream, g, c->f[IPG_P], c->f[IPG_PNEW], c->f[IPG_DCDT], c->f[IPG_C]);
  IPG_HIP(hipGetLastError());
  return 0;
}

// The grid of the z-marching checks: 64 x 4 (i, j) tiles, and enough z
// blocks of kb planes to reach about 1024 workgroups (at most 1024, the
// partials the reduction scratch holds).
static dim3 march_grid(const Grid &g, int &kb) {
  const int bx = (g.nx + 63) / 64, by = (g.ny + 3) / 4;
  const int nkb = std::max(1, std::min(g.nz, 1024 / std::max(1, bx * by)));
  kb = (g.nz + nkb - 1) / nkb;
  return dim3(bx, by, (g.nz + kb - 1) / kb);
}

int ipg_launch_chkdt(ipg_ctx *c, double *dtmax) {
  int kb;
  const dim3 grid = march_grid(c->g, kb);
  const int nb = (int)(grid.x * grid.y * grid.z);
  if (nb > 2048) { ipg_set_error("chkdt: grid too wide for the reduction scratch"); return 1; }
  hipLaunchKernelGGL(k_chkdt_march, grid, dim3(256), 0, c->stream, c->g, c->f[IPG_U], c->f[IPG_V], c->f[IPG_W], c->red,
                     kb);
  IPG_HIP(hipGetLastError());
  std::vector<double> h(nb);
  IPG_HIP(hipMemcpyAsync(h.data(), c->red, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  double m = h[0];
  for (int q = 1; q < nb; q++) m = h[q] < m ? h[q] : m;
  *dtmax = m;
  return 0;
}

int ipg_launch_chkdiv(ipg_ctx *c, double *tot, double *mx) {
  int kb;
  const dim3 grid = march_grid(c->g, kb);
  const int nb = (int)(grid.x * grid.y * grid.z);
  if (nb > 2048) { ipg_set_error("chkdiv: grid too wide for the reduction scratch"); return 1; }
  hipLaunchKernelGGL(k_chkdiv_march, grid, dim3(256), 0, c->stream, c->g, c->f[IPG_U], c->f[IPG_V], c->f[IPG_W], c->red,
                     kb);
  IPG_HIP(hipGetLastError());
  std::vector<double> h(2 * (size_t)nb);
  IPG_HIP(hipMemcpyAsync(h.data(), c->red, 2 * nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  double a = 0.0, b = 0.0;
  for (int q = 0; q < nb; q++) { a += h[2 * q]; b = h[2 * q + 1] > b ? h[2 * q + 1] : b; }
  *tot = a; *mx = b;
  return 0;
}

// One post3d quantity into host memory (interior, i fastest).  The Poisson
// spectrum, dead between substeps, holds the result on the device.
int ipg_launch_post3d(ipg_ctx *c, int q, double *host) {
  const Grid &g = c->g;
  double *out = (double *)c->ps2.spec;
  if (q == IPG_POST_SOL) {
    if (ipg_launch_phase_sol(c, out)) return 1;
  } else {
    hipLaunchKernelGGL(k_post3d, cell_grid(g), dim3(BX, BY), 0, c->stream, g, q, c->f[IPG_U], c->f[IPG_V], c->f[IPG_W],
                       c->f[IPG_PNEW], c->f[IPG_C], out);
  }
  IPG_HIP(hipGetLastError());
  IPG_HIP(hipMemcpyAsync(host, out, (size_t)g.nx * g.ny * g.nz * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  IPG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

// ---- divergence check without allocation ------------------------------------
// ipg_chkdiv_noalloc: src/chkdiv.f90:23-39 (k_chkdiv, as ipg_chkdiv launches
// it) on caller-owned device arrays.  Its per-block partial sums go to a
// file-scope device array rather than to a context's scratch, so a call
// allocates nothing.  Each GPU holds its own copy of the array, at its own
// address, found once per device with hipGetSymbolAddress; host code uses that
// address only.  Every partial the final reduction reads is written by the
// same call, so nothing needs clearing; calls that overlap on one GPU share
// the array and race.
namespace {

constexpr int NOALLOC_BLOCKS = 16384;            // per-block partials the array holds
constexpr long NOALLOC_CELLS_PER_BLOCK = 8192;   // 256 threads, at most 32 cells each
constexpr int NOALLOC_MAX_DEVICES = 64;

// [2 b], [2 b + 1]: block b's total and maximum; the last pair: the result
static __device__ __align__(256) double scratch[2 * NOALLOC_BLOCKS + 2];

std::mutex g_scratch_mu;
double *g_scratch_ptr[NOALLOC_MAX_DEVICES];

// one block: thread t reduces partials t, t + 256, ... in order, then a tree
__global__ __launch_bounds__(256) void k_chkdiv_final(const double *__restrict__ part, int nb, double *__restrict__ out) {
  __shared__ double s1[256], s2[256];
  double a = 0.0, b = 0.0;
  for (int q = threadIdx.x; q < nb; q += 256) { a += part[2 * q]; b = fmax(b, part[2 * q + 1]); }
  s1[threadIdx.x] = a;
  s2[threadIdx.x] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      s1[threadIdx.x] += s1[threadIdx.x + h];
      s2[threadIdx.x] = fmax(s2[threadIdx.x], s2[threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out[0] = s1[0]; out[1] = s2[0]; }
}

}  // namespace

extern "C" int ipg_chkdiv_noalloc(const double *u, const double *v, const double *w, int itot, int jtot, int ktot, double lx,
                                  double ly, double lz, double *divtot, double *divmax) {
  if (!u || !v || !w || !divtot || !divmax) { ipg_set_error("chkdiv_noalloc: null argument"); return 1; }
  if (itot < 1 || jtot < 1 || ktot < 1 || !(lx > 0.0) || !(ly > 0.0) || !(lz > 0.0)) {
    ipg_set_error("chkdiv_noalloc: bad grid");
    return 1;
  }
  const long cells = (long)itot * jtot * ktot;
  const long nb = (cells + NOALLOC_CELLS_PER_BLOCK - 1) / NOALLOC_CELLS_PER_BLOCK;
  if (nb > NOALLOC_BLOCKS) {
    ipg_set_error("chkdiv_noalloc: " + std::to_string(cells) + " cells exceed its maximum of " +
                  std::to_string((long)NOALLOC_BLOCKS * NOALLOC_CELLS_PER_BLOCK) + " (512^3)");
    return 2;
  }
  int dev = 0;
  IPG_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) { ipg_set_error("chkdiv_noalloc: device index out of range"); return 1; }
  double *part;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
      void *p = nullptr;
      IPG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
      g_scratch_ptr[dev] = (double *)p;
    }
    part = g_scratch_ptr[dev];
  }
  // the reference's (0:i1,0:j1,0:k1) layout, i fastest: Grid indexing with
  // the row pitch itot + 2 from element (1, 0, 0)
  Grid g{};
  g.nx = itot; g.ny = jtot; g.nz = ktot;
  g.sx = (long)itot + 2;
  g.sxy = g.sx * ((long)jtot + 2);
  g.dxi = itot / lx; g.dyi = jtot / ly; g.dzi = ktot / lz;
  hipLaunchKernelGGL(k_chkdiv, dim3((unsigned)nb), dim3(256), 0, 0, g, u + 1, v + 1, w + 1, part);
  IPG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_chkdiv_final, dim3(1), dim3(256), 0, 0, part, (int)nb, part + 2 * NOALLOC_BLOCKS);
  IPG_HIP(hipGetLastError());
  double r[2];
  IPG_HIP(hipMemcpy(r, part + 2 * NOALLOC_BLOCKS, sizeof(r), hipMemcpyDeviceToHost));
  *divtot = r[0];
  *divmax = r[1];
  return 0;
}
