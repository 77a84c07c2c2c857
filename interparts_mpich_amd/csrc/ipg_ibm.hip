// ipg_ibm.hip -- direct-forcing immersed boundary block of one substep
// (src/main.f90:286-338): three passes of marker interpolation
// (eulr2lagr, src/interp_spread.f90:12-228), marker force update
// (complagrforces / updtlagrforces, src/forcing.f90:9-391) and spreading
// (lagr2eulr, src/interp_spread.f90:230-474).
//
// The reference copies dudt into a 2-ghost scratch field, zeroes it before
// every spread, scatters with atomics, folds the periodic ghosts and adds the
// saved dudt back (9 full-field passes per IBM pass).  Here the spread
// contributions go, with periodic wrap folded into the index, into sparse
// force buffers G (zero outside the spheres' stencils), and the velocity the
// markers see is dudt + G: the same sums, fl(G + dudtold), without the
// full-field passes.  The last pass spreads straight into dudt.  Work is
// organised one workgroup per (sphere, velocity component): the stencil cells
// of the sphere's box are staged in LDS for the gathers, and the spread is
// summed in LDS before one global atomic per touched cell.  The w-stencil of a marker within h/2 of the top
// wall reaches the k1 = kmax+1 plane (src/interp_spread.f90:100 has no upper
// clamp); that plane never receives dudt and keeps the last spread until the
// next one, so it is held in its own planes K.
#include "ipg_internal.h"
#include <cmath>

namespace {

struct Stencil {
  int ilows, jlows, klows, ilow, jlow, klow;
  double phix[3], phiy[3], phiz[3], phixh[3], phiyh[3], phizh[3];
};

// src/coordsfp.f90:49-68 / src/intgr_nwtn_eulr.f90:1451-1465 marker position
// and the isout test; false when the marker is not kept.
__device__ __forceinline__ bool marker_pos(const Grid &g, double xp, double yp, double zp, double ex, double ey, double ez,
                                           double &xl, double &yl, double &zl) {
  xl = xp + ex; yl = yp + ey; zl = zp + ez;
  // a non-finite or far-away marker (a diverged run) is dropped rather than
  // turned into an out-of-range index
  if (!(fabs(xl) < 1e6 * g.lx && fabs(yl) < 1e6 * g.ly && fabs(zl) < 1e6 * g.lz)) return false;
  double isx = 0.0, isy = 0.0;
  if (xl < 0.0 + 0.5 * g.dx) isx = 1.0;
  if (xl >= g.lx + 0.5 * g.dx) isx = -1.0;
  if (yl < 0.0 + 0.5 * g.dy) isy = 1.0;
  if (yl >= g.ly + 0.5 * g.dy) isy = -1.0;
  const double cx = (xl + isx * g.lx - 0.0) * g.dxi;
  if (nint_f(cx) < 1 || nint_f(cx) > g.nx) return false;
  const double cy = (yl + isy * g.ly - 0.0) * g.dyi;
  if (nint_f(cy) < 1 || nint_f(cy) > g.ny) return false;
  return true;
}

// src/interp_spread.f90:76-130
__device__ __forceinline__ void stencil(const Grid &g, double xfp, double yfp, double zfp, Stencil &t) {
  const double cx = xfp * g.dxi - (double)((int)floor((xfp - 0.5 * g.dx) / g.lx) * g.nx) - 0.0;
  const double cy = yfp * g.dyi - (double)((int)floor((yfp - 0.5 * g.dy) / g.ly) * g.ny) - 0.0;
  const double cz = zfp * g.dzi;
  t.ilows = nint_f(cx) - 1;
  t.jlows = nint_f(cy) - 1;
  t.klows = max(nint_f(cz) - 1, 1);
  t.ilow = (int)floor(cx);
  t.jlow = (int)floor(cy);
  t.klow = max(min((int)floor(cz), g.nz - 2), 1);
#pragma unroll
  for (int q = 0; q < 3; q++) {
    t.phix[q] = roma_kernel(((double)t.ilows + (double)q) - cx);
    t.phiy[q] = roma_kernel(((double)t.jlows + (double)q) - cy);
    t.phiz[q] = roma_kernel(((double)t.klows + (double)q) - cz);
    t.phixh[q] = roma_kernel((((double)t.ilow + (double)q) - 0.5) - cx);
    t.phiyh[q] = roma_kernel((((double)t.jlow + (double)q) - 0.5) - cy);
    t.phizh[q] = roma_kernel((((double)t.klow + (double)q) - 0.5) - cz);
  }
}

struct IbmArgs {
  Grid g;
  int np, nl;
  double dVlagr, dVeul, radfp;
  int bmx, bmy, bmz;  // largest box extent per direction (LDS sizing)
  const double *px, *py, *pz, *pu, *pv, *pw, *pomx, *pomy, *pomz, *pphi;
  const double *mk;   // [3][nl] marker offsets from the sphere centre
  double *fl;         // [3][np*nl] marker forces fxl, fyl, fzl
  double *fld[3];     // dudt dvdt dwdt
  double *Gread[3];   // spread of the previous pass, added to fld when gathering (null at pass 0)
  double *Gspread[3]; // spread target; null: spread straight into fld (last pass)
  double *Gzero[3];   // buffer cleared on this sphere's cells (null: none)
  double *Kread, *Kspread, *Kzero;  // dwdtf(.,.,k1) planes
  const double *dtp;
};

__device__ __forceinline__ long gidx(const Grid &g, int wi, int wj, int k) {
  return (long)(wi - 1) + (long)g.nx * ((long)(wj - 1) + (long)g.ny * (long)(k - 1));
}
// periodic wrap of any unwrapped index onto 1..n
__device__ __forceinline__ int wrapany(int i, int n) {
  int r = i - 1;
  if ((unsigned)r >= (unsigned)n) {
    r %= n;
    if (r < 0) r += n;
  }
  return r + 1;
}

// The per-sphere box of one velocity component: cells [b0, b0+bn) per
// direction in unwrapped index space (z clipped to the channel, 1..nz).  It
// holds every stencil cell of the sphere's markers with one cell to spare.
struct Box { int bi0, bj0, bk0, bx, by, bz; };
__device__ __forceinline__ Box sphere_box(const IbmArgs &A, double xp, double yp, double zp) {
  const Grid &g = A.g;
  Box b;
  const int x0 = (int)floor((xp - A.radfp) * g.dxi), x1 = (int)floor((xp + A.radfp) * g.dxi);
  const int y0 = (int)floor((yp - A.radfp) * g.dyi), y1 = (int)floor((yp + A.radfp) * g.dyi);
  const int z0 = (int)floor((zp - A.radfp) * g.dzi), z1 = (int)floor((zp + A.radfp) * g.dzi);
  b.bi0 = x0 - 2; b.bx = min(x1 + 3 - b.bi0 + 1, A.bmx);
  b.bj0 = y0 - 2; b.by = min(y1 + 3 - b.bj0 + 1, A.bmy);
  b.bk0 = max(z0 - 2, 1);
  b.bz = min(min(z1 + 3, g.nz) - b.bk0 + 1, A.bmz);
  if (b.bz < 0) b.bz = 0;
  return b;
}

// local box index of an unwrapped cell, -1 when outside the box
__device__ __forceinline__ int box_at(const Box &b, int iu, int ju, int k) {
  const int li = iu - b.bi0, lj = ju - b.bj0, lk = k - b.bk0;
  if ((unsigned)li >= (unsigned)b.bx || (unsigned)lj >= (unsigned)b.by || (unsigned)lk >= (unsigned)b.bz) return -1;
  return li + b.bx * (lj + b.by * lk);
}

// Marker stencil of one component in unwrapped indices: base cell (i0,j0,k0)
// and the three weights per direction, in the reference's order
// (src/interp_spread.f90:98-130).
struct CompStencil { int i0, j0, k0; double wx[3], wy[3], wz[3]; };
__device__ __forceinline__ void comp_stencil(const Grid &g, double xl, double yl, double zl, int comp, CompStencil &c) {
  Stencil s;
  stencil(g, xl, yl, zl, s);
  const int sx = (int)floor((xl - 0.5 * g.dx) / g.lx) * g.nx;
  const int sy = (int)floor((yl - 0.5 * g.dy) / g.ly) * g.ny;
  if (comp == 0) {
    c.i0 = s.ilows + sx; c.j0 = s.jlow + sy; c.k0 = s.klow;
    for (int q = 0; q < 3; q++) { c.wx[q] = s.phix[q]; c.wy[q] = s.phiyh[q]; c.wz[q] = s.phizh[q]; }
  } else if (comp == 1) {
    c.i0 = s.ilow + sx; c.j0 = s.jlows + sy; c.k0 = s.klow;
    for (int q = 0; q < 3; q++) { c.wx[q] = s.phixh[q]; c.wy[q] = s.phiy[q]; c.wz[q] = s.phizh[q]; }
  } else {
    c.i0 = s.ilow + sx; c.j0 = s.jlow + sy; c.k0 = s.klows;
    for (int q = 0; q < 3; q++) { c.wx[q] = s.phixh[q]; c.wy[q] = s.phiyh[q]; c.wz[q] = s.phiz[q]; }
  }
}

// Marker preparation, once per substep (markers do not move during the IBM
// block): the validity flag (the isout test of src/coordsfp.f90:55-65), the
// unwrapped stencil base of each velocity component, the nine stencil weights
// of each component and the rigid-body velocity at the marker
// (src/intgr_nwtn_eulr.f90:1471-1476).  Layout (SoA over markers m):
// mw[(comp*9 + q)*M + m] weights (q = 0..2 x, 3..5 y, 6..8 z), mb[(comp*3+d)*M + m]
// bases, mv[comp*M + m] rigid velocity, mb[9*M + m] validity.
__global__ __launch_bounds__(256) void k_ibm_prep(IbmArgs A, double *mw, int *mb, double *mv) {
  const Grid &g = A.g;
  const long M = (long)A.np * A.nl;
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int p = (int)(m / A.nl), l = (int)(m - (long)p * A.nl);
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  double xl, yl, zl;
  const bool ok = marker_pos(g, xp, yp, zp, A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl);
  mb[9 * M + m] = ok ? 1 : 0;
  if (!ok) return;
  for (int comp = 0; comp < 3; comp++) {
    CompStencil c;
    comp_stencil(g, xl, yl, zl, comp, c);
    mb[(comp * 3 + 0) * M + m] = c.i0;
    mb[(comp * 3 + 1) * M + m] = c.j0;
    mb[(comp * 3 + 2) * M + m] = c.k0;
    for (int q = 0; q < 3; q++) {
      mw[(comp * 9 + q) * M + m] = c.wx[q];
      mw[(comp * 9 + 3 + q) * M + m] = c.wy[q];
      mw[(comp * 9 + 6 + q) * M + m] = c.wz[q];
    }
  }
  mv[m] = A.pu[p] + A.pomy[p] * (zl - zp) - A.pomz[p] * (yl - yp);
  mv[M + m] = A.pv[p] + A.pomz[p] * (xl - xp) - A.pomx[p] * (zl - zp);
  mv[2 * M + m] = A.pw[p] + A.pomx[p] * (yl - yp) - A.pomy[p] * (xl - xp);
}

// One IBM pass for one (sphere, velocity component) workgroup.  The cells of
// the sphere's box that some marker's stencil touches are flagged first (one
// bit per cell in LDS); only those are staged from (dudt + previous spread),
// which is a shell about 2.5 cells thick, not the whole box.  Every marker
// gathers its 27-point stencil from LDS (src/interp_spread.f90:132-154) and
// updates its force (src/forcing.f90:63-65 / 255-257); then the markers'
// spread (src/interp_spread.f90:329-355) is summed in the same LDS box with
// LDS atomics and the flagged cells are added to the global target once.
// gather/spread select the halves; the last pass runs them as two launches
// because its spread goes straight into dudt, which other spheres may still
// be gathering from.
__global__ __launch_bounds__(256) void k_ibm_pass(IbmArgs A, const double *__restrict__ mw, const int *__restrict__ mb,
                                                  const double *__restrict__ mv, int pass, int gather, int spread) {
  extern __shared__ double box[];
  const Grid &g = A.g;
  const int p = blockIdx.x / 3, comp = blockIdx.x % 3;
  const long M = (long)A.np * A.nl;
  if (A.Kzero) {
    const long nplane = (long)g.nx * g.ny;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nplane; t += (long)gridDim.x * blockDim.x) A.Kzero[t] = 0.0;
  }
  const double dt = *A.dtp, dti = 1.0 / dt;
  const double c0 = (A.dVlagr / A.dVeul) * dt;
  const Box b = sphere_box(A, A.px[p], A.py[p], A.pz[p]);
  const int nbox = b.bx * b.by * b.bz, bxy = b.bx * b.by;
  unsigned *flag = (unsigned *)(box + A.bmx * A.bmy * A.bmz);
  const int nflag = (nbox + 31) >> 5;
  const double *F = A.fld[comp];
  double *fl = A.fl + comp * M;
  const long m0 = (long)p * A.nl;
  const double *W = mw + (long)comp * 9 * M;
  const int *I0 = mb + (long)comp * 3 * M, *J0 = I0 + M, *K0 = J0 + M, *OK = mb + 9 * M;
  for (int t = threadIdx.x; t < nflag; t += blockDim.x) flag[t] = 0u;
  __syncthreads();
  for (int l = threadIdx.x; l < A.nl; l += blockDim.x) {
    const long m = m0 + l;
    if (!OK[m]) continue;
    const int i0 = I0[m], j0 = J0[m], k0 = K0[m];
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          const int q = box_at(b, i0 + i, j0 + j, k0 + k);
          if (q >= 0) atomicOr(&flag[q >> 5], 1u << (q & 31));
        }
  }
  __syncthreads();
  auto flagged = [&](int t) { return (flag[t >> 5] >> (t & 31)) & 1u; };
  auto cell_of = [&](int t, int &wi, int &wj, int &k) {
    const int lk = t / bxy, r = t - lk * bxy, lj = r / b.bx, li = r - lj * b.bx;
    wi = wrapany(b.bi0 + li, g.nx); wj = wrapany(b.bj0 + lj, g.ny); k = b.bk0 + lk;
  };
  if (gather) {
    // staged in chunks of 8 cells per thread so that the loads of a chunk are
    // all in flight together
    const double *Gr = A.Gread[comp];
    constexpr int CH = 8;
    for (int t0 = threadIdx.x; t0 < nbox; t0 += CH * (int)blockDim.x) {
      double v[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * (int)blockDim.x;
        v[u] = 0.0;
        if (t < nbox && flagged(t)) {
          int wi, wj, k;
          cell_of(t, wi, wj, k);
          v[u] = F[g.at(wi, wj, k)];
          if (Gr) v[u] = v[u] + Gr[gidx(g, wi, wj, k)];
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * (int)blockDim.x;
        if (t < nbox) box[t] = v[u];
      }
    }
    __syncthreads();
    for (int l = threadIdx.x; l < A.nl; l += blockDim.x) {
      const long m = m0 + l;
      if (!OK[m]) continue;
      const int i0 = I0[m], j0 = J0[m], k0 = K0[m];
      double wx[3], wy[3], wz[3];
      for (int q = 0; q < 3; q++) { wx[q] = W[q * M + m]; wy[q] = W[(3 + q) * M + m]; wz[q] = W[(6 + q) * M + m]; }
      double acc = 0.0;
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            const int iu = i0 + i, ju = j0 + j, kk = k0 + k;
            double f;
            const int q = box_at(b, iu, ju, kk);
            if (q >= 0) f = box[q];
            else if (kk >= 1 && kk <= g.nz) {
              const int wi = wrapany(iu, g.nx), wj = wrapany(ju, g.ny);
              f = F[g.at(wi, wj, kk)];
              if (A.Gread[comp]) f = f + A.Gread[comp][gidx(g, wi, wj, kk)];
            } else if (kk == g.nz + 1 && comp == 2) {
              f = A.Kread[(long)(wrapany(iu, g.nx) - 1) + (long)g.nx * (wrapany(ju, g.ny) - 1)];
            } else f = 0.0;
            acc = acc + f * wx[i] * wy[j] * wz[k];
          }
      const double ub = mv[comp * M + m];
      if (pass == 0) fl[m] = (ub - acc) * dti;
      else fl[m] = fl[m] + (ub - acc) * dti;
    }
    __syncthreads();
  }
  if (spread) {
    double *Gw = A.Gspread[comp];
    for (int t = threadIdx.x; t < nbox; t += blockDim.x) box[t] = 0.0;
    __syncthreads();
    for (int l = threadIdx.x; l < A.nl; l += blockDim.x) {
      const long m = m0 + l;
      if (!OK[m]) continue;
      const int i0 = I0[m], j0 = J0[m], k0 = K0[m];
      double wx[3], wy[3], wz[3];
      for (int q = 0; q < 3; q++) { wx[q] = W[q * M + m]; wy[q] = W[(3 + q) * M + m]; wz[q] = W[(6 + q) * M + m]; }
      const double fc = fl[m] * c0;
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            const int iu = i0 + i, ju = j0 + j, kk = k0 + k;
            const double v = fc * wx[i] * wy[j] * wz[k];
            const int q = box_at(b, iu, ju, kk);
            if (q >= 0) unsafeAtomicAdd(&box[q], v);
            else if (kk >= 1 && kk <= g.nz) {
              const int wi = wrapany(iu, g.nx), wj = wrapany(ju, g.ny);
              if (Gw) unsafeAtomicAdd(&Gw[gidx(g, wi, wj, kk)], v);
              else unsafeAtomicAdd(&A.fld[comp][g.at(wi, wj, kk)], v);
            } else if (kk == g.nz + 1 && comp == 2) {
              unsafeAtomicAdd(&A.Kspread[(long)(wrapany(iu, g.nx) - 1) + (long)g.nx * (wrapany(ju, g.ny) - 1)], v);
            }
          }
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nbox; t += blockDim.x) {
      if (!flagged(t)) continue;
      const double v = box[t];
      if (v != 0.0) {
        int wi, wj, k;
        cell_of(t, wi, wj, k);
        if (Gw) unsafeAtomicAdd(&Gw[gidx(g, wi, wj, k)], v);
        else unsafeAtomicAdd(&A.fld[comp][g.at(wi, wj, k)], v);
      }
    }
  }
  if (A.Gzero[comp]) {
    // the cells this sphere spread into in an earlier pass: its flagged cells
    // (the markers have not moved) and any stencil cell outside the box
    double *Gz = A.Gzero[comp];
    for (int t = threadIdx.x; t < nbox; t += blockDim.x) {
      if (!flagged(t)) continue;
      int wi, wj, k;
      cell_of(t, wi, wj, k);
      Gz[gidx(g, wi, wj, k)] = 0.0;
    }
    for (int l = threadIdx.x; l < A.nl; l += blockDim.x) {
      const long m = m0 + l;
      if (!OK[m]) continue;
      const int i0 = I0[m], j0 = J0[m], k0 = K0[m];
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            const int iu = i0 + i, ju = j0 + j, kk = k0 + k;
            if (box_at(b, iu, ju, kk) < 0 && kk >= 1 && kk <= g.nz) Gz[gidx(g, wrapany(iu, g.nx), wrapany(ju, g.ny), kk)] = 0.0;
          }
    }
  }
}

// Per-sphere force and torque sums after the last pass, src/forcing.f90:258-287,
// one block per sphere.
__global__ __launch_bounds__(256) void k_ibm_totals(IbmArgs A, double *fxt, double *fyt, double *fzt, double *tqx, double *tqy,
                                                    double *tqz, double *tqt) {
  const Grid &g = A.g;
  if (A.Kzero) {
    const long nplane = (long)g.nx * g.ny;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nplane; t += (long)gridDim.x * blockDim.x) A.Kzero[t] = 0.0;
  }
  const int p = blockIdx.x;
  const long M = (long)A.np * A.nl;
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int l = threadIdx.x; l < A.nl; l += blockDim.x) {
    double xl, yl, zl;
    if (!marker_pos(g, xp, yp, zp, A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl)) continue;
    const long m = (long)p * A.nl + l;
    const double fx = A.fl[m], fy = A.fl[M + m], fz = A.fl[2 * M + m];
    s[0] += fx; s[1] += fy; s[2] += fz;
    s[3] += (yl - yp) * fz - (zl - zp) * fy;
    s[4] += (zl - zp) * fx - (xl - xp) * fz;
    s[5] += (xl - xp) * fy - (yl - yp) * fx;
  }
  __shared__ double sh[6][4];
  for (int q = 0; q < 6; q++) {
    double v = s[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) sh[q][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r[6];
    for (int q = 0; q < 6; q++) { r[q] = 0.0; for (int w = 0; w < (int)(blockDim.x >> 6); w++) r[q] += sh[q][w]; }
    fxt[p] = r[0] * A.dVlagr; fyt[p] = r[1] * A.dVlagr; fzt[p] = r[2] * A.dVlagr;
    tqx[p] = r[3] * A.dVlagr; tqy[p] = r[4] * A.dVlagr; tqz[p] = r[5] * A.dVlagr;
    tqt[p] = (tqy[p] * cos(A.pphi[p])) - (tqx[p] * sin(A.pphi[p]));
  }
}

}  // namespace

// LDS of the per-sphere box (one velocity component at a time), set once at
// marker set-up so that graph capture never meets an attribute change.
int ipg_ibm_setup(ipg_ctx *c) {
  const double r2 = 2.0 * c->pc.radfp;
  const size_t cells = (size_t)((int)std::ceil(r2 * c->g.dxi) + 7) * ((int)std::ceil(r2 * c->g.dyi) + 7) *
                       ((int)std::ceil(r2 * c->g.dzi) + 7);
  const size_t lds = cells * sizeof(double) + (cells + 31) / 32 * sizeof(unsigned);
  if (lds > 160 * 1024) { ipg_set_error("ibm: sphere box exceeds the 160 KiB LDS of a CU"); return 1; }
  IPG_HIP(hipFuncSetAttribute((const void *)k_ibm_pass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  c->ibm_lds_set = true;
  return 0;
}

namespace {

// Common arguments of the IBM launches for this context.
bool ibm_args(ipg_ctx *c, IbmArgs &A, size_t &lds) {
  if (!c->markers_ready || !c->ibm_lds_set) { ipg_set_error("ibm: ipg_coordsfp has not been called"); return false; }
  A = IbmArgs{};
  A.g = c->g; A.np = c->pc.np; A.nl = c->pc.nl; A.dVlagr = c->pc.dVlagr; A.dVeul = c->pc.dVeul; A.radfp = c->pc.radfp;
  const double r2 = 2.0 * c->pc.radfp;
  A.bmx = (int)std::ceil(r2 * c->g.dxi) + 7;
  A.bmy = (int)std::ceil(r2 * c->g.dyi) + 7;
  A.bmz = (int)std::ceil(r2 * c->g.dzi) + 7;
  const size_t cells = (size_t)A.bmx * A.bmy * A.bmz;
  lds = cells * sizeof(double) + (cells + 31) / 32 * sizeof(unsigned);
  double **q = c->ps.q;
  A.px = q[IPG_PX]; A.py = q[IPG_PY]; A.pz = q[IPG_PZ]; A.pu = q[IPG_PU]; A.pv = q[IPG_PV]; A.pw = q[IPG_PW];
  A.pomx = q[IPG_POMX]; A.pomy = q[IPG_POMY]; A.pomz = q[IPG_POMZ]; A.pphi = q[IPG_PPHI];
  A.mk = c->mk; A.fl = c->fl;
  A.fld[0] = c->f[IPG_DUDT]; A.fld[1] = c->f[IPG_DVDT]; A.fld[2] = c->f[IPG_DWDT];
  A.dtp = c->dscal;
  return true;
}

// Buffer roles of pass `pass` (fixed, so a captured graph replays them):
// pass 0 gathers dudt and spreads into GA; pass 1 gathers dudt+GA and spreads
// into GB; pass 2 gathers dudt+GB (clearing GA), then spreads into dudt
// (clearing GB).  The k1 plane: K0 carries it between substeps, K1/K2 hold
// passes 0/1; K2 is cleared by the last force sums.
void ibm_roles(ipg_ctx *c, IbmArgs &A, int pass, int gather, int spread) {
  double **GA = c->G[0], **GB = c->G[1], **K = c->K1;
  double **rd = nullptr, **sp = nullptr, **zr = nullptr;
  double *kr = nullptr, *ks = nullptr, *kz = nullptr;
  if (pass == 0) { kr = K[0]; sp = GA; ks = K[1]; }
  else if (pass == 1) { rd = GA; kr = K[1]; sp = GB; ks = K[2]; }
  else {
    if (gather) { rd = GB; kr = K[2]; zr = GA; kz = K[0]; }
    else { ks = K[0]; zr = GB; kz = K[1]; }
  }
  if (!gather) { rd = nullptr; kr = nullptr; }
  if (!spread) { sp = nullptr; ks = nullptr; }
  for (int d = 0; d < 3; d++) {
    A.Gread[d] = rd ? rd[d] : nullptr;
    A.Gspread[d] = sp ? sp[d] : nullptr;
    A.Gzero[d] = zr ? zr[d] : nullptr;
  }
  A.Kread = kr; A.Kspread = ks; A.Kzero = kz;
}

}  // namespace

// src/main.f90:301-303 (dudtf = dudt): the marker stencils of this substep.
int ipg_launch_ibm_begin(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  const long M = (long)A.np * A.nl;
  hipLaunchKernelGGL(k_ibm_prep, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, c->stream, A, c->mw, c->mb, c->mw + 27 * M);
  ipg_mark(c, "ibm_prep");
  IPG_HIP(hipGetLastError());
  return 0;
}

// One IBM pass or half pass: gather (+ marker force update) and/or spread.
int ipg_launch_ibm_pass(ipg_ctx *c, int pass, int gather, int spread) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  if (pass == 2 && gather && spread) { ipg_set_error("ibm: the last pass gathers and spreads in two launches"); return 1; }
  ibm_roles(c, A, pass, gather, spread);
  const long M = (long)A.np * A.nl;
  hipLaunchKernelGGL(k_ibm_pass, dim3(3 * A.np), dim3(256), lds, c->stream, A, c->mw, c->mb, c->mw + 27 * M, pass, gather,
                     spread);
  static const char *names[3][2] = {{"ibm_pass0_spread", "ibm_pass0"}, {"ibm_pass1_spread", "ibm_pass1"},
                                    {"ibm_pass2_spread", "ibm_pass2_gather"}};
  ipg_mark(c, gather ? (spread ? names[pass][1] : (pass == 2 ? names[2][1] : "ibm_gather")) : names[pass][0]);
  IPG_HIP(hipGetLastError());
  return 0;
}

// Per-sphere force and torque sums (src/forcing.f90:66-89, 258-287); after the
// last pass this also clears the pass-1 k1 plane.
int ipg_launch_ibm_totals(ipg_ctx *c, int last) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  for (int d = 0; d < 3; d++) { A.Gread[d] = A.Gspread[d] = A.Gzero[d] = nullptr; }
  A.Kread = A.Kspread = nullptr;
  A.Kzero = last ? c->K1[2] : nullptr;
  double **q = c->ps.q;
  hipLaunchKernelGGL(k_ibm_totals, dim3(A.np), dim3(256), 0, c->stream, A, q[IPG_PFXLTOT], q[IPG_PFYLTOT], q[IPG_PFZLTOT],
                     q[IPG_PTORQX], q[IPG_PTORQY], q[IPG_PTORQZ], q[IPG_PTORQTHETA]);
  ipg_mark(c, "ibm_totals");
  IPG_HIP(hipGetLastError());
  return 0;
}

// The whole IBM block of one substep, src/main.f90:296-338.
int ipg_launch_ibm(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  if (ipg_launch_ibm_begin(c)) return 1;
  if (ipg_launch_ibm_pass(c, 0, 1, 1)) return 1;
  if (ipg_launch_ibm_pass(c, 1, 1, 1)) return 1;
  if (ipg_launch_ibm_pass(c, 2, 1, 0)) return 1;
  if (ipg_launch_ibm_pass(c, 2, 0, 1)) return 1;
  return ipg_launch_ibm_totals(c, 1);
}
