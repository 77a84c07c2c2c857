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
// (src/interp_spread.f90:76-130).  Only the 9 weights the component uses are
// evaluated; each is the same expression stencil() evaluates.
struct CompStencil { int i0, j0, k0; double wx[3], wy[3], wz[3]; };
__device__ __forceinline__ void comp_stencil(const Grid &g, double xl, double yl, double zl, int comp, CompStencil &c) {
  const int sxw = (int)floor((xl - 0.5 * g.dx) / g.lx) * g.nx;
  const int syw = (int)floor((yl - 0.5 * g.dy) / g.ly) * g.ny;
  const double cx = xl * g.dxi - (double)sxw - 0.0;
  const double cy = yl * g.dyi - (double)syw - 0.0;
  const double cz = zl * g.dzi;
  // component 0 sits on x faces: the x weights are the face (s) ones, y and z
  // the half-shifted ones; likewise for 1 (y) and 2 (z)
  if (comp == 0) {
    c.i0 = nint_f(cx) - 1; c.j0 = (int)floor(cy); c.k0 = max(min((int)floor(cz), g.nz - 2), 1);
  } else if (comp == 1) {
    c.i0 = (int)floor(cx); c.j0 = nint_f(cy) - 1; c.k0 = max(min((int)floor(cz), g.nz - 2), 1);
  } else {
    c.i0 = (int)floor(cx); c.j0 = (int)floor(cy); c.k0 = max(nint_f(cz) - 1, 1);
  }
#pragma unroll
  for (int q = 0; q < 3; q++) {
    c.wx[q] = comp == 0 ? roma_kernel(((double)c.i0 + (double)q) - cx) : roma_kernel((((double)c.i0 + (double)q) - 0.5) - cx);
    c.wy[q] = comp == 1 ? roma_kernel(((double)c.j0 + (double)q) - cy) : roma_kernel((((double)c.j0 + (double)q) - 0.5) - cy);
    c.wz[q] = comp == 2 ? roma_kernel(((double)c.k0 + (double)q) - cz) : roma_kernel((((double)c.k0 + (double)q) - 0.5) - cz);
  }
  c.i0 += sxw;
  c.j0 += syw;
}

// One IBM pass for one (sphere, velocity component) workgroup of IBM_THREADS
// threads; thread t owns markers t, t + IBM_THREADS, ... (at most IBM_MPT).
// Each marker's position (src/coordsfp.f90:49-68), stencil base and nine
// weights (src/interp_spread.f90:76-130) and rigid-body velocity
// (src/intgr_nwtn_eulr.f90:1471-1476) are computed in registers: the markers
// do not move inside the IBM block, and recomputing costs less than reading a
// stored table.  The cells of the sphere's box that some stencil touches are
// flagged (one bit per cell in LDS); only those are staged from (dudt +
// previous spread), a shell about 2.5 cells thick.  Every marker gathers its
// 27-point stencil from LDS (src/interp_spread.f90:132-154) and updates its
// force (src/forcing.f90:63-65 / 255-257); the markers' spread
// (src/interp_spread.f90:329-355) is summed in the same LDS box with LDS
// atomics and the flagged cells are added to the global target once.
// gather/spread select the halves; the last pass runs them as two launches
// because its spread goes straight into dudt, which other spheres may still
// be gathering from.
constexpr int IBM_THREADS = 512, IBM_MPT = 2;

// iso (may be null) marks the isolated spheres, which k_ibm_local handles:
// their workgroups return at once.
__global__ __launch_bounds__(IBM_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_ibm_pass(IbmArgs A, int pass, int gather, int spread,
                                                                                                 const int *__restrict__ iso) {
  extern __shared__ double box[];
  const Grid &g = A.g;
  const int p = blockIdx.x / 3, comp = blockIdx.x % 3;
  const long M = (long)A.np * A.nl;
  if (A.Kzero) {
    const long nplane = (long)g.nx * g.ny;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nplane; t += (long)gridDim.x * blockDim.x) A.Kzero[t] = 0.0;
  }
  if (iso && iso[p]) return;
  const double dt = *A.dtp, dti = 1.0 / dt;
  const double c0 = (A.dVlagr / A.dVeul) * dt;
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  const Box b = sphere_box(A, xp, yp, zp);
  const int nbox = b.bx * b.by * b.bz, bxy = b.bx * b.by;
  unsigned *flag = (unsigned *)(box + A.bmx * A.bmy * A.bmz);
  const int nflag = (nbox + 31) >> 5;
  const double *F = A.fld[comp];
  double *fl = A.fl + comp * M;
  const long m0 = (long)p * A.nl;
  // this thread's markers
  CompStencil cs[IBM_MPT];
  bool ok[IBM_MPT];
  double ub[IBM_MPT];
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    const int l = (int)threadIdx.x + q * IBM_THREADS;
    double xl = 0.0, yl = 0.0, zl = 0.0;
    ok[q] = l < A.nl && marker_pos(g, xp, yp, zp, A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl);
    ub[q] = 0.0;
    if (ok[q]) {
      comp_stencil(g, xl, yl, zl, comp, cs[q]);
      if (comp == 0) ub[q] = A.pu[p] + A.pomy[p] * (zl - zp) - A.pomz[p] * (yl - yp);
      else if (comp == 1) ub[q] = A.pv[p] + A.pomz[p] * (xl - xp) - A.pomx[p] * (zl - zp);
      else ub[q] = A.pw[p] + A.pomx[p] * (yl - yp) - A.pomy[p] * (xl - xp);
    }
  }
  for (int t = threadIdx.x; t < nflag; t += blockDim.x) flag[t] = 0u;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    if (!ok[q]) continue;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const int e = box_at(b, cs[q].i0 + i, cs[q].j0 + j, cs[q].k0 + k);
          if (e >= 0) atomicOr(&flag[e >> 5], 1u << (e & 31));
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
#pragma unroll
    for (int q = 0; q < IBM_MPT; q++) {
      if (!ok[q]) continue;
      const long m = m0 + (int)threadIdx.x + q * IBM_THREADS;
      const CompStencil &c = cs[q];
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const int iu = c.i0 + i, ju = c.j0 + j, kk = c.k0 + k;
            double f;
            const int e = box_at(b, iu, ju, kk);
            if (e >= 0) f = box[e];
            else if (kk >= 1 && kk <= g.nz) {
              const int wi = wrapany(iu, g.nx), wj = wrapany(ju, g.ny);
              f = F[g.at(wi, wj, kk)];
              if (A.Gread[comp]) f = f + A.Gread[comp][gidx(g, wi, wj, kk)];
            } else if (kk == g.nz + 1 && comp == 2) {
              f = A.Kread[(long)(wrapany(iu, g.nx) - 1) + (long)g.nx * (wrapany(ju, g.ny) - 1)];
            } else f = 0.0;
            acc = acc + f * c.wx[i] * c.wy[j] * c.wz[k];
          }
      if (pass == 0) fl[m] = (ub[q] - acc) * dti;
      else fl[m] = fl[m] + (ub[q] - acc) * dti;
    }
    __syncthreads();
  }
  if (spread) {
    double *Gw = A.Gspread[comp];
    for (int t = threadIdx.x; t < nbox; t += blockDim.x) box[t] = 0.0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IBM_MPT; q++) {
      if (!ok[q]) continue;
      const long m = m0 + (int)threadIdx.x + q * IBM_THREADS;
      const CompStencil &c = cs[q];
      const double fc = fl[m] * c0;
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const int iu = c.i0 + i, ju = c.j0 + j, kk = c.k0 + k;
            const double v = fc * c.wx[i] * c.wy[j] * c.wz[k];
            const int e = box_at(b, iu, ju, kk);
            if (e >= 0) unsafeAtomicAdd(&box[e], v);
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
#pragma unroll
    for (int q = 0; q < IBM_MPT; q++) {
      if (!ok[q]) continue;
      const CompStencil &c = cs[q];
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const int iu = c.i0 + i, ju = c.j0 + j, kk = c.k0 + k;
            if (box_at(b, iu, ju, kk) < 0 && kk >= 1 && kk <= g.nz) Gz[gidx(g, wrapany(iu, g.nx), wrapany(ju, g.ny), kk)] = 0.0;
          }
    }
  }
}

// The three IBM passes (src/main.f90:299-334) of an isolated sphere
// (k_ibm_isolate): its spreads touch no cell any other sphere reads or writes,
// so the workgroup keeps them in LDS between the passes and adds the last
// one to dudt at the end.  Workgroups of the other spheres return at once.
// Every stencil of an isolated sphere lies inside its box (the box spans
// the markers' cells with two cells to spare and stops below the k1 plane),
// so a marker's 27 cells are e0 + i + bx (j + by k) in LDS.
__global__ __launch_bounds__(IBM_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_ibm_local(IbmArgs A, const int *__restrict__ iso) {
  extern __shared__ double box[];
  const Grid &g = A.g;
  const int p = blockIdx.x / 3, comp = blockIdx.x % 3;
  const long M = (long)A.np * A.nl;
  if (!iso[p]) return;
  const double dt = *A.dtp, dti = 1.0 / dt;
  const double c0 = (A.dVlagr / A.dVeul) * dt;
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  const Box b = sphere_box(A, xp, yp, zp);
  const int nbox = b.bx * b.by * b.bz, bxy = b.bx * b.by;
  unsigned *flag = (unsigned *)(box + A.bmx * A.bmy * A.bmz);
  const int nflag = (nbox + 31) >> 5;
  const double *F = A.fld[comp];
  double *Fw = A.fld[comp];
  double *fl = A.fl + comp * M;
  const long m0 = (long)p * A.nl;
  double wx[IBM_MPT][3], wy[IBM_MPT][3], wz[IBM_MPT][3], ub[IBM_MPT];
  int e0[IBM_MPT];
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    const int l = (int)threadIdx.x + q * IBM_THREADS;
    double xl = 0.0, yl = 0.0, zl = 0.0;
    e0[q] = -1;
    ub[q] = 0.0;
    if (l < A.nl && marker_pos(g, xp, yp, zp, A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl)) {
      CompStencil c;
      comp_stencil(g, xl, yl, zl, comp, c);
      for (int r = 0; r < 3; r++) { wx[q][r] = c.wx[r]; wy[q][r] = c.wy[r]; wz[q][r] = c.wz[r]; }
      // inside by construction for an isolated sphere; a marker that were not
      // would be left out rather than index outside the box
      if (box_at(b, c.i0, c.j0, c.k0) >= 0 && box_at(b, c.i0 + 2, c.j0 + 2, c.k0 + 2) >= 0)
        e0[q] = box_at(b, c.i0, c.j0, c.k0);
      if (comp == 0) ub[q] = A.pu[p] + A.pomy[p] * (zl - zp) - A.pomz[p] * (yl - yp);
      else if (comp == 1) ub[q] = A.pv[p] + A.pomz[p] * (xl - xp) - A.pomx[p] * (zl - zp);
      else ub[q] = A.pw[p] + A.pomx[p] * (yl - yp) - A.pomy[p] * (xl - xp);
    }
  }
  for (int t = threadIdx.x; t < nflag; t += blockDim.x) flag[t] = 0u;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    if (e0[q] < 0) continue;
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          const int e = e0[q] + i + b.bx * j + bxy * k;
          atomicOr(&flag[e >> 5], 1u << (e & 31));
        }
  }
  __syncthreads();
  auto flagged = [&](int t) { return (flag[t >> 5] >> (t & 31)) & 1u; };
  auto cell_of = [&](int t) {
    const int lk = t / bxy, r = t - lk * bxy, lj = r / b.bx, li = r - lj * b.bx;
    return g.at(wrapany(b.bi0 + li, g.nx), wrapany(b.bj0 + lj, g.ny), b.bk0 + lk);
  };
  // Before each gather the box holds the previous pass's spread S (zero at
  // pass 0) and becomes fl(dudt + S) on the flagged cells, the dudtf the
  // reference gathers from; after each spread it holds the new S.
  constexpr int CH = 8;
#pragma unroll 1
  for (int it = 0; it < 3; it++) {
    for (int t0 = threadIdx.x; t0 < nbox; t0 += CH * (int)blockDim.x) {
      double v[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * (int)blockDim.x;
        v[u] = 0.0;
        if (t < nbox && flagged(t)) {
          v[u] = F[cell_of(t)];
          if (it > 0) v[u] = v[u] + box[t];
        }
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * (int)blockDim.x;
        if (t < nbox) box[t] = v[u];
      }
    }
    __syncthreads();
    double fnew[IBM_MPT];
#pragma unroll
    for (int q = 0; q < IBM_MPT; q++) {
      fnew[q] = 0.0;
      if (e0[q] < 0) continue;
      const long m = m0 + (int)threadIdx.x + q * IBM_THREADS;
      double acc = 0.0;
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) acc = acc + box[e0[q] + i + b.bx * j + bxy * k] * wx[q][i] * wy[q][j] * wz[q][k];
      fnew[q] = it == 0 ? (ub[q] - acc) * dti : fl[m] + (ub[q] - acc) * dti;
      fl[m] = fnew[q];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < nbox; t += blockDim.x) box[t] = 0.0;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < IBM_MPT; q++) {
      if (e0[q] < 0) continue;
      const double fc = fnew[q] * c0;
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) unsafeAtomicAdd(&box[e0[q] + i + b.bx * j + bxy * k], fc * wx[q][i] * wy[q][j] * wz[q][k]);
    }
    __syncthreads();
  }
  // dudt = dudtf = fl(dudt + S) on this sphere's cells (src/main.f90:332-334)
  for (int t = threadIdx.x; t < nbox; t += blockDim.x) {
    if (!flagged(t)) continue;
    const long cc = cell_of(t);
    Fw[cc] = Fw[cc] + box[t];
  }
}

// a box [a0, a0+an) and a box [b0, b0+bn) of one periodic direction of n
// cells overlap once wrapped
__device__ __forceinline__ bool overlap_periodic(int a0, int an, int b0, int bn, int n) {
  for (int s = -2; s <= 2; s++) {
    const int c0 = b0 + s * n;
    if (a0 < c0 + bn && c0 < a0 + an) return true;
  }
  return false;
}

// iso[p] = 1 when sphere p's box overlaps no other sphere's box (periodic in
// x and y), does not wrap onto itself, and no w stencil of its markers can
// reach the k1 = kmax+1 plane (src/interp_spread.f90:100); its IBM passes
// then touch cells no other sphere touches.
__global__ __launch_bounds__(256) void k_ibm_isolate(IbmArgs A, int *iso) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.np) return;
  const Grid &g = A.g;
  const Box bp = sphere_box(A, A.px[p], A.py[p], A.pz[p]);
  bool ok = bp.bx < g.nx && bp.by < g.ny && (A.pz[p] + A.radfp) * g.dzi + 2.0 < (double)g.nz;
  for (int q = 0; q < A.np && ok; q++) {
    if (q == p) continue;
    const Box bq = sphere_box(A, A.px[q], A.py[q], A.pz[q]);
    if (bq.bk0 + bq.bz <= bp.bk0 || bp.bk0 + bp.bz <= bq.bk0) continue;
    if (!overlap_periodic(bp.bi0, bp.bx, bq.bi0, bq.bx, g.nx)) continue;
    if (!overlap_periodic(bp.bj0, bp.by, bq.bj0, bq.by, g.ny)) continue;
    ok = false;
  }
  iso[p] = ok ? 1 : 0;
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
  if (c->pc.nl > IBM_THREADS * IBM_MPT) { ipg_set_error("ibm: more than 1024 markers per sphere"); return 1; }
  IPG_HIP(hipFuncSetAttribute((const void *)k_ibm_pass, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  IPG_HIP(hipFuncSetAttribute((const void *)k_ibm_local, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
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
// The reference copies dudtf = dudt before the passes (src/main.f90:301-303);
// here the passes read dudt and the force buffers directly, so there is
// nothing to launch.  Kept as the IBM block's entry for the split interface.
int ipg_launch_ibm_begin(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  return 0;
}

// One IBM pass or half pass: gather (+ marker force update) and/or spread.
int ipg_launch_ibm_pass(ipg_ctx *c, int pass, int gather, int spread) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  if (pass == 2 && gather && spread) { ipg_set_error("ibm: the last pass gathers and spreads in two launches"); return 1; }
  ibm_roles(c, A, pass, gather, spread);
  hipLaunchKernelGGL(k_ibm_pass, dim3(3 * A.np), dim3(IBM_THREADS), lds, c->stream, A, pass, gather, spread,
                     (const int *)nullptr);
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

// The whole IBM block of one substep, src/main.f90:296-338.  Isolated
// spheres run their three passes in the first launch; the passes through the
// global force buffers then serve only the spheres near another sphere or the
// top wall.
int ipg_launch_ibm(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  hipLaunchKernelGGL(k_ibm_isolate, dim3((A.np + 255) / 256), dim3(256), 0, c->stream, A, c->iso);
  ipg_mark(c, "ibm_isolate");
  hipLaunchKernelGGL(k_ibm_local, dim3(3 * A.np), dim3(IBM_THREADS), lds, c->stream, A, (const int *)c->iso);
  ipg_mark(c, "ibm_local");
  const int roles[4][3] = {{0, 1, 1}, {1, 1, 1}, {2, 1, 0}, {2, 0, 1}};
  static const char *names[4] = {"ibm_pass0", "ibm_pass1", "ibm_pass2_gather", "ibm_pass2_spread"};
  for (int r = 0; r < 4; r++) {
    IbmArgs B = A;
    ibm_roles(c, B, roles[r][0], roles[r][1], roles[r][2]);
    hipLaunchKernelGGL(k_ibm_pass, dim3(3 * A.np), dim3(IBM_THREADS), lds, c->stream, B, roles[r][0], roles[r][1],
                       roles[r][2], (const int *)c->iso);
    ipg_mark(c, names[r]);
  }
  IPG_HIP(hipGetLastError());
  return ipg_launch_ibm_totals(c, 1);
}
