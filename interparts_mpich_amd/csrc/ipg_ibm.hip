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
// force buffers G (zero outside marker stencils), and the velocity the markers
// see is dudt + G: the same sums, fl(G + dudtold), without the full-field
// passes.  Two G buffers alternate so a pass reads one while spreading into
// the other; cells are cleared by the markers that touched them.  The
// w-stencil of a marker within h/2 of the top wall reaches the k1 = kmax+1
// plane (src/interp_spread.f90:100 has no upper clamp); that plane never
// receives dudt and keeps the last spread until the next one, so it is held
// in its own double-buffered plane K1.
#include "ipg_internal.h"

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
  double dVlagr, dVeul;
  const double *px, *py, *pz, *pu, *pv, *pw, *pomx, *pomy, *pomz, *pphi;
  const double *mk;  // [3][nl]
  double *fl;        // [3][np*nl]
  double *fld[3];    // dudt dvdt dwdt
  double *Gr[3], *Gw[3];
  double *K1r, *K1w;
  const double *dtp;
};

__device__ __forceinline__ long gidx(const Grid &g, int wi, int wj, int k) {
  return (long)(wi - 1) + (long)g.nx * ((long)(wj - 1) + (long)g.ny * (long)(k - 1));
}

// Gather (src/interp_spread.f90:132-154) + force update (src/forcing.f90:63-65,
// 255-257).  pass 2 also clears the spread buffer of pass 0; every pass
// clears the K1 plane it will spread into.
__global__ __launch_bounds__(256) void k_ibm_interp(IbmArgs A, int pass) {
  const Grid &g = A.g;
  const long nplane = (long)g.nx * g.ny;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nplane; t += (long)gridDim.x * blockDim.x) A.K1w[t] = 0.0;
  const long M = (long)A.np * A.nl;
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int p = (int)(m / A.nl), l = (int)(m - (long)p * A.nl);
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  double xl, yl, zl;
  if (!marker_pos(g, xp, yp, zp, A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl)) return;
  Stencil s;
  stencil(g, xl, yl, zl, s);
  double au = 0.0, av = 0.0, aw = 0.0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        const int wi = wrapi(s.ilows + i, g.nx), wj = wrapi(s.jlow + j, g.ny), kk = s.klow + k;
        const double f = A.fld[0][g.at(wi, wj, kk)] + A.Gr[0][gidx(g, wi, wj, kk)];
        au = au + f * s.phix[i] * s.phiyh[j] * s.phizh[k];
      }
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        const int wi = wrapi(s.ilow + i, g.nx), wj = wrapi(s.jlows + j, g.ny), kk = s.klow + k;
        const double f = A.fld[1][g.at(wi, wj, kk)] + A.Gr[1][gidx(g, wi, wj, kk)];
        av = av + f * s.phixh[i] * s.phiy[j] * s.phizh[k];
      }
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        const int wi = wrapi(s.ilow + i, g.nx), wj = wrapi(s.jlow + j, g.ny), kk = s.klows + k;
        double f;
        if (kk == g.nz + 1) f = A.K1r[(long)(wi - 1) + (long)g.nx * (wj - 1)];
        else if (kk > g.nz + 1) f = 0.0;  // dwdtf(.,.,k1+1): never holds dudt (marker outside the channel)
        else f = A.fld[2][g.at(wi, wj, kk)] + A.Gr[2][gidx(g, wi, wj, kk)];
        aw = aw + f * s.phixh[i] * s.phiyh[j] * s.phiz[k];
      }
  // rigid-body velocity at the marker, src/intgr_nwtn_eulr.f90:1471-1476
  const double ul = A.pu[p] + A.pomy[p] * (zl - zp) - A.pomz[p] * (yl - yp);
  const double vl = A.pv[p] + A.pomz[p] * (xl - xp) - A.pomx[p] * (zl - zp);
  const double wl = A.pw[p] + A.pomx[p] * (yl - yp) - A.pomy[p] * (xl - xp);
  const double dti = 1.0 / *A.dtp;
  double *fx = A.fl, *fy = A.fl + M, *fz = A.fl + 2 * M;
  if (pass == 0) {
    fx[m] = (ul - au) * dti; fy[m] = (vl - av) * dti; fz[m] = (wl - aw) * dti;
  } else {
    fx[m] = fx[m] + (ul - au) * dti; fy[m] = fy[m] + (vl - av) * dti; fz[m] = fz[m] + (wl - aw) * dti;
  }
  if (pass == 2) {
    for (int k = 0; k < 3; k++)
      for (int j = 0; j < 3; j++)
        for (int i = 0; i < 3; i++) {
          A.Gw[0][gidx(g, wrapi(s.ilows + i, g.nx), wrapi(s.jlow + j, g.ny), s.klow + k)] = 0.0;
          A.Gw[1][gidx(g, wrapi(s.ilow + i, g.nx), wrapi(s.jlows + j, g.ny), s.klow + k)] = 0.0;
          const int kk = s.klows + k;
          if (kk <= g.nz) A.Gw[2][gidx(g, wrapi(s.ilow + i, g.nx), wrapi(s.jlow + j, g.ny), kk)] = 0.0;
        }
  }
}

// Spread, src/interp_spread.f90:329-355 with the periodic fold of
// updthalos_ibm icase 2 (src/bound.f90:276-305) folded into the index.
__global__ __launch_bounds__(256) void k_ibm_spread(IbmArgs A) {
  const Grid &g = A.g;
  const long M = (long)A.np * A.nl;
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int p = (int)(m / A.nl), l = (int)(m - (long)p * A.nl);
  double xl, yl, zl;
  if (!marker_pos(g, A.px[p], A.py[p], A.pz[p], A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl)) return;
  Stencil s;
  stencil(g, xl, yl, zl, s);
  const double c0 = (A.dVlagr / A.dVeul) * *A.dtp;
  const double fxc = A.fl[m] * c0, fyc = A.fl[M + m] * c0, fzc = A.fl[2 * M + m] * c0;
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        unsafeAtomicAdd(&A.Gw[0][gidx(g, wrapi(s.ilows + i, g.nx), wrapi(s.jlow + j, g.ny), s.klow + k)],
                        fxc * s.phix[i] * s.phiyh[j] * s.phizh[k]);
        unsafeAtomicAdd(&A.Gw[1][gidx(g, wrapi(s.ilow + i, g.nx), wrapi(s.jlows + j, g.ny), s.klow + k)],
                        fyc * s.phixh[i] * s.phiy[j] * s.phizh[k]);
        const int wi = wrapi(s.ilow + i, g.nx), wj = wrapi(s.jlow + j, g.ny), kk = s.klows + k;
        const double cw = fzc * s.phixh[i] * s.phiyh[j] * s.phiz[k];
        if (kk == g.nz + 1) unsafeAtomicAdd(&A.K1w[(long)(wi - 1) + (long)g.nx * (wj - 1)], cw);
        else if (kk <= g.nz) unsafeAtomicAdd(&A.Gw[2][gidx(g, wi, wj, kk)], cw);
      }
}

// End of the IBM block (src/main.f90:332-334): dudt(1:n) = dudtf(1:n), i.e.
// dudt += G of the last pass on the touched cells.  The first marker to reach
// a cell takes its G value (atomic exchange with zero) and adds it; the
// spread buffer of the middle pass is cleared on the same cells.
__global__ __launch_bounds__(256) void k_ibm_apply(IbmArgs A) {
  const Grid &g = A.g;
  // the last spread's k1 plane becomes the plane the next substep reads
  const long nplane = (long)g.nx * g.ny;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nplane; t += (long)gridDim.x * blockDim.x) A.K1w[t] = A.K1r[t];
  const long M = (long)A.np * A.nl;
  const long m = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (m >= M) return;
  const int p = (int)(m / A.nl), l = (int)(m - (long)p * A.nl);
  double xl, yl, zl;
  if (!marker_pos(g, A.px[p], A.py[p], A.pz[p], A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl)) return;
  Stencil s;
  stencil(g, xl, yl, zl, s);
  for (int k = 0; k < 3; k++)
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) {
        for (int comp = 0; comp < 3; comp++) {
          int wi, wj, kk;
          if (comp == 0) { wi = wrapi(s.ilows + i, g.nx); wj = wrapi(s.jlow + j, g.ny); kk = s.klow + k; }
          else if (comp == 1) { wi = wrapi(s.ilow + i, g.nx); wj = wrapi(s.jlows + j, g.ny); kk = s.klow + k; }
          else { wi = wrapi(s.ilow + i, g.nx); wj = wrapi(s.jlow + j, g.ny); kk = s.klows + k; }
          if (kk > g.nz) continue;
          const long q = gidx(g, wi, wj, kk);
          A.Gw[comp][q] = 0.0;
          const double f = atomicExch(&A.Gr[comp][q], 0.0);
          if (f != 0.0) {
            double *fp = &A.fld[comp][g.at(wi, wj, kk)];
            *fp = *fp + f;
          }
        }
      }
}

// Per-sphere force and torque sums after the last pass, src/forcing.f90:258-287,
// one block per sphere.
__global__ __launch_bounds__(256) void k_ibm_totals(IbmArgs A, double *fxt, double *fyt, double *fzt, double *tqx, double *tqy,
                                                    double *tqz, double *tqt) {
  const Grid &g = A.g;
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

int ipg_launch_ibm(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  if (!c->markers_ready) { ipg_set_error("ibm: ipg_coordsfp has not been called"); return 1; }
  IbmArgs A{};
  A.g = c->g; A.np = c->pc.np; A.nl = c->pc.nl; A.dVlagr = c->pc.dVlagr; A.dVeul = c->pc.dVeul;
  double **q = c->ps.q;
  A.px = q[IPG_PX]; A.py = q[IPG_PY]; A.pz = q[IPG_PZ]; A.pu = q[IPG_PU]; A.pv = q[IPG_PV]; A.pw = q[IPG_PW];
  A.pomx = q[IPG_POMX]; A.pomy = q[IPG_POMY]; A.pomz = q[IPG_POMZ]; A.pphi = q[IPG_PPHI];
  A.mk = c->mk; A.fl = c->fl;
  A.fld[0] = c->f[IPG_DUDT]; A.fld[1] = c->f[IPG_DVDT]; A.fld[2] = c->f[IPG_DWDT];
  A.dtp = c->dscal;
  const long M = (long)A.np * A.nl;
  const int blocks = (int)((M + 255) / 256);
  // Fixed buffer roles every substep (so a captured graph replays correctly):
  // pass 0 reads G0 (all zero) and spreads into G1, pass 1 reads G1 into G0,
  // pass 2 reads G0 (clearing G1) into G1; the apply step consumes G1 and
  // clears G0.  K1[0] holds the k1 plane between substeps.
  c->gcur = 0;
  c->kcur = 0;
  for (int pass = 0; pass < 3; pass++) {
    for (int d = 0; d < 3; d++) { A.Gr[d] = c->G[c->gcur][d]; A.Gw[d] = c->G[1 - c->gcur][d]; }
    A.K1r = c->K1[c->kcur]; A.K1w = c->K1[1 - c->kcur];
    hipLaunchKernelGGL(k_ibm_interp, dim3(blocks), dim3(256), 0, c->stream, A, pass);
    ipg_mark(c, "ibm_interp");
    hipLaunchKernelGGL(k_ibm_spread, dim3(blocks), dim3(256), 0, c->stream, A);
    ipg_mark(c, "ibm_spread");
    c->gcur = 1 - c->gcur;
    c->kcur = 1 - c->kcur;
  }
  for (int d = 0; d < 3; d++) { A.Gr[d] = c->G[c->gcur][d]; A.Gw[d] = c->G[1 - c->gcur][d]; }
  A.K1r = c->K1[c->kcur]; A.K1w = c->K1[1 - c->kcur];
  hipLaunchKernelGGL(k_ibm_totals, dim3(A.np), dim3(256), 0, c->stream, A, q[IPG_PFXLTOT], q[IPG_PFYLTOT], q[IPG_PFZLTOT],
                     q[IPG_PTORQX], q[IPG_PTORQY], q[IPG_PTORQZ], q[IPG_PTORQTHETA]);
  ipg_mark(c, "ibm_totals");
  hipLaunchKernelGGL(k_ibm_apply, dim3(blocks), dim3(256), 0, c->stream, A);
  ipg_mark(c, "ibm_apply");
  c->gcur = 0;
  c->kcur = 0;
  IPG_HIP(hipGetLastError());
  return 0;
}
