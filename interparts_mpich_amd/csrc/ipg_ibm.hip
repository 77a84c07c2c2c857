// ipg_ibm.hip -- direct-forcing immersed boundary block of one substep
// (src/main.f90:286-338): three passes of marker interpolation
// (eulr2lagr, src/interp_spread.f90:12-228), marker force update
// (complagrforces / updtlagrforces, src/forcing.f90:9-391) and spreading
// (lagr2eulr, src/interp_spread.f90:230-474).
//
// The reference copies dudt into a 2-ghost scratch field dudtf, zeroes it
// before every spread, scatters the marker forces with atomics, folds the
// periodic ghosts and adds the saved dudt back: 9 full-field passes per IBM
// pass.  Here nothing full-field is touched.  One workgroup works on one
// (sphere, velocity component): the cells its markers' 27-point stencils touch
// lie in a box around the sphere, and that box is held in LDS.  The spread S
// of a pass at those cells is the sum over the sphere's own markers and the
// markers of the spheres whose boxes overlap its box (a neighbour list built
// once per substep), so each workgroup rebuilds S in LDS from the marker
// forces of the previous pass and gathers from fl(dudt + S), the dudtf of the
// reference.  After the last pass each cell's owner (the lowest-numbered
// sphere whose stencils touch it) adds S to dudt, once.  No force field is
// kept, and no global atomics are needed except on the top-wall plane: the
// w-stencil of a marker within h/2 of the top wall reaches the k1 = kmax+1
// plane (src/interp_spread.f90:100 has no upper clamp), which never receives
// dudt and keeps the last spread until the next one (planes K).
#include "ipg_internal.h"
#include "ipg_ibm_geom.h"
#include <cmath>

namespace {


struct IbmArgs {
  Grid g;
  int np, nl;
  double dVlagr, dVeul, radfp;
  int bmx, bmy, bmz;  // largest box extent per direction (LDS sizing)
  const double *px, *py, *pz, *pu, *pv, *pw, *pomx, *pomy, *pomz, *pphi;
  const double *mk;   // [3][nl] marker offsets from the sphere centre
  void *fl;           // [3][np*nl] marker forces fxl, fyl, fzl written by this launch (T, the launch's type)
  const void *flp;    // [3][np*nl] marker forces of the previous pass (read; another buffer)
  double *fld[3];     // dudt dvdt dwdt
  const int *nbl;     // [np][IBM_MAXNB] neighbour images: q*25 + (sx+2)*5 + (sy+2)
  const int *nbc;     // [np] neighbour image count
  int *nml;           // [np*3][IBM_MAXNM] neighbour markers touching the box: (q*nl + l)*25 + shift
  int *nmc;           // [np*3] their count
  int *err;           // device error word: bits 1, 2: a neighbour list overflowed; 4: a stencil outside its box
  unsigned *flg;      // [np*3][nflagmax] the box cells the own stencils touch, set by pass 0
  double *Kread, *Kspread, *Kzero;  // dwdtf(.,.,k1) planes
  double *psum;       // [3 passes][3 np][IBM_WAVES]: per-wavefront sums of the pass's new marker forces
  const double *dtp;
};

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

// Neighbour images of each sphere, once per substep: every (q, sx, sy) with
// q's box, shifted by (sx nx, sy ny), overlapping p's box; q != p (a box
// never meets its own image: it is narrower than the domain).  One block per
// sphere, one thread per partner; the order of the entries is immaterial, as
// the spreads they feed are sums.
constexpr int IBM_NB_THREADS = 256;
__device__ __forceinline__ void ibm_neighbours_body(const IbmArgs &A, int *nbl, int *nbc, int *err, const VB vb) {
  const Grid &g = A.g;
  const int p = vb.x;
  __shared__ int count;
  if (threadIdx.x == 0) count = 0;
  __syncthreads();
  const Box bp = reach_box(A.g, A.radfp, A.px[p], A.py[p], A.pz[p]);
  for (int q = threadIdx.x; q < A.np; q += IBM_NB_THREADS) {
    if (q == p) continue;
    const Box bq = reach_box(A.g, A.radfp, A.px[q], A.py[q], A.pz[q]);
    unsigned mx, my;
    if (!image_masks(g, bp, bq, mx, my)) continue;
    const int n = __popc(mx) * __popc(my);
    int slot = atomicAdd(&count, n);
    for (int sx = 0; sx < 5; sx++)
      for (int sy = 0; sy < 5; sy++)
        if (((mx >> sx) & 1u) && ((my >> sy) & 1u)) {
          if (slot < IBM_MAXNB) nbl[(long)p * IBM_MAXNB + slot] = q * 25 + sx * 5 + sy;
          slot++;
        }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    nbc[p] = min(count, IBM_MAXNB);
    if (count > IBM_MAXNB) atomicOr(err, 1);
  }
}

struct IbmNbArgs { IbmArgs A; int *nbl, *nbc, *err; };
struct IbmNbK {
  using Args = IbmNbArgs;
  static constexpr int THREADS = IBM_NB_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) { ibm_neighbours_body(a.A, a.nbl, a.nbc, a.err, v); }
};
__global__ __launch_bounds__(IBM_NB_THREADS) void k_ibm_neighbours(IbmNbArgs) { run_single<IbmNbK>(); }

// Stencil of marker l of sphere q, seen through the periodic shift (sxs nx,
// sys ny) in the unwrapped frame of box b; false when it misses the box.
template <bool WEIGHTS = true, typename T = double>
__device__ __forceinline__ bool neighbour_stencil(const IbmArgs &A, const Box &b, int comp, int qs, int l, int sxs, int sys,
                                                  CompStencilT<T> &c) {
  return image_stencil<WEIGHTS, T>(A.g, b, comp, A.px[qs], A.py[qs], A.pz[qs], A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l],
                                   sxs, sys, c);
}

// One IBM launch for one (sphere, velocity component) workgroup.
//   MODE 0 (pass 0): gather from dudt, set the marker forces; list the
//     neighbours' markers whose stencils touch this box (for the next launches).
//   MODE 1 (passes 1, 2): sum dudt and the previous pass's spread S in LDS
//     (dudtf = dudt + S), gather from it, add to the marker forces.
//   MODE 2 (after pass 2): sum dudt and the last spread S in LDS, and store
//     dudt + S on the cells this sphere owns.
// Thread t owns markers t, t + IBM_THREADS; their positions
// (src/coordsfp.f90:49-68), stencils and weights (src/interp_spread.f90:76-130)
// and rigid-body velocities (src/intgr_nwtn_eulr.f90:1471-1476) are computed in
// registers, as the markers do not move inside the IBM block.  Gathers follow
// src/interp_spread.f90:132-154 (((f phi_x) phi_y) phi_z); the force update
// src/forcing.f90:63-65 / 255-257; spreads src/interp_spread.f90:329-355
// ((((F c0) phi_x) phi_y) phi_z, c0 = dVlagr / dVeul dt).
//
// T is the arithmetic type of the weights, the gathers, the spreads and the
// LDS box: double (FP64 path) or float (mixed path; positions, stencil base
// cells and the offsets the weights take are still formed in FP64).  With
// float, MODE 2 sums only the spread S in the box and adds it to the FP64
// dudt as it stores, so that dudt is never rounded to FP32.
template <int MODE, typename T>
__device__ __forceinline__ void ibm_body(const IbmArgs &A, int pass, const VB vb) {
  constexpr bool F64 = sizeof(T) == sizeof(double);
  extern __shared__ __align__(16) unsigned char ibm_lds[];
  T *box = reinterpret_cast<T *>(ibm_lds);
  const Grid &g = A.g;
  const int p = vb.x / 3, comp = vb.x % 3;
  const long M = (long)A.np * A.nl;
  if (A.Kzero) {
    const long nplane = (long)g.nx * g.ny;
    for (long t = vb.x * (long)IBM_THREADS + threadIdx.x; t < nplane; t += (long)vb.gx * IBM_THREADS) A.Kzero[t] = 0.0;
  }
  const double dt = *A.dtp;
  const T dti = (T)(1.0 / dt);
  const T c0 = (T)((A.dVlagr / A.dVeul) * dt);
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  Box b = comp_box(A.g, A.radfp, xp, yp, zp, comp);
  if (b.bx > A.bmx || b.by > A.bmy || b.bz > A.bmz) {
    // cannot happen for markers within radfp of the centre; kept as a check
    if (MODE == 0 && threadIdx.x == 0) atomicOr(A.err, 4);
    b.bx = min(b.bx, A.bmx); b.by = min(b.by, A.bmy); b.bz = min(b.bz, A.bmz);
  }
  const int nbox = b.bx * b.by * b.bz, bxy = b.bx * b.by;
  const IbmLds L = ibm_lds_layout(A.bmx, A.bmy, A.bmz, sizeof(T));
  const int nflag = (nbox + 31) >> 5, nflagmax = L.nflagmax;
  unsigned *flag = reinterpret_cast<unsigned *>(ibm_lds + L.flag);
  unsigned *taken = reinterpret_cast<unsigned *>(ibm_lds + L.taken);  // MODE 2: cells a lower-numbered sphere owns
  // field offset of a box cell = rowoff[row] + icol[li] (box_rowoff, box_icol)
  int *rowoff = reinterpret_cast<int *>(ibm_lds + L.rowoff);
  int *icol = reinterpret_cast<int *>(ibm_lds + L.icol);
  double *F = A.fld[comp];
  T *fl = (T *)A.fl + comp * M;
  // the previous pass's forces: a buffer no workgroup of this launch writes,
  // as neighbours read them while their owners compute the new ones
  const T *flp = (const T *)A.flp + comp * M;
  const long m0 = (long)p * A.nl;
  // Own marker q of this thread: stencil, LDS index e0 of the stencil's first
  // cell, how many of its z planes lie in the box (the rest is the k1 plane),
  // and the rigid-body velocity ub.  Recomputed in each phase that needs it
  // rather than kept live across the launch.  The box holds the stencil in x
  // and y, and in z up to kmax (comp_box); a stencil that were not held is
  // reported (error bit 4) and left out rather than index outside the box.
  // The own markers' offsets and previous-pass forces, loaded first so that
  // their latency overlaps the set-up below rather than following the
  // staging (indexed by the unrolled q only: registers).
  double mkx[IBM_MPT], mky[IBM_MPT], mkz[IBM_MPT];
  T fprev[IBM_MPT];
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    const int l = (int)threadIdx.x + q * IBM_THREADS;
    mkx[q] = mky[q] = mkz[q] = 0.0;
    fprev[q] = (T)0;
    if (l < A.nl) {
      mkx[q] = A.mk[l]; mky[q] = A.mk[A.nl + l]; mkz[q] = A.mk[2 * A.nl + l];
      if (MODE >= 1) fprev[q] = flp[m0 + l];
    }
  }
  // likewise the count of listed neighbour markers and this thread's first
  // entry (the list holds IBM_MAXNM >= IBM_THREADS entries, so the load is in
  // bounds whatever the count)
  static_assert(IBM_MAXNM >= IBM_THREADS, "first neighbour-list entry per thread");
  int nbcnt = 0, code0 = 0;
  if (MODE >= 1) {
    nbcnt = A.nmc[vb.x];
    code0 = A.nml[(long)vb.x * IBM_MAXNM + threadIdx.x];
  } else {
    // pass 0: the sphere's neighbour images, and the image of this thread's
    // first candidate marker (entry threadIdx.x / nl of the list)
    nbcnt = A.nbc[p];
    code0 = A.nbl[(long)p * IBM_MAXNB + min((int)threadIdx.x / A.nl, IBM_MAXNB - 1)];
  }
  auto own = [&](int q, CompStencilT<T> &c, int &e0, int &nzin, T &ub) -> bool {
    const int l = (int)threadIdx.x + q * IBM_THREADS;
    double xl, yl, zl;
    if (l >= A.nl || !marker_pos(g, xp, yp, zp, mkx[q], mky[q], mkz[q], xl, yl, zl)) return false;
    comp_stencil<true, T>(g, xl, yl, zl, comp, c);
    if (!own_cells(b, c.i0, c.j0, c.k0, e0, nzin)) {
      if (MODE == 0) atomicOr(A.err, 4);
      return false;
    }
    double u;
    if (comp == 0) u = A.pu[p] + A.pomy[p] * (zl - zp) - A.pomz[p] * (yl - yp);
    else if (comp == 1) u = A.pv[p] + A.pomz[p] * (xl - xp) - A.pomx[p] * (zl - zp);
    else u = A.pw[p] + A.pomx[p] * (yl - yp) - A.pomy[p] * (xl - xp);
    ub = (T)u;
    return true;
  };
  // the cells the own stencils touch: flagged by pass 0, read back by the
  // later launches (the markers do not move inside the IBM block)
  unsigned *gflag = A.flg + (long)vb.x * nflagmax;
  for (int t = threadIdx.x; t < nflag; t += IBM_THREADS) {
    flag[t] = MODE == 0 ? 0u : gflag[t];
    if (MODE == 2) taken[t] = 0u;
  }
  for (int t = threadIdx.x; t < b.by * b.bz; t += IBM_THREADS) rowoff[t] = box_rowoff(g, b, t);
  for (int t = threadIdx.x; t < b.bx; t += IBM_THREADS) icol[t] = box_icol(g, b, t);
  if (MODE >= 1)
    for (int t = threadIdx.x; t < nbox; t += IBM_THREADS) box[t] = (T)0;
  __shared__ int lcount;
  if (MODE == 0 && threadIdx.x == 0) lcount = 0;
  __syncthreads();
  auto flagged = [&](int t) { return (flag[t >> 5] >> (t & 31)) & 1u; };
  // the operands of this thread's first listed neighbour marker (its
  // sphere's centre, its offsets, its previous force), loaded before the
  // staging so that their latency overlaps it
  double nbq[3] = {0.0, 0.0, 0.0}, nbe[3] = {0.0, 0.0, 0.0};
  T nbf = (T)0;
  if (MODE >= 1 && (int)threadIdx.x < nbcnt) {
    const int mq = code0 / 25, qs = mq / A.nl, l = mq - qs * A.nl;
    nbq[0] = A.px[qs]; nbq[1] = A.py[qs]; nbq[2] = A.pz[qs];
    nbe[0] = A.mk[l]; nbe[1] = A.mk[A.nl + l]; nbe[2] = A.mk[2 * A.nl + l];
    nbf = flp[mq];
  }
  // the box cells of this thread: t = threadIdx.x + n IBM_THREADS
  const BoxWalk w0(b.bx, IBM_THREADS, (int)threadIdx.x);
  constexpr int CH = IBM_CH;
  if (MODE == 1 || (MODE == 2 && F64)) {
    // dudt on the cells the own stencils touch (flagged by pass 0), added
    // into the box together with the spreads below, so that its loads overlap
    // them: the box ends as dudt + S, the reference's dudtf, with the terms
    // summed in another order (as the LDS atomics already order S's terms)
    BoxWalk w = w0;
    for (int t0 = threadIdx.x; t0 < nbox; t0 += CH * IBM_THREADS) {
      double v[CH];
      unsigned fm = 0u;
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * IBM_THREADS;
        v[u] = 0.0;
        if (t < nbox && flagged(t)) {
          v[u] = F[(long)rowoff[w.row] + icol[w.li]];
          fm |= 1u << u;
        }
        w.advance();
      }
#pragma unroll
      for (int u = 0; u < CH; u++)
        if ((fm >> u) & 1u) unsafeAtomicAdd(&box[t0 + u * IBM_THREADS], (T)v[u]);
    }
  }
  const long lbase = (long)vb.x * IBM_MAXNM;
  if (MODE == 0) {
    // list the neighbours' markers whose stencils touch this box; the
    // markers do not move inside the IBM block, so the three launches
    // that spread them reuse the list.  A thread's first candidate takes
    // the image loaded first, and, from the first image, the offsets of
    // its own marker (the same marker number).
    const int tot = nbcnt * A.nl;
    for (int idx = threadIdx.x; idx < tot; idx += IBM_THREADS) {
      const int n = idx / A.nl, l = idx - n * A.nl;
      const bool first = idx == (int)threadIdx.x;
      const int code = first ? code0 : A.nbl[(long)p * IBM_MAXNB + n];
      const int qs = code / 25, sh = code % 25;
      double ex, ey, ez;
      if (first && n == 0) {
        ex = mkx[0]; ey = mky[0]; ez = mkz[0];
      } else {
        ex = A.mk[l]; ey = A.mk[A.nl + l]; ez = A.mk[2 * A.nl + l];
      }
      CompStencil c;
      const bool keep = image_stencil<false, double>(g, b, comp, A.px[qs], A.py[qs], A.pz[qs], ex, ey, ez, sh / 5 - 2,
                                                     sh % 5 - 2, c);
      // one LDS atomic per wavefront
      const unsigned long long msk = __ballot(keep);
      if (!keep) continue;
      const int lead = __ffsll((long long)msk) - 1, lane = (int)(threadIdx.x & 63);
      int base = 0;
      if (lane == lead) base = atomicAdd(&lcount, (int)__popcll(msk));
      base = __shfl(base, lead, 64);
      const int slot = base + (int)__popcll(msk & ((1ull << lane) - 1ull));
      if (slot < IBM_MAXNM) A.nml[lbase + slot] = (qs * A.nl + l) * 25 + sh;
    }
  } else {
    // the listed neighbour markers' contributions to this box
    for (int idx = threadIdx.x; idx < nbcnt; idx += IBM_THREADS) {
      const bool first = idx == (int)threadIdx.x;
      const int code = first ? code0 : A.nml[lbase + idx];
      const int mq = code / 25, sh = code % 25;
      const int qs = mq / A.nl, l = mq - qs * A.nl;
      double xq, yq, zq, ex, ey, ez;
      T f;
      if (first) {
        xq = nbq[0]; yq = nbq[1]; zq = nbq[2]; ex = nbe[0]; ey = nbe[1]; ez = nbe[2]; f = nbf;
      } else {
        xq = A.px[qs]; yq = A.py[qs]; zq = A.pz[qs];
        ex = A.mk[l]; ey = A.mk[A.nl + l]; ez = A.mk[2 * A.nl + l];
        f = flp[mq];
      }
      CompStencilT<T> c;
      if (!image_stencil<true, T>(g, b, comp, xq, yq, zq, ex, ey, ez, sh / 5 - 2, sh % 5 - 2, c)) continue;
      const T fc = f * c0;
#pragma unroll
      for (int k = 0; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++) {
            const int e = box_at(b, c.i0 + i, c.j0 + j, c.k0 + k);
            if (e < 0) continue;
            unsafeAtomicAdd(&box[e], fc * c.wx[i] * c.wy[j] * c.wz[k]);
            if (MODE == 2 && qs < p) atomicOr(&taken[e >> 5], 1u << (e & 31));
          }
    }
  }
  // flag the own stencil cells; with MODE >= 1 also spread the own forces of
  // the previous pass into S (after the neighbours' part, so that the first
  // marker's stencil, kept in registers for the gather below, is not live
  // across the neighbour loop).
  CompStencilT<T> cs0{};
  int e00 = 0, nz00 = 0;
  T ub00 = (T)0;
  bool have0 = false;
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    CompStencilT<T> c;
    int e0, nzin;
    T ub;
    if (!own(q, c, e0, nzin, ub)) continue;
    if (q == 0 && MODE <= 1) { cs0 = c; e00 = e0; nz00 = nzin; ub00 = ub; have0 = true; }
    const T fc = MODE >= 1 ? fprev[q] * c0 : (T)0;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++) {
          if (k >= nzin) continue;
          const int e = e0 + i + b.bx * j + bxy * k;
          if (MODE == 0) atomicOr(&flag[e >> 5], 1u << (e & 31));
          else unsafeAtomicAdd(&box[e], fc * c.wx[i] * c.wy[j] * c.wz[k]);
        }
  }
  __syncthreads();
  if (MODE == 0)
    for (int t = threadIdx.x; t < nflag; t += IBM_THREADS) gflag[t] = flag[t];
  if (MODE == 2) {
    // dudt = dudtf = fl(dudt + S) (src/main.f90:332-334), once per cell
    BoxWalk w = w0;
    for (int t = threadIdx.x; t < nbox; t += IBM_THREADS) {
      if (flagged(t) && !((taken[t >> 5] >> (t & 31)) & 1u)) {
        const long x = (long)rowoff[w.row] + icol[w.li];
        if (F64) F[x] = (double)box[t];
        else F[x] = F[x] + (double)box[t];
      }
      w.advance();
    }
    return;
  }
  if (MODE == 0) {
    // stage dudt on the flagged cells, in chunks of IBM_CH cells per thread
    // so that the loads of a chunk are all in flight together
    BoxWalk w = w0;
    for (int t0 = threadIdx.x; t0 < nbox; t0 += CH * IBM_THREADS) {
      double v[CH];
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * IBM_THREADS;
        v[u] = 0.0;
        if (t < nbox && flagged(t)) v[u] = F[(long)rowoff[w.row] + icol[w.li]];
        w.advance();
      }
#pragma unroll
      for (int u = 0; u < CH; u++) {
        const int t = t0 + u * IBM_THREADS;
        if (t < nbox) box[t] = (T)v[u];
      }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      A.nmc[vb.x] = min(lcount, IBM_MAXNM);
      if (lcount > IBM_MAXNM) atomicOr(A.err, 2);
    }
  }
  double fsum = 0.0;
#pragma unroll
  for (int q = 0; q < IBM_MPT; q++) {
    CompStencilT<T> c;
    int e0, nzin;
    T ub;
    if (q == 0 && MODE <= 1) {
      if (!have0) continue;
      c = cs0; e0 = e00; nzin = nz00; ub = ub00;
    } else if (!own(q, c, e0, nzin, ub)) {
      continue;
    }
    const long m = m0 + (int)threadIdx.x + q * IBM_THREADS;
    T acc = (T)0;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int i = 0; i < 3; i++)
          if (k < nzin) acc = acc + box[e0 + i + b.bx * j + bxy * k] * c.wx[i] * c.wy[j] * c.wz[k];
    // the stencil planes above kmax: the k1 plane (the same k order, so the
    // same sum), and this pass's spread into it
    if (nzin < 3) {
#pragma unroll
      for (int k = 1; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++)
            if (k >= nzin) acc = acc + (T)A.Kread[kplane(g, c.i0 + i, c.j0 + j)] * c.wx[i] * c.wy[j] * c.wz[k];
    }
    const T fnew = MODE == 0 ? (ub - acc) * dti : fprev[q] + (ub - acc) * dti;
    fl[m] = fnew;
    fsum += (double)fnew;
    if (nzin < 3) {
      const T fc = fnew * c0;
#pragma unroll
      for (int k = 1; k < 3; k++)
#pragma unroll
        for (int j = 0; j < 3; j++)
#pragma unroll
          for (int i = 0; i < 3; i++)
            if (k >= nzin)
              unsafeAtomicAdd(&A.Kspread[kplane(g, c.i0 + i, c.j0 + j)], (double)(fc * c.wx[i] * c.wy[j] * c.wz[k]));
    }
  }
  // this pass's force total of the sphere's markers (src/forcing.f90:66 /
  // 258, before the factor dVlagr), one partial sum per wavefront, for the
  // per-pass output of updtintermediatevel (k_ibm_force_sums)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) fsum += __shfl_xor(fsum, o, 64);
  if ((threadIdx.x & 63) == 0) A.psum[((long)pass * vb.gx + vb.x) * IBM_WAVES + (threadIdx.x >> 6)] = fsum;
}

struct IbmPassArgs { IbmArgs A; int pass; };
template <int MODE, typename T>
struct IbmK {
  using Args = IbmPassArgs;
  static constexpr int THREADS = IBM_THREADS, WPE = 4;
  static __device__ __forceinline__ void run(const Args &a, const VB v) { ibm_body<MODE, T>(a.A, a.pass, v); }
};
template <int MODE, typename T>
__global__ __launch_bounds__(IBM_THREADS) __attribute__((amdgpu_waves_per_eu(4))) void k_ibm(IbmPassArgs) {
  run_single<IbmK<MODE, T>>();
}

// src/forcing.f90:440-458 (updtintermediatevel): forcextot, forceytot,
// forceztot, the sums over the spheres of fxltot.. (the marker-force sums
// times dVlagr, src/forcing.f90:66-77 / 258-275), for each of the three passes
// of the last IBM block: out[3 pass + comp].  One block per (pass,
// component); the order of the sums is fixed, so the result is reproducible.
__global__ __launch_bounds__(256) void k_ibm_force_sums(const double *__restrict__ psum, int np, double dVlagr,
                                                        double *__restrict__ out) {
  const int pass = blockIdx.x / 3, comp = blockIdx.x % 3;
  double s = 0.0;
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    const double *w = psum + ((long)pass * 3 * np + 3 * p + comp) * IBM_WAVES;
    double t = 0.0;
    for (int q = 0; q < IBM_WAVES; q++) t += w[q];
    s += t * dVlagr;
  }
  __shared__ double red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// Per-sphere force and torque sums after the last pass, src/forcing.f90:258-287,
// one block per sphere.
constexpr int IBM_TOT_THREADS = 256;
template <typename T>
__device__ __forceinline__ void ibm_totals_body(const IbmArgs &A, double *fxt, double *fyt, double *fzt, double *tqx,
                                                double *tqy, double *tqz, double *tqt, const VB vb) {
  const Grid &g = A.g;
  const int p = vb.x;
  const long M = (long)A.np * A.nl;
  const double xp = A.px[p], yp = A.py[p], zp = A.pz[p];
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int l = threadIdx.x; l < A.nl; l += IBM_TOT_THREADS) {
    double xl, yl, zl;
    if (!marker_pos(g, xp, yp, zp, A.mk[l], A.mk[A.nl + l], A.mk[2 * A.nl + l], xl, yl, zl)) continue;
    const long m = (long)p * A.nl + l;
    const T *f = (const T *)A.fl;
    const double fx = f[m], fy = f[M + m], fz = f[2 * M + m];
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
    for (int q = 0; q < 6; q++) { r[q] = 0.0; for (int w = 0; w < IBM_TOT_THREADS / 64; w++) r[q] += sh[q][w]; }
    fxt[p] = r[0] * A.dVlagr; fyt[p] = r[1] * A.dVlagr; fzt[p] = r[2] * A.dVlagr;
    tqx[p] = r[3] * A.dVlagr; tqy[p] = r[4] * A.dVlagr; tqz[p] = r[5] * A.dVlagr;
    tqt[p] = (tqy[p] * cos(A.pphi[p])) - (tqx[p] * sin(A.pphi[p]));
  }
}

struct IbmTotArgs { IbmArgs A; double *fxt, *fyt, *fzt, *tqx, *tqy, *tqz, *tqt; };
template <typename T>
struct IbmTotK {
  using Args = IbmTotArgs;
  static constexpr int THREADS = IBM_TOT_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    ibm_totals_body<T>(a.A, a.fxt, a.fyt, a.fzt, a.tqx, a.tqy, a.tqz, a.tqt, v);
  }
};
template <typename T>
__global__ __launch_bounds__(IBM_TOT_THREADS) void k_ibm_totals(IbmTotArgs) { run_single<IbmTotK<T>>(); }

// dynamic LDS of one k_ibm workgroup (ibm_lds_layout)
size_t ibm_lds_bytes(int bmx, int bmy, int bmz, size_t esize) { return ibm_lds_layout(bmx, bmy, bmz, esize).bytes; }


}  // namespace

// LDS of the per-sphere box (one velocity component at a time) and its two
// cell masks, set once at marker set-up so that graph capture never meets an
// attribute change.
int ipg_ibm_setup(ipg_ctx *c) {
  const int bmx = ibm_box_max(c->pc.radfp, c->g.dxi), bmy = ibm_box_max(c->pc.radfp, c->g.dyi);
  const int bmz = ibm_box_max(c->pc.radfp, c->g.dzi);
  const size_t lds = ibm_lds_bytes(bmx, bmy, bmz, c->ibm32 ? sizeof(float) : sizeof(double));
  // with the kernels' own static LDS (pass 0's list counter)
  hipFuncAttributes fa{};
  IPG_HIP(hipFuncGetAttributes(&fa, c->ibm32 ? (const void *)k_ibm<0, float> : (const void *)k_ibm<0, double>));
  if (lds + fa.sharedSizeBytes > 160 * 1024) { ipg_set_error("ibm: sphere box exceeds the 160 KiB LDS of a CU"); return 1; }
  // the box's row offsets are kept as 32-bit integers
  if ((double)c->g.sxy * (c->g.nz + 2) >= 2147483647.0) { ipg_set_error("ibm: field larger than 2^31 cells"); return 1; }
  // the reach box (one cell wider than a component's box) must not meet its
  // own periodic image
  if (bmx + 1 > c->g.nx || bmy + 1 > c->g.ny) { ipg_set_error("ibm: sphere box wider than the domain"); return 1; }
  if (c->pc.nl > IBM_THREADS * IBM_MPT) { ipg_set_error("ibm: more than 1024 markers per sphere"); return 1; }
  if (c->ibm32) {
    IPG_HIP(hipFuncSetAttribute((const void *)k_ibm<0, float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    IPG_HIP(hipFuncSetAttribute((const void *)k_ibm<1, float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    IPG_HIP(hipFuncSetAttribute((const void *)k_ibm<2, float>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  } else {
    IPG_HIP(hipFuncSetAttribute((const void *)k_ibm<0, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    IPG_HIP(hipFuncSetAttribute((const void *)k_ibm<1, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    IPG_HIP(hipFuncSetAttribute((const void *)k_ibm<2, double>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  }
  // the captured substeps hold the IBM launch arguments (marker radius, box
  // sizes, dVlagr) by value: a new marker set-up drops them (and a batch's
  // merged launches, which hold them too)
  c->gen++;
  for (int r = 0; r < IPG_NGRAPH; r++)
    if (c->graph_ok[r]) { (void)hipGraphExecDestroy(c->graph[r]); c->graph_ok[r] = false; }
  // the pass-0 cell flags; a larger box (a second set-up with a larger
  // marker radius) takes a new buffer
  const size_t nflag = 3 * (size_t)c->pc.np * (((size_t)bmx * bmy * bmz + 31) / 32);
  if (nflag > c->ibm_flg_n) {
    if (c->ibm_flg) (void)hipFree(c->ibm_flg);
    c->ibm_flg = nullptr;
    c->ibm_flg_n = 0;
    if (hipMalloc(&c->ibm_flg, nflag * sizeof(unsigned)) != hipSuccess) { ipg_set_error("ibm: out of device memory"); return 1; }
    c->ibm_flg_n = nflag;
  }
  c->ibm_lds_set = true;
  return 0;
}

namespace {

// Common arguments of the IBM launches for this context.
bool ibm_args(ipg_ctx *c, IbmArgs &A, size_t &lds) {
  if (!c->markers_ready || !c->ibm_lds_set) { ipg_set_error("ibm: ipg_coordsfp has not been called"); return false; }
  A = IbmArgs{};
  A.g = c->g; A.np = c->pc.np; A.nl = c->pc.nl; A.dVlagr = c->pc.dVlagr; A.dVeul = c->pc.dVeul; A.radfp = c->pc.radfp;
  A.bmx = ibm_box_max(c->pc.radfp, c->g.dxi);
  A.bmy = ibm_box_max(c->pc.radfp, c->g.dyi);
  A.bmz = ibm_box_max(c->pc.radfp, c->g.dzi);
  lds = ibm_lds_bytes(A.bmx, A.bmy, A.bmz, c->ibm32 ? sizeof(float) : sizeof(double));
  double **q = c->ps.q;
  A.px = q[IPG_PX]; A.py = q[IPG_PY]; A.pz = q[IPG_PZ]; A.pu = q[IPG_PU]; A.pv = q[IPG_PV]; A.pw = q[IPG_PW];
  A.pomx = q[IPG_POMX]; A.pomy = q[IPG_POMY]; A.pomz = q[IPG_POMZ]; A.pphi = q[IPG_PPHI];
  A.mk = c->mk; A.fl = c->fl;
  A.fld[0] = c->f[IPG_DUDT]; A.fld[1] = c->f[IPG_DVDT]; A.fld[2] = c->f[IPG_DWDT];
  A.nbl = c->nbl; A.nbc = c->nbc; A.nml = c->nml; A.nmc = c->nmc; A.err = c->ibm_err; A.flg = c->ibm_flg;
  A.psum = c->ibm_psum;
  A.dtp = c->dscal;
  return true;
}

}  // namespace

// Neighbour lists of this substep (the markers do not move inside the IBM
// block).  This is where the reference copies dudtf = dudt
// (src/main.f90:301-303), which is not needed here.
int ipg_launch_ibm_begin(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  if (ipg_go<IbmNbK>(c, k_ibm_neighbours, dim3(A.np), dim3(IBM_NB_THREADS), 0, IbmNbArgs{A, c->nbl, c->nbc, c->ibm_err},
                     "ibm_neighbours"))
    return 1;
  IPG_HIP(hipGetLastError());
  return 0;
}

// Pass `pass` (0..2): eulr2lagr and the marker force update.  The k1 planes
// (fixed roles, so a captured graph replays them): K0 carries the last
// spread of the previous substep into pass 0, pass 0 spreads into K1, pass 1
// into K2, pass 2 into K0; each plane is cleared once nothing reads it.
int ipg_launch_ibm_gather(ipg_ctx *c, int pass) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  double **K = c->K1;
  // forces: pass 0 -> fl, pass 1 fl -> fl2, pass 2 fl2 -> fl
  A.fl = pass == 1 ? c->fl2 : c->fl;
  A.flp = pass == 2 ? c->fl2 : c->fl;
  A.Kread = K[pass];
  A.Kspread = K[(pass + 1) % 3];
  A.Kzero = pass == 0 ? nullptr : K[pass - 1];
  static const char *names[3] = {"ibm_pass0", "ibm_pass1", "ibm_pass2"};
  const dim3 grid(3 * A.np), blk(IBM_THREADS);
  const IbmPassArgs a{A, pass};
  int rc;
  if (c->ibm32) {
    if (pass == 0) rc = ipg_go<IbmK<0, float>>(c, k_ibm<0, float>, grid, blk, lds, a, names[pass]);
    else rc = ipg_go<IbmK<1, float>>(c, k_ibm<1, float>, grid, blk, lds, a, names[pass]);
  } else {
    if (pass == 0) rc = ipg_go<IbmK<0, double>>(c, k_ibm<0, double>, grid, blk, lds, a, names[pass]);
    else rc = ipg_go<IbmK<1, double>>(c, k_ibm<1, double>, grid, blk, lds, a, names[pass]);
  }
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

// lagr2eulr of the last pass: dudt += S, once per cell.  The spreads of
// passes 0 and 1 need no launch of their own: the next gather rebuilds them.
int ipg_launch_ibm_final(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  A.Kzero = c->K1[2];
  A.flp = c->fl;
  const IbmPassArgs a{A, 2};
  const int rc = c->ibm32 ? ipg_go<IbmK<2, float>>(c, k_ibm<2, float>, dim3(3 * A.np), dim3(IBM_THREADS), lds, a, "ibm_spread")
                          : ipg_go<IbmK<2, double>>(c, k_ibm<2, double>, dim3(3 * A.np), dim3(IBM_THREADS), lds, a, "ibm_spread");
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

// The per-pass force totals of the last IBM block into out9 (device memory).
int ipg_launch_ibm_forces(ipg_ctx *c, double *out9) {
  if (c->pc.np <= 0) return 0;
  hipLaunchKernelGGL(k_ibm_force_sums, dim3(9), dim3(256), 0, c->stream, c->ibm_psum, c->pc.np, c->pc.dVlagr, out9);
  IPG_HIP(hipGetLastError());
  return 0;
}

// Per-sphere force and torque sums (src/forcing.f90:66-89, 258-287) of the
// forces after pass `pass`.
int ipg_launch_ibm_totals(ipg_ctx *c, int pass) {
  if (c->pc.np <= 0) return 0;
  IbmArgs A; size_t lds;
  if (!ibm_args(c, A, lds)) return 1;
  if (pass == 1) A.fl = c->fl2;
  double **q = c->ps.q;
  const IbmTotArgs a{A, q[IPG_PFXLTOT], q[IPG_PFYLTOT], q[IPG_PFZLTOT], q[IPG_PTORQX], q[IPG_PTORQY], q[IPG_PTORQZ],
                     q[IPG_PTORQTHETA]};
  const int rc = c->ibm32 ? ipg_go<IbmTotK<float>>(c, k_ibm_totals<float>, dim3(A.np), dim3(IBM_TOT_THREADS), 0, a, "ibm_totals")
                          : ipg_go<IbmTotK<double>>(c, k_ibm_totals<double>, dim3(A.np), dim3(IBM_TOT_THREADS), 0, a, "ibm_totals");
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

// The device error word.  An IBM neighbour list that overflowed (more than 64
// neighbour images: a cluster far denser than random packing), a stencil
// outside its sphere's box, or a collision candidate list that overflowed
// (more than 48 spheres within reach of one) would have dropped contributions:
// fail.
int ipg_ibm_check(ipg_ctx *c) {
  if (c->pc.np <= 0 || !c->ibm_err) return 0;
  int e = 0;
  IPG_HIP(hipMemcpy(&e, c->ibm_err, sizeof(int), hipMemcpyDeviceToHost));
  if (e) {
    ipg_set_error(e & 1   ? "ibm: more than 64 neighbour images around a sphere"
                  : e & 2 ? "ibm: more than 4096 neighbour markers touching one sphere's stencils"
                  : e & 4 ? "ibm: a marker stencil outside its sphere's box"
                          : "collisions: more than 48 spheres within reach of one sphere");
    return 3;
  }
  return 0;
}

// The three passes, the final spread and the force sums (after the
// neighbour lists of ipg_launch_ibm_begin).
int ipg_launch_ibm_passes(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  for (int pass = 0; pass < 3; pass++)
    if (ipg_launch_ibm_gather(c, pass)) return 1;
  if (ipg_launch_ibm_final(c)) return 1;
  return ipg_launch_ibm_totals(c, 2);
}

// The whole IBM block of one substep, src/main.f90:296-338.
int ipg_launch_ibm(ipg_ctx *c) {
  if (c->pc.np <= 0) return 0;
  if (ipg_launch_ibm_begin(c)) return 1;
  return ipg_launch_ibm_passes(c);
}
