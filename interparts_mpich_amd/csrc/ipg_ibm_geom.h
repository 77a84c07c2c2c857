// ipg_ibm_geom.h -- marker positions, stencil base cells and weights, and the
// per-sphere boxes of the IBM (src/coordsfp.f90, src/interp_spread.f90:76-130),
// host and device: the kernels in ipg_ibm.hip use them, and a host build of
// the same functions checks that every stencil lies in its sphere's box.
#pragma once
#include "ipg_internal.h"
#include <cmath>

__host__ __device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// src/coordsfp.f90:49-68 / src/intgr_nwtn_eulr.f90:1451-1465 marker position
// and the isout test; false when the marker is not kept.
__host__ __device__ __forceinline__ bool marker_pos(const Grid &g, double xp, double yp, double zp, double ex, double ey, double ez,
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

// Cells [b0, b0+bn) per direction in unwrapped index space (z clipped to the
// channel, 1..nz).
struct Box { int bi0, bj0, bk0, bx, by, bz; };

// The reach of a sphere: a box holding every stencil cell of its markers, of
// any component (cells floor(X) - 1 .. floor(X) + 2 for a marker at X cells),
// with one cell to spare at each end.  Used to find the spheres whose markers
// may reach another sphere's boxes.
__host__ __device__ __forceinline__ Box reach_box(const Grid &g, double radfp, double xp, double yp, double zp) {
  Box b;
  const int x0 = (int)floor((xp - radfp) * g.dxi), x1 = (int)floor((xp + radfp) * g.dxi);
  const int y0 = (int)floor((yp - radfp) * g.dyi), y1 = (int)floor((yp + radfp) * g.dyi);
  const int z0 = (int)floor((zp - radfp) * g.dzi), z1 = (int)floor((zp + radfp) * g.dzi);
  b.bi0 = x0 - 2; b.bx = x1 + 3 - b.bi0 + 1;
  b.bj0 = y0 - 2; b.by = y1 + 3 - b.bj0 + 1;
  b.bk0 = imax(z0 - 2, 1);
  b.bz = imax(imin(z1 + 3, g.nz) - b.bk0 + 1, 0);
  return b;
}

// First stencil cell of one direction over the markers of a sphere, [lo, hi].
// A marker at X = x / h cells (x = centre + offset, offset in [-radfp,
// radfp]) starts its stencil at floor(X) in a cell-centred direction and at
// nint(X) - 1 in the component's own (face) direction.  Xa = (centre -
// radfp) / h and Xb = (centre + radfp) / h bound X, but each of X, Xa and Xb
// is rounded, and a rounding can push a floor or nint over an integer by one
// (upward only for X, whose wrapped coordinate is positive); hence one cell
// more than the exact bounds at each end:
//   cell-centred: [floor(Xa) - 1, floor(Xb) + 1]
//   face:         [floor(Xa + 1/2) - 2, floor(Xb + 1/2)]
// so the stencils span at most ceil(2 radfp / h) + 5 cells (ibm_box_max).
__host__ __device__ __forceinline__ void first_cell_range(double lo_pos, double hi_pos, double hinv, bool face, int &lo, int &hi) {
  if (face) {
    lo = (int)floor(lo_pos * hinv + 0.5) - 2;
    hi = (int)floor(hi_pos * hinv + 0.5);
  } else {
    lo = (int)floor(lo_pos * hinv) - 1;
    hi = (int)floor(hi_pos * hinv) + 1;
  }
}

// The box of one (sphere, velocity component): the cells its markers'
// stencils of that component can touch, z clipped to 1..nz (the planes above
// kmax are the k1 plane, outside every box).  The first z plane of a stencil
// is clamped to 1 (and to nz - 2 for u and v) as comp_stencil clamps it.
__host__ __device__ __forceinline__ Box comp_box(const Grid &g, double radfp, double xp, double yp, double zp, int comp) {
  int x0, x1, y0, y1, z0, z1;
  first_cell_range(xp - radfp, xp + radfp, g.dxi, comp == 0, x0, x1);
  first_cell_range(yp - radfp, yp + radfp, g.dyi, comp == 1, y0, y1);
  first_cell_range(zp - radfp, zp + radfp, g.dzi, comp == 2, z0, z1);
  if (comp != 2) { z0 = imin(z0, g.nz - 2); z1 = imin(z1, g.nz - 2); }
  z0 = imax(z0, 1);
  z1 = imax(z1, 1);
  Box b;
  b.bi0 = x0; b.bx = x1 + 2 - x0 + 1;
  b.bj0 = y0; b.by = y1 + 2 - y0 + 1;
  b.bk0 = z0; b.bz = imax(imin(z1 + 2, g.nz) - z0 + 1, 0);
  return b;
}

// local box index of an unwrapped cell, -1 when outside the box
__host__ __device__ __forceinline__ int box_at(const Box &b, int iu, int ju, int k) {
  const int li = iu - b.bi0, lj = ju - b.bj0, lk = k - b.bk0;
  if ((unsigned)li >= (unsigned)b.bx || (unsigned)lj >= (unsigned)b.by || (unsigned)lk >= (unsigned)b.bz) return -1;
  return li + b.bx * (lj + b.by * lk);
}

// Marker stencil of one component in unwrapped indices: base cell (i0,j0,k0)
// and the three weights per direction, in the reference's order
// (src/interp_spread.f90:76-130).  Only the 9 weights the component uses are
// evaluated; each is the same expression stencil() evaluates.
// With WEIGHTS = false only the base cell is found.  T is the type of the
// weights: the offsets r are formed in FP64 either way and rounded to T for
// the kernel (FP32 in the mixed path).
template <typename T>
struct CompStencilT { int i0, j0, k0; T wx[3], wy[3], wz[3]; };

// The three Roma weights (src/kernel.f90:7-25) of a stencil direction whose
// middle point lies within half a cell of the marker (|r1| <= 1/2, the
// outer points at r1 -+ 1): the outer branch's 1 - |r| of either outer point
// is -+r1, so all three take the one square root sqrt(1 - 3 r1^2) that the
// middle point's inner branch takes.  The weights are the kernel's, each
// formula as roma_kernel writes it; only the outer points' square-root
// argument is r1^2 rather than (1 - |r|)^2 rounded on its own, a difference
// in the last bit.  Off-centre stencils (clamped at a wall) take roma_kernel
// per point.
__host__ __device__ __forceinline__ double roma_root(double x) { return sqrt(x); }
__host__ __device__ __forceinline__ float roma_root(float x) { return sqrt_f32(x); }
template <typename T>
__host__ __device__ __forceinline__ void roma3(T r0, T r1, T r2, T *w) {
  const T s = roma_root(-((T)3 * (r1 * r1)) + (T)1);
  w[1] = ((T)1 / (T)3) * ((T)1 + s);
  w[0] = ((T)1 / (T)6) * (((T)5 - (T)3 * (T)fabs(r0)) - s);
  w[2] = ((T)1 / (T)6) * (((T)5 - (T)3 * (T)fabs(r2)) - s);
}
using CompStencil = CompStencilT<double>;
template <bool WEIGHTS = true, typename T = double>
__host__ __device__ __forceinline__ void comp_stencil(const Grid &g, double xl, double yl, double zl, int comp,
                                                      CompStencilT<T> &c) {
  // floor(a / L) is 0 for 0 <= a < L without the division: there a / L < 1
  // exactly, and its rounding stays below 1 (a <= L's predecessor, so a / L
  // <= 1 - 2^-53); a division only for markers beyond a periodic edge
  const double ax = xl - 0.5 * g.dx, ay = yl - 0.5 * g.dy;
  const int sxw = ax >= 0.0 && ax < g.lx ? 0 : (int)floor(ax / g.lx) * g.nx;
  const int syw = ay >= 0.0 && ay < g.ly ? 0 : (int)floor(ay / g.ly) * g.ny;
  const double cx = xl * g.dxi - (double)sxw - 0.0;
  const double cy = yl * g.dyi - (double)syw - 0.0;
  const double cz = zl * g.dzi;
  // component 0 sits on x faces: the x weights are the face (s) ones, y and z
  // the half-shifted ones; likewise for 1 (y) and 2 (z)
  if (comp == 0) {
    c.i0 = nint_f(cx) - 1; c.j0 = (int)floor(cy); c.k0 = imax(imin((int)floor(cz), g.nz - 2), 1);
  } else if (comp == 1) {
    c.i0 = (int)floor(cx); c.j0 = nint_f(cy) - 1; c.k0 = imax(imin((int)floor(cz), g.nz - 2), 1);
  } else {
    c.i0 = (int)floor(cx); c.j0 = (int)floor(cy); c.k0 = imax(nint_f(cz) - 1, 1);
  }
  if (WEIGHTS) {
    T rx[3], ry[3], rz[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      rx[q] = (T)(comp == 0 ? ((double)c.i0 + (double)q) - cx : (((double)c.i0 + (double)q) - 0.5) - cx);
      ry[q] = (T)(comp == 1 ? ((double)c.j0 + (double)q) - cy : (((double)c.j0 + (double)q) - 0.5) - cy);
      rz[q] = (T)(comp == 2 ? ((double)c.k0 + (double)q) - cz : (((double)c.k0 + (double)q) - 0.5) - cz);
    }
    // x and y are periodic: their stencils are always centred on the marker
    roma3(rx[0], rx[1], rx[2], c.wx);
    roma3(ry[0], ry[1], ry[2], c.wy);
    if (fabs((double)rz[1]) <= 0.5) {
      roma3(rz[0], rz[1], rz[2], c.wz);
    } else {
#pragma unroll
      for (int q = 0; q < 3; q++) c.wz[q] = roma_kernel(rz[q]);
    }
  }
  c.i0 += sxw;
  c.j0 += syw;
}

// Largest comp_box extent in a direction of spacing 1/hinv: the first stencil
// cells span floor(Xb) - floor(Xa) + 2 (face: the same with Xa, Xb + 1/2) <=
// ceil(Xb - Xa) + 2, and a stencil is 3 cells; Xb - Xa exceeds 2 radfp hinv by
// roundings only, which an exact integer 2 radfp hinv would carry over to the
// next integer, hence floor(.. + 1e-9) + 1 rather than ceil.
inline int ibm_box_max(double radfp, double hinv) { return (int)std::floor(2.0 * radfp * hinv + 1e-9) + 1 + 5; }

// ---------------------------------------------------------------------------
// Index arithmetic of the k_ibm workgroups (ipg_ibm.hip), host and device, so
// that a host emulation of every thread checks each LDS and global index the
// kernel forms.

constexpr int IBM_THREADS = 512;  // threads per (sphere, component) workgroup
constexpr int IBM_MPT = 2;        // markers per thread: nl <= IBM_THREADS * IBM_MPT
constexpr int IBM_MAXNB = 64;     // neighbour images per sphere
constexpr int IBM_MAXNM = 4096;   // neighbour markers touching one (sphere, component) box
constexpr int IBM_CH = 12;        // cells per thread whose loads one staging chunk keeps in flight
constexpr int IBM_WAVES = IBM_THREADS / 64;  // wavefronts per workgroup

// periodic wrap of any unwrapped index onto 1..n
__host__ __device__ __forceinline__ int wrapany(int i, int n) {
  int r = i - 1;
  if ((unsigned)r >= (unsigned)n) {
    r %= n;
    if (r < 0) r += n;
  }
  return r + 1;
}

// offset in a k1 plane (nx*ny) of the unwrapped stencil cell (i, j).  A
// marker lies within radfp of its sphere's centre, which is inside the
// domain, and a box is narrower than the domain, so i is in [1-nx, 2nx] and
// one conditional shift per direction wraps it.
__host__ __device__ __forceinline__ long kplane(const Grid &g, int i, int j) {
  int r = i - 1, s = j - 1;
  r += r < 0 ? g.nx : 0;
  r -= r >= g.nx ? g.nx : 0;
  s += s < 0 ? g.ny : 0;
  s -= s >= g.ny ? g.ny : 0;
  return (long)r + (long)g.nx * s;
}

// a box [a0, a0+an) and a box [b0 + s n, b0 + s n + bn) overlap
__host__ __device__ __forceinline__ bool overlap(int a0, int an, int b0, int bn) { return a0 < b0 + bn && b0 < a0 + an; }

// The periodic images (sx nx, sy ny), sx, sy in -2..2, of sphere q's reach box
// bq that overlap sphere p's reach box bp: bit sx+2 of mx, bit sy+2 of my
// (an image is a pair of bits); false when none does.
__host__ __device__ __forceinline__ bool image_masks(const Grid &g, const Box &bp, const Box &bq, unsigned &mx, unsigned &my) {
  mx = 0u;
  my = 0u;
  if (!overlap(bp.bk0, bp.bz, bq.bk0, bq.bz)) return false;
#pragma unroll
  for (int sh = -2; sh <= 2; sh++) {
    if (overlap(bp.bi0, bp.bx, bq.bi0 + sh * g.nx, bq.bx)) mx |= 1u << (sh + 2);
    if (overlap(bp.bj0, bp.by, bq.bj0 + sh * g.ny, bq.by)) my |= 1u << (sh + 2);
  }
  return mx != 0u && my != 0u;
}

// LDS of one k_ibm workgroup, byte offsets from the start of its dynamic LDS:
// the box (bmx*bmy*bmz cells of esize bytes), two cell masks of one bit per
// cell (the cells the own stencils touch, and the cells a lower-numbered
// sphere owns), the field offsets of the box rows and the wrapped columns.
struct IbmLds { size_t box, flag, taken, rowoff, icol, bytes; int nflagmax; };
__host__ __device__ __forceinline__ IbmLds ibm_lds_layout(int bmx, int bmy, int bmz, size_t esize) {
  IbmLds L;
  const size_t cells = (size_t)bmx * bmy * bmz;
  L.nflagmax = (int)((cells + 31) / 32);
  L.box = 0;
  L.flag = cells * esize;
  L.taken = L.flag + (size_t)L.nflagmax * sizeof(unsigned);
  L.rowoff = L.taken + (size_t)L.nflagmax * sizeof(unsigned);
  L.icol = L.rowoff + (size_t)bmy * bmz * sizeof(int);
  L.bytes = L.icol + (size_t)bmx * sizeof(int);
  return L;
}

// Field offset of box cell li + bx (lj + by lk) = rowoff(row) + icol(li), row
// = lj + by lk, from a field base pointer at cell (1, 0, 0) (Grid::at): the
// periodic wrap and the multiplications once per row and column.
__host__ __device__ __forceinline__ int box_rowoff(const Grid &g, const Box &b, int row) {
  const int lk = row / b.by, lj = row - lk * b.by;
  return (int)(g.sx * wrapany(b.bj0 + lj, g.ny) + g.sxy * (b.bk0 + lk));
}
__host__ __device__ __forceinline__ int box_icol(const Grid &g, const Box &b, int li) { return wrapany(b.bi0 + li, g.nx) - 1; }

// Cell t = li + bx row of a box, walked by a stride of `threads` cells from
// t = tid: (li, row) advance by (rb, qb) with a carry instead of a division
// per cell.
struct BoxWalk {
  int li, row, rb, qb, bx;
  __host__ __device__ __forceinline__ BoxWalk(int bx_, int threads, int tid) : bx(bx_) {
    qb = threads / bx;
    rb = threads - qb * bx;
    row = tid / bx;
    li = tid - row * bx;
  }
  __host__ __device__ __forceinline__ void advance() {
    li += rb;
    row += qb;
    if (li >= bx) { li -= bx; row++; }
  }
};

// A marker's own stencil in its sphere's box b: the LDS index e0 of its first
// cell and how many of its z planes lie in the box (the rest are above kmax:
// the k1 plane).  False when the stencil is not held in x and y, or its first
// plane is not in the box (cannot happen for markers within radfp of the
// centre; the kernel reports it).
__host__ __device__ __forceinline__ bool own_cells(const Box &b, int i0, int j0, int k0, int &e0, int &nzin) {
  e0 = box_at(b, i0, j0, k0);
  if (e0 < 0 || box_at(b, i0 + 2, j0 + 2, k0) < 0) return false;
  nzin = imin(3, b.bk0 + b.bz - k0);
  return true;
}

// Stencil of the marker at offset (ex, ey, ez) from sphere q's centre (xq,
// yq, zq), seen through the periodic shift (sxs nx, sys ny) in the unwrapped
// frame of box b; false when it misses the box.
template <bool WEIGHTS = true, typename T = double>
__host__ __device__ __forceinline__ bool image_stencil(const Grid &g, const Box &b, int comp, double xq, double yq, double zq,
                                                      double ex, double ey, double ez, int sxs, int sys, CompStencilT<T> &c) {
  // a stencil spans less than two cells either side of its marker
  const double xs = (xq + ex) * g.dxi + sxs * g.nx, ys = (yq + ey) * g.dyi + sys * g.ny;
  const double zs = (zq + ez) * g.dzi;
  if (xs < b.bi0 - 3.0 || xs > b.bi0 + b.bx + 2.0 || ys < b.bj0 - 3.0 || ys > b.bj0 + b.by + 2.0 ||
      zs < b.bk0 - 3.0 || zs > b.bk0 + b.bz + 2.0)
    return false;
  double xl, yl, zl;
  if (!marker_pos(g, xq, yq, zq, ex, ey, ez, xl, yl, zl)) return false;
  comp_stencil<WEIGHTS, T>(g, xl, yl, zl, comp, c);
  c.i0 += sxs * g.nx;
  c.j0 += sys * g.ny;
  return !(c.i0 + 2 < b.bi0 || c.i0 >= b.bi0 + b.bx || c.j0 + 2 < b.bj0 || c.j0 >= b.bj0 + b.by ||
           c.k0 + 2 < b.bk0 || c.k0 >= b.bk0 + b.bz);
}
