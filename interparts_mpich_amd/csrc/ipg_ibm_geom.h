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
using CompStencil = CompStencilT<double>;
template <bool WEIGHTS = true, typename T = double>
__host__ __device__ __forceinline__ void comp_stencil(const Grid &g, double xl, double yl, double zl, int comp,
                                                      CompStencilT<T> &c) {
  const int sxw = (int)floor((xl - 0.5 * g.dx) / g.lx) * g.nx;
  const int syw = (int)floor((yl - 0.5 * g.dy) / g.ly) * g.ny;
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
#pragma unroll
  for (int q = 0; q < 3 && WEIGHTS; q++) {
    c.wx[q] = roma_kernel((T)(comp == 0 ? ((double)c.i0 + (double)q) - cx : (((double)c.i0 + (double)q) - 0.5) - cx));
    c.wy[q] = roma_kernel((T)(comp == 1 ? ((double)c.j0 + (double)q) - cy : (((double)c.j0 + (double)q) - 0.5) - cy));
    c.wz[q] = roma_kernel((T)(comp == 2 ? ((double)c.k0 + (double)q) - cz : (((double)c.k0 + (double)q) - 0.5) - cz));
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
