"""Every index the IBM workgroups form stays inside what it indexes: a host
emulation of every thread of k_ibm (interparts_mpich_amd/csrc/ipg_ibm.hip),
built from the same index functions the kernel calls
(interparts_mpich_amd/csrc/ipg_ibm_geom.h: ibm_lds_layout, BoxWalk,
box_rowoff/box_icol, own_cells, image_stencil, image_masks, kplane).

The emulation follows the kernel's phases between its barriers, thread by
thread, for the neighbour search, pass 0 (MODE 0), passes 1 and 2 (MODE 1) and
the final spread (MODE 2), for workgroups of 384, 448 and 512 threads with 2
or 3 markers per thread and for FP64 and FP32 boxes, and checks:
  * the LDS regions (box, two cell masks, row offsets, columns) lie inside
    the dynamic LDS the launch asks for, do not overlap, and fit 160 KiB with
    the kernel's static LDS;
  * every LDS access is inside its region and inside the box of this
    workgroup, and reads only what an earlier phase wrote (no read of
    uninitialised LDS);
  * the box walk's (li, row) equals (t mod bx, t div bx) at every cell t;
  * every field access is an interior cell (1..nx, 1..ny, 1..nz) of the
    field's allocation and is the cell the box index names;
  * marker, force, neighbour-list and cell-flag indices stay inside their
    allocations, also when the neighbour lists overflow;
  * k1-plane indices stay inside the nx*ny plane;
  * the gathers read only cells that hold dudt (flagged by pass 0);
  * the final spread writes each touched cell exactly once (one owner), and
    every neighbour marker whose stencil touches a box is in its list
    (brute force over all spheres and images for some workgroups).

Cases: the 256^3 channel with 1000 spheres and the 32^3 case the Fortran
driver tests run (their seeded placements and the marker file as written),
the small parity grids, spheres pressed against both walls and the periodic
edges, and a cluster dense enough to overflow the neighbour lists.
"""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from interparts_mpich_amd.inputs import CONFIGS, Case, markers_for, place_spheres

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "interparts_mpich_amd/csrc/ipg_ibm_geom.h"
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static long nfail = 0, nchecks = 0;
#define CHECK(cond, ...)                                                  \
  do {                                                                    \
    nchecks++;                                                            \
    if (!(cond)) {                                                        \
      if (nfail < 25) {                                                   \
        std::printf("FAIL line %d [%s] ", __LINE__, cur.c_str());          \
        std::printf(__VA_ARGS__);                                         \
        std::printf("\n");                                                \
      }                                                                   \
      nfail++;                                                            \
    }                                                                     \
  } while (0)
static std::string cur;

struct CaseData {
  std::string name;
  int nx, ny, nz, np, nl, dense;
  double lx, ly, lz, radfp;
  std::vector<double> x, y, z, mk;
};

// allocation sizes of the context (ipg_init) the kernel indexes into
struct Globals {
  long nfield, M;
  std::vector<int> nbl, nbc, nml, nmc;
  std::vector<char> nbl_w, nml_w, nmc_w, flg_w;
  std::vector<unsigned> flg;
  std::vector<uint8_t> writes[3];  // final-spread stores per field cell
};

struct Ctx {
  const CaseData *cs;
  Grid g;
  int bm[3];
  int NT, MPT;
  size_t esize;
  long nflagmax;
  Globals *G;
  long touched_above = 0, overflowed = 0;
};

// interior cell of a field offset from the base pointer at cell (1, 0, 0)
static bool interior(const Grid &g, long nfield, long x, int &i, int &j, int &k) {
  if (x + 16 < 0 || x + 16 >= nfield) return false;
  k = (int)(x / g.sxy);
  const long r = x - (long)k * g.sxy;
  j = (int)(r / g.sx);
  i = (int)(r - (long)j * g.sx) + 1;
  return i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && k >= 1 && k <= g.nz;
}

static void neighbours(Ctx &C) {
  const CaseData &cs = *C.cs;
  Globals &G = *C.G;
  for (int p = 0; p < cs.np; p++) {
    const Box bp = reach_box(C.g, cs.radfp, cs.x[p], cs.y[p], cs.z[p]);
    int count = 0;
    for (int q = 0; q < cs.np; q++) {
      if (q == p) continue;
      const Box bq = reach_box(C.g, cs.radfp, cs.x[q], cs.y[q], cs.z[q]);
      unsigned mx, my;
      if (!image_masks(C.g, bp, bq, mx, my)) continue;
      int slot = count;
      count += __builtin_popcount(mx) * __builtin_popcount(my);
      for (int sx = 0; sx < 5; sx++)
        for (int sy = 0; sy < 5; sy++)
          if (((mx >> sx) & 1u) && ((my >> sy) & 1u)) {
            if (slot < IBM_MAXNB) {
              const long at = (long)p * IBM_MAXNB + slot;
              CHECK(at >= 0 && at < (long)G.nbl.size(), "nbl index %ld", at);
              G.nbl[at] = q * 25 + sx * 5 + sy;
              G.nbl_w[at] = 1;
            }
            slot++;
          }
    }
    G.nbc[p] = std::min(count, IBM_MAXNB);
    if (count > IBM_MAXNB) C.overflowed++;
  }
}

// One launch of k_ibm<MODE> (IBM pass `pass`) for workgroup (p, comp), phase by phase.
static void block(Ctx &C, int MODE, int pass, int p, int comp) {
  const CaseData &cs = *C.cs;
  const Grid &g = C.g;
  Globals &G = *C.G;
  const int NT = C.NT, MPT = C.MPT, blk = 3 * p + comp;
  const bool F64 = C.esize == 8;
  Box b = comp_box(g, cs.radfp, cs.x[p], cs.y[p], cs.z[p], comp);
  CHECK(b.bx <= C.bm[0] && b.by <= C.bm[1] && b.bz <= C.bm[2], "box %dx%dx%d beyond %d %d %d", b.bx, b.by, b.bz, C.bm[0],
        C.bm[1], C.bm[2]);
  b.bx = imin(b.bx, C.bm[0]); b.by = imin(b.by, C.bm[1]); b.bz = imin(b.bz, C.bm[2]);
  const int nbox = b.bx * b.by * b.bz, bxy = b.bx * b.by;
  const IbmLds L = ibm_lds_layout(C.bm[0], C.bm[1], C.bm[2], C.esize);
  const int nflag = (nbox + 31) >> 5;
  CHECK(nflag <= L.nflagmax, "nflag %d > %d", nflag, L.nflagmax);
  const long M = G.M;
  // LDS, with a written mark per element
  std::vector<double> box(nbox, 0.0);
  std::vector<char> box_w(nbox, 0);
  std::vector<unsigned> flag(L.nflagmax, 0u), taken(L.nflagmax, 0u);
  std::vector<char> flag_w(L.nflagmax, 0), taken_w(L.nflagmax, 0);
  std::vector<int> rowoff(C.bm[1] * C.bm[2], 0), icol(C.bm[0], 0);
  std::vector<char> rowoff_w(rowoff.size(), 0), icol_w(icol.size(), 0);
  auto box_ok = [&](long e) { return e >= 0 && e < nbox; };
  auto flagged = [&](int t) -> bool {
    CHECK(t >= 0 && t < nbox && (t >> 5) < nflag, "flag read t=%d", t);
    CHECK(flag_w[t >> 5], "flag word %d read before written", t >> 5);
    return (flag[t >> 5] >> (t & 31)) & 1u;
  };
  auto field_at = [&](int row, int li, int t) -> long {
    CHECK(row >= 0 && row < b.by * b.bz && row < (int)rowoff.size(), "row %d", row);
    CHECK(li >= 0 && li < b.bx && li < (int)icol.size(), "col %d", li);
    CHECK(rowoff_w[row] && icol_w[li], "row/col offsets read before written");
    const long x = (long)rowoff[row] + icol[li];
    int i, j, k;
    CHECK(interior(g, G.nfield, x, i, j, k), "field offset %ld not interior", x);
    // the cell the box index t names
    const int lk = t / bxy, lj = (t - lk * bxy) / b.bx, lii = t - lk * bxy - lj * b.bx;
    CHECK(i == wrapany(b.bi0 + lii, g.nx) && j == wrapany(b.bj0 + lj, g.ny) && k == b.bk0 + lk,
          "field cell (%d %d %d) is not box cell %d", i, j, k, t);
    return x;
  };
  auto walk_ok = [&](const BoxWalk &w, int t) {
    CHECK(w.li == t % b.bx && w.row == t / b.bx, "walk (%d %d) at t=%d bx=%d", w.li, w.row, t, b.bx);
  };
  auto mark_of = [&](int l, double &xl, double &yl, double &zl) {
    CHECK(l >= 0 && l < cs.nl, "marker %d", l);
    return marker_pos(g, cs.x[p], cs.y[p], cs.z[p], cs.mk[l], cs.mk[cs.nl + l], cs.mk[2 * cs.nl + l], xl, yl, zl);
  };
  // own marker q of thread tid (the kernel's own())
  auto own = [&](int tid, int q, CompStencil &c, int &e0, int &nzin) -> bool {
    const int l = tid + q * NT;
    double xl, yl, zl;
    if (l >= cs.nl || !mark_of(l, xl, yl, zl)) return false;
    comp_stencil<true, double>(g, xl, yl, zl, comp, c);
    const bool ok = own_cells(b, c.i0, c.j0, c.k0, e0, nzin);
    CHECK(ok, "own stencil of marker %d outside the box", l);
    if (!ok) return false;
    CHECK(nzin >= 1 && nzin <= 3, "nzin %d", nzin);
    return true;
  };
  auto own_cell = [&](const CompStencil &c, int e0, int i, int j, int k) -> int {
    const int e = e0 + i + b.bx * j + bxy * k;
    CHECK(box_ok(e) && box_at(b, c.i0 + i, c.j0 + j, c.k0 + k) == e, "own cell %d", e);
    return e;
  };
  // Kzero: every index of the grid-stride loop is < nx*ny by its bound
  // ---- phase 1
  const long gbase = (long)blk * L.nflagmax;
  for (int tid = 0; tid < NT; tid++) {
    for (int t = tid; t < nflag; t += NT) {
      CHECK(t < L.nflagmax, "flag init %d", t);
      if (MODE == 0) flag[t] = 0u;
      else {
        CHECK(gbase + t < (long)G.flg.size() && G.flg_w[gbase + t], "gflag read %ld", gbase + t);
        flag[t] = G.flg[gbase + t];
      }
      flag_w[t] = 1;
      if (MODE == 2) { taken[t] = 0u; taken_w[t] = 1; }
    }
    for (int t = tid; t < b.by * b.bz; t += NT) {
      CHECK(t < (int)rowoff.size(), "rowoff %d", t);
      rowoff[t] = box_rowoff(g, b, t);
      rowoff_w[t] = 1;
    }
    for (int t = tid; t < b.bx; t += NT) {
      CHECK(t < (int)icol.size(), "icol %d", t);
      icol[t] = box_icol(g, b, t);
      icol_w[t] = 1;
    }
    if (MODE >= 1)
      for (int t = tid; t < nbox; t += NT) { box[t] = 0.0; box_w[t] = 1; }
  }
  // ---- phase 2
  int lcount = 0;
  const long lbase = (long)blk * IBM_MAXNM;
  for (int tid = 0; tid < NT; tid++) {
    const BoxWalk w0(b.bx, NT, tid);
    if (MODE == 1 || (MODE == 2 && F64)) {
      BoxWalk w = w0;
      for (int t0 = tid; t0 < nbox; t0 += IBM_CH * NT)
        for (int u = 0; u < IBM_CH; u++) {
          const int t = t0 + u * NT;
          if (t < nbox) walk_ok(w, t);
          if (t < nbox && flagged(t)) {
            field_at(w.row, w.li, t);
            CHECK(box_w[t], "box %d added to before written", t);
          }
          w.advance();
        }
    }
    // passes 1, 2 and the final spread load each thread's first list entry
    // early, before the count is known: in bounds for every thread
    if (MODE >= 1) CHECK(lbase + tid < (long)G.nml.size() && tid < IBM_MAXNM, "early nml read %ld", lbase + tid);
    if (MODE == 0) {
      // pass 0 loads the image of each thread's first candidate early
      const long at0 = (long)p * IBM_MAXNB + std::min(tid / cs.nl, IBM_MAXNB - 1);
      CHECK(at0 >= 0 && at0 < (long)G.nbl.size(), "early nbl read %ld", at0);
      if (tid < G.nbc[p] * cs.nl) CHECK(G.nbl_w[at0], "early nbl read %ld before written", at0);
      CHECK(G.nbc[p] >= 0 && G.nbc[p] <= IBM_MAXNB, "nbc %d", G.nbc[p]);
      const int tot = G.nbc[p] * cs.nl;
      for (int idx = tid; idx < tot; idx += NT) {
        const int n = idx / cs.nl, l = idx - n * cs.nl;
        const long at = (long)p * IBM_MAXNB + n;
        CHECK(at < (long)G.nbl.size() && G.nbl_w[at], "nbl read %ld", at);
        const int code = G.nbl[at];
        const int qs = code / 25, sh = code % 25;
        CHECK(qs >= 0 && qs < cs.np && qs != p, "neighbour sphere %d", qs);
        CompStencil c;
        if (!image_stencil<false>(g, b, comp, cs.x[qs], cs.y[qs], cs.z[qs], cs.mk[l], cs.mk[cs.nl + l], cs.mk[2 * cs.nl + l],
                                  sh / 5 - 2, sh % 5 - 2, c))
          continue;
        const int slot = lcount++;
        if (slot < IBM_MAXNM) {
          CHECK(lbase + slot < (long)G.nml.size(), "nml write %ld", lbase + slot);
          G.nml[lbase + slot] = (qs * cs.nl + l) * 25 + sh;
          G.nml_w[lbase + slot] = 1;
        }
      }
    } else {
      CHECK(G.nmc_w[blk], "nmc read before written");
      const int cnt = G.nmc[blk];
      CHECK(cnt >= 0 && cnt <= IBM_MAXNM, "nmc %d", cnt);
      for (int idx = tid; idx < cnt; idx += NT) {
        CHECK(G.nml_w[lbase + idx], "nml read before written");
        const int code = G.nml[lbase + idx];
        const int mq = code / 25, sh = code % 25;
        const int qs = mq / cs.nl, l = mq - qs * cs.nl;
        CHECK(mq >= 0 && mq < M && (long)comp * M + mq < 3 * M, "neighbour force index %d", mq);
        CompStencil c;
        if (!image_stencil<true, double>(g, b, comp, cs.x[qs], cs.y[qs], cs.z[qs], cs.mk[l], cs.mk[cs.nl + l],
                                         cs.mk[2 * cs.nl + l], sh / 5 - 2, sh % 5 - 2, c)) {
          CHECK(false, "listed neighbour marker misses the box");
          continue;
        }
        for (int k = 0; k < 3; k++)
          for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
              const int e = box_at(b, c.i0 + i, c.j0 + j, c.k0 + k);
              if (e < 0) continue;
              CHECK(box_ok(e) && box_w[e], "neighbour spread into %d", e);
              if (MODE == 2 && qs < p) {
                CHECK((e >> 5) < nflag && taken_w[e >> 5], "taken word %d", e >> 5);
                taken[e >> 5] |= 1u << (e & 31);
              }
            }
      }
    }
    for (int q = 0; q < MPT; q++) {
      CompStencil c;
      int e0, nzin;
      if (!own(tid, q, c, e0, nzin)) continue;
      const long m = (long)p * cs.nl + tid + q * NT;
      CHECK(m < M && (long)comp * M + m < 3 * M, "force index %ld", m);
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++) {
            if (k >= nzin) continue;
            const int e = own_cell(c, e0, i, j, k);
            if (MODE == 0) {
              CHECK((e >> 5) < nflag && flag_w[e >> 5], "flag word %d", e >> 5);
              flag[e >> 5] |= 1u << (e & 31);
            } else {
              CHECK(box_w[e], "own spread into unwritten %d", e);
            }
          }
    }
  }
  // ---- phase 3
  if (MODE == 0) {
    for (int tid = 0; tid < NT; tid++)
      for (int t = tid; t < nflag; t += NT) {
        CHECK(gbase + t < (long)G.flg.size(), "gflag write %ld", gbase + t);
        G.flg[gbase + t] = flag[t];
        G.flg_w[gbase + t] = 1;
      }
    G.nmc[blk] = std::min(lcount, IBM_MAXNM);
    G.nmc_w[blk] = 1;
    if (lcount > IBM_MAXNM) C.overflowed++;
  }
  if (MODE == 2) {
    for (int tid = 0; tid < NT; tid++) {
      BoxWalk w(b.bx, NT, tid);
      for (int t = tid; t < nbox; t += NT) {
        walk_ok(w, t);
        if (flagged(t)) {
          CHECK(taken_w[t >> 5], "taken read before written");
          if (!((taken[t >> 5] >> (t & 31)) & 1u)) {
            CHECK(box_w[t], "box %d stored before written", t);
            const long x = field_at(w.row, w.li, t);
            if (x + 16 >= 0 && x + 16 < G.nfield) {
              uint8_t &n = G.writes[comp][x + 16];
              if (n < 255) n++;
            }
          }
        }
        w.advance();
      }
    }
    return;
  }
  if (MODE == 0) {
    for (int tid = 0; tid < NT; tid++) {
      BoxWalk w(b.bx, NT, tid);
      for (int t0 = tid; t0 < nbox; t0 += IBM_CH * NT) {
        for (int u = 0; u < IBM_CH; u++) {
          const int t = t0 + u * NT;
          if (t < nbox) walk_ok(w, t);
          if (t < nbox && flagged(t)) field_at(w.row, w.li, t);
          w.advance();
        }
        for (int u = 0; u < IBM_CH; u++) {
          const int t = t0 + u * NT;
          if (t < nbox) box_w[t] = 1;
        }
      }
    }
  }
  // ---- phase 4: gathers and the k1 plane
  const long nplane = (long)g.nx * g.ny;
  for (int tid = 0; tid < NT; tid++)
    for (int q = 0; q < MPT; q++) {
      CompStencil c;
      int e0, nzin;
      if (!own(tid, q, c, e0, nzin)) continue;
      for (int k = 0; k < 3; k++)
        for (int j = 0; j < 3; j++)
          for (int i = 0; i < 3; i++)
            if (k < nzin) {
              const int e = own_cell(c, e0, i, j, k);
              CHECK(box_w[e], "gather reads unwritten box cell %d", e);
              CHECK(flagged(e), "gather reads a cell that holds no dudt: %d", e);
            }
      if (nzin < 3) {
        for (int k = 1; k < 3; k++)
          for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++)
              if (k >= nzin) {
                const int iu = c.i0 + i, ju = c.j0 + j;
                CHECK(iu >= 1 - g.nx && iu <= 2 * g.nx && ju >= 1 - g.ny && ju <= 2 * g.ny, "k1 cell (%d %d)", iu, ju);
                const long kp = kplane(g, iu, ju);
                CHECK(kp >= 0 && kp < nplane, "k1 plane index %ld", kp);
                if (c.k0 + k != g.nz + 1) C.touched_above++;
              }
      }
      const long m = (long)p * cs.nl + tid + q * NT;
      CHECK(m < M, "force store %ld", m);
    }
  // the per-wavefront force sums: psum[pass][3 np][waves]
  for (int tid = 0; tid < NT; tid += 64) {
    const long at = ((long)pass * 3 * cs.np + blk) * (NT / 64) + tid / 64;
    CHECK(at >= 0 && at < 9L * cs.np * (NT / 64) && NT % 64 == 0, "psum index %ld", at);
  }
}

// Every neighbour marker whose stencil touches box (p, comp), through any
// image of any other sphere, is in the list of that box.
static void completeness(Ctx &C, int p, int comp) {
  const CaseData &cs = *C.cs;
  Globals &G = *C.G;
  const int blk = 3 * p + comp;
  const Box b = comp_box(C.g, cs.radfp, cs.x[p], cs.y[p], cs.z[p], comp);
  std::vector<int> listed(G.nml.begin() + (long)blk * IBM_MAXNM, G.nml.begin() + (long)blk * IBM_MAXNM + G.nmc[blk]);
  std::sort(listed.begin(), listed.end());
  CHECK(std::adjacent_find(listed.begin(), listed.end()) == listed.end(), "a neighbour marker listed twice");
  long want = 0;
  for (int q = 0; q < cs.np; q++) {
    if (q == p) continue;
    for (int sx = -2; sx <= 2; sx++)
      for (int sy = -2; sy <= 2; sy++)
        for (int l = 0; l < cs.nl; l++) {
          CompStencil c;
          if (!image_stencil<false>(C.g, b, comp, cs.x[q], cs.y[q], cs.z[q], cs.mk[l], cs.mk[cs.nl + l], cs.mk[2 * cs.nl + l],
                                    sx, sy, c))
            continue;
          want++;
          const int code = (q * cs.nl + l) * 25 + (sx + 2) * 5 + (sy + 2);
          CHECK(std::binary_search(listed.begin(), listed.end(), code), "sphere %d marker %d image (%d %d) not listed for %d/%d",
                q, l, sx, sy, p, comp);
        }
  }
  CHECK(want == (long)listed.size(), "listed %zu, touching %ld", listed.size(), want);
}

// Each cell some own stencil touches (z <= nz) is stored by exactly one
// workgroup of the final spread.
static void ownership(Ctx &C) {
  const CaseData &cs = *C.cs;
  const Grid &g = C.g;
  Globals &G = *C.G;
  for (int comp = 0; comp < 3; comp++) {
    std::vector<char> touched(G.nfield, 0);
    for (int p = 0; p < cs.np; p++)
      for (int l = 0; l < cs.nl; l++) {
        double xl, yl, zl;
        if (!marker_pos(g, cs.x[p], cs.y[p], cs.z[p], cs.mk[l], cs.mk[cs.nl + l], cs.mk[2 * cs.nl + l], xl, yl, zl)) continue;
        CompStencil c;
        comp_stencil<false>(g, xl, yl, zl, comp, c);
        for (int k = 0; k < 3; k++)
          for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
              if (c.k0 + k > g.nz) continue;
              touched[16 + g.at(wrapany(c.i0 + i, g.nx), wrapany(c.j0 + j, g.ny), c.k0 + k)] = 1;
            }
      }
    long bad = 0;
    for (long x = 0; x < G.nfield; x++)
      if ((touched[x] ? 1 : 0) != G.writes[comp][x]) {
        if (bad < 3) std::printf("  comp %d cell %ld touched %d stored %d times\n", comp, x - 16, touched[x], G.writes[comp][x]);
        bad++;
      }
    CHECK(bad == 0, "%ld cells of component %d not stored exactly once", bad, comp);
  }
}

static bool read_case(FILE *f, CaseData &c) {
  int hdr[6];
  if (std::fread(hdr, sizeof(int), 6, f) != 6) return false;
  c.nx = hdr[0]; c.ny = hdr[1]; c.nz = hdr[2]; c.np = hdr[3]; c.nl = hdr[4]; c.dense = hdr[5];
  double d[4];
  if (std::fread(d, sizeof(double), 4, f) != 4) return false;
  c.lx = d[0]; c.ly = d[1]; c.lz = d[2]; c.radfp = d[3];
  char name[32];
  if (std::fread(name, 1, 32, f) != 32) return false;
  c.name = std::string(name, strnlen(name, 32));
  c.x.resize(c.np); c.y.resize(c.np); c.z.resize(c.np); c.mk.resize(3 * (size_t)c.nl);
  return std::fread(c.x.data(), 8, c.np, f) == (size_t)c.np && std::fread(c.y.data(), 8, c.np, f) == (size_t)c.np &&
         std::fread(c.z.data(), 8, c.np, f) == (size_t)c.np && std::fread(c.mk.data(), 8, 3 * c.nl, f) == (size_t)(3 * c.nl);
}

int main(int argc, char **argv) {
  FILE *f = std::fopen(argv[1], "rb");
  std::vector<CaseData> cases;
  CaseData c;
  while (read_case(f, c)) cases.push_back(c);
  std::fclose(f);
  // one workgroup shape per run: threads, markers per thread
  const int threads[] = {std::atoi(argv[2])}, mpts[] = {std::atoi(argv[3])};
  bool shipped = false;
  for (const CaseData &cs : cases) {
    Grid g{};
    g.nx = cs.nx; g.ny = cs.ny; g.nz = cs.nz;
    g.lx = cs.lx; g.ly = cs.ly; g.lz = cs.lz;
    g.dx = cs.lx / cs.nx; g.dy = cs.ly / cs.ny; g.dz = cs.lz / cs.nz;
    g.dxi = 1.0 / g.dx; g.dyi = 1.0 / g.dy; g.dzi = 1.0 / g.dz;
    g.sx = ((long)cs.nx + 2 + 15) / 16 * 16;  // ipg_init's row pitch
    g.sxy = g.sx * (cs.ny + 2);
    const int bm[3] = {ibm_box_max(cs.radfp, g.dxi), ibm_box_max(cs.radfp, g.dyi), ibm_box_max(cs.radfp, g.dzi)};
    for (int NT : threads)
      for (int MPT : mpts)
        for (size_t esize : {(size_t)8, (size_t)4}) {
          char buf[160];
          std::snprintf(buf, sizeof buf, "%s threads=%d mpt=%d esize=%zu", cs.name.c_str(), NT, MPT, esize);
          cur = buf;
          if (cs.nl > NT * MPT) continue;  // ipg_ibm_setup refuses it
          if (NT == IBM_THREADS && MPT == IBM_MPT) shipped = true;
          const IbmLds L = ibm_lds_layout(bm[0], bm[1], bm[2], esize);
          // setup's own checks (ipg_ibm_setup); the kernel's static LDS is at most 16 bytes
          if (L.bytes + 16 > 160 * 1024 || bm[0] + 1 > cs.nx || bm[1] + 1 > cs.ny) { std::printf("%s: refused by setup\n", buf); continue; }
          const size_t cells = (size_t)bm[0] * bm[1] * bm[2];
          CHECK(L.box == 0 && L.flag >= cells * esize && L.taken >= L.flag + L.nflagmax * 4 && L.rowoff >= L.taken + L.nflagmax * 4 &&
                L.icol >= L.rowoff + (size_t)bm[1] * bm[2] * 4 && L.bytes >= L.icol + (size_t)bm[0] * 4,
                "LDS regions overlap");
          CHECK(L.flag % 4 == 0 && L.rowoff % 4 == 0 && L.icol % 4 == 0, "LDS region misaligned");
          CHECK((size_t)L.nflagmax * 32 >= cells, "flag words %d for %zu cells", L.nflagmax, cells);
          Globals G;
          G.nfield = (16 + g.sxy * (cs.nz + 2) + 16 + 15) / 16 * 16;
          G.M = (long)cs.np * cs.nl;
          G.nbl.assign((size_t)cs.np * IBM_MAXNB, -1); G.nbl_w.assign(G.nbl.size(), 0);
          G.nbc.assign(cs.np, 0);
          G.nml.assign((size_t)cs.np * 3 * IBM_MAXNM, -1); G.nml_w.assign(G.nml.size(), 0);
          G.nmc.assign((size_t)cs.np * 3, 0); G.nmc_w.assign(G.nmc.size(), 0);
          G.flg.assign((size_t)cs.np * 3 * L.nflagmax, 0u); G.flg_w.assign(G.flg.size(), 0);
          for (int q = 0; q < 3; q++) G.writes[q].assign(G.nfield, 0);
          Ctx C{&cs, g, {bm[0], bm[1], bm[2]}, NT, MPT, esize, L.nflagmax, &G};
          neighbours(C);
          const int modes[4] = {0, 1, 1, 2}, passes[4] = {0, 1, 2, 2};
          for (int launch = 0; launch < 4; launch++)
            for (int blk = 0; blk < 3 * cs.np; blk++) block(C, modes[launch], passes[launch], blk / 3, blk % 3);
          if (!cs.dense) {
            CHECK(C.overflowed == 0, "a neighbour list overflowed");
            ownership(C);
            if (esize == 8 && MPT == 2)
              for (int blk = 0; blk < std::min(3 * cs.np, 24); blk++) completeness(C, blk / 3, blk % 3);
          } else {
            CHECK(C.overflowed > 0, "the dense case did not overflow a list");
          }
          std::printf("%s: np=%d nl=%d box max %d %d %d, LDS %zu B, overflowed %ld, stencil planes above k1 %ld\n", buf, cs.np,
                      cs.nl, bm[0], bm[1], bm[2], L.bytes, C.overflowed, C.touched_above);
        }
  }
  std::printf("shipped shape %d %d: %s\n", IBM_THREADS, IBM_MPT, shipped ? "checked" : "not this run");
  std::printf("checks %ld failures %ld\n", nchecks, nfail);
  return nfail != 0;
}
"""


def _marker_offsets(case):
    """Marker offsets from the centre as ipg_coordsfp forms them from the
    marker file (src/coordsfp.f90:49-51)."""
    theta, phi, radfp = markers_for(case)
    mk = np.concatenate([radfp * np.sin(0.0 + theta) * np.cos(0.0 + phi),
                         radfp * np.sin(0.0 + theta) * np.sin(0.0 + phi),
                         radfp * np.cos(0.0 + theta)])
    return mk, radfp


def _pack(name, case, x, y, z, dense=False):
    mk, radfp = _marker_offsets(case)
    hdr = struct.pack("6i", case.itot, case.jtot, case.ktot, len(x), case.nl, int(dense))
    hdr += struct.pack("4d", case.lx, case.ly, case.lz, radfp)
    hdr += name.encode()[:32].ljust(32, b"\0")
    return hdr + b"".join(np.ascontiguousarray(a, dtype=np.float64).tobytes() for a in (x, y, z, mk))


def _cases():
    out = []
    for name in ("ch256_1000sph", "t32", "t16", "t24x16x32", "ch64_8sph"):
        case = CONFIGS[name]
        out.append(_pack(name, case, *place_spheres(case)))
    rng = np.random.default_rng(20261204)
    # spheres pressed against the walls (centres down to radfp - h/2 from
    # either wall, markers beyond it) and on the periodic edges
    case = CONFIGS["t32"]
    h, R = case.lx / case.itot, case.radius
    radfp = R - 0.3 * h
    zs = np.array([radfp - 0.5 * h, radfp, R, case.lz - R, case.lz - radfp, case.lz - radfp + 0.5 * h])
    xs = np.array([0.0, 1e-15, case.lx - 1e-15, 0.5 * h, case.lx - 0.5 * h, 0.5 * case.lx])
    ys = np.array([case.ly - 1e-15, 0.0, 0.5 * h, 0.25 * case.ly, case.ly - 0.5 * h, 1e-15])
    out.append(_pack("walls_edges_32", case, xs, ys, zs))
    # many spheres on a 64^3 grid, random centres (overlaps allowed)
    case = Case(name="rand64", itot=64, jtot=64, ktot=64, np=60, diameter_h=8.0)
    R = case.radius
    out.append(_pack("random_64", case, rng.uniform(0, case.lx, 60), rng.uniform(0, case.ly, 60),
                     rng.uniform(R, case.lz - R, 60)))
    # a cluster dense enough to overflow the neighbour lists
    case = Case(name="dense64", itot=64, jtot=64, ktot=64, np=90, diameter_h=8.0)
    out.append(_pack("dense_overflow_64", case, rng.uniform(0.40, 0.52, 90), rng.uniform(0.40, 0.52, 90),
                     rng.uniform(0.40, 0.52, 90), dense=True))
    return out


def test_every_ibm_index_in_bounds(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    data = tmp_path / "cases.bin"
    data.write_bytes(b"".join(_cases()))
    src = tmp_path / "ibm_index.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "ibm_index"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        f"-I{ROOT}", str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    # one process per workgroup shape, side by side
    shapes = [(t, m) for t in (384, 448, 512) for m in (2, 3)]
    runs = [subprocess.Popen([str(exe), str(data), str(t), str(m)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                             text=True) for t, m in shapes]
    outs = [r.communicate(timeout=900)[0] for r in runs]
    shipped = 0
    for (t, m), r, out in zip(shapes, runs, outs):
        print(f"--- {t} threads, {m} markers per thread\n" + out[-3000:])
        assert r.returncode == 0, out[-4000:]
        assert "failures 0" in out
        shipped += ": checked" in out
    # the library's own shape is among those checked
    assert shipped == 1
