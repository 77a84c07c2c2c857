"""The IBM's per-(sphere, component) boxes hold every stencil of the sphere's
markers (host build of interparts_mpich_amd/csrc/ipg_ibm_geom.h, the same
functions the kernels use).

Each (sphere, velocity component) workgroup keeps the cells its markers'
27-point stencils (src/interp_spread.f90:76-130) touch in an LDS box sized by
ibm_box_max; a stencil outside its box would be reported as an error.  This
test places sphere centres at random, and so that a marker on an axis (at
+-radfp from the centre) lands on, or an ulp from, an integer or half-integer
cell coordinate, and checks for every marker and component that the stencil
lies in the box, that the box is no larger than ibm_box_max, and that it lies
inside the sphere's reach box (the neighbour search's box, at most one cell
wider).  ibm_box_max is tight: one cell less fails here.  The box's one-cell
margin at each end covers a rounding of the marker coordinate across an
integer, which these placements do not produce (with centres inside the
domain it cannot reach the box's extreme cells); the check passes with either
margin removed, so the margins are insurance, not pinned by this test.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "interparts_mpich_amd/csrc/ipg_ibm_geom.h"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

// grid points per direction, domain lengths, sphere radius in cells, markers
struct Cfg { int nx, ny, nz; double lx, ly, radius_h; int nl; };

int main(int argc, char **argv) {
  const int trials = std::atoi(argv[1]);
  const Cfg cfgs[] = {{64, 64, 64, 1.0, 1.0, 4.0, 201},    {256, 256, 256, 1.0, 1.0, 6.0, 409},
                      {512, 512, 512, 1.0, 1.0, 8.0, 746},  {96, 96, 64, 1.5, 1.5, 5.5, 300},
                      {128, 128, 192, 2.0, 2.0, 6.0, 452},  {64, 64, 64, 1.0, 1.0, 6.3, 452},
                      {64, 64, 64, 1.0, 1.0, 5.8, 452},     {80, 64, 64, 1.25, 1.0, 4.3, 201}};
  std::mt19937_64 rng(20261203);
  std::uniform_real_distribution<double> U(0.0, 1.0);
  long tested = 0, bad = 0;
  for (const Cfg &cf : cfgs) {
    Grid g{};
    g.nx = cf.nx; g.ny = cf.ny; g.nz = cf.nz;
    g.lx = cf.lx; g.ly = cf.ly;
    g.dx = cf.lx / cf.nx; g.dy = cf.ly / cf.ny; g.dz = g.dx; g.lz = g.dz * cf.nz;
    g.dxi = 1.0 / g.dx; g.dyi = 1.0 / g.dy; g.dzi = 1.0 / g.dz;
    // radfp = radius - 0.3 h (the marker shell of lagrangianforcepoints2)
    const double radfp = (cf.radius_h - 0.3) * g.dx;
    // a golden-angle spiral, as the marker file, plus the six markers on the
    // axes (at exactly +-radfp along one axis), which the box bounds must hold
    // when the placement below puts them within an ulp of a cell boundary
    const double pi = 3.14159265358979323846;
    const int nl = cf.nl + 6;
    std::vector<double> th(nl), ph(nl), mk(3 * (size_t)nl);
    const double ga = pi * (3.0 - std::sqrt(5.0));
    for (int l = 0; l < cf.nl; l++) {
      th[l] = std::acos(1.0 - (2.0 * l + 1.0) / cf.nl);
      ph[l] = std::fmod(l * ga, 2 * pi);
    }
    const double axes[6][2] = {{pi / 2, 0.0}, {pi / 2, pi}, {pi / 2, pi / 2}, {pi / 2, 3 * pi / 2}, {0.0, 0.0}, {pi, 0.0}};
    for (int q = 0; q < 6; q++) { th[cf.nl + q] = axes[q][0]; ph[cf.nl + q] = axes[q][1]; }
    for (int l = 0; l < nl; l++) {  // src/coordsfp.f90:49-51
      mk[l] = radfp * std::sin(th[l]) * std::cos(ph[l]);
      mk[nl + l] = radfp * std::sin(th[l]) * std::sin(ph[l]);
      mk[2 * nl + l] = radfp * std::cos(th[l]);
    }
    const int bm[3] = {ibm_box_max(radfp, g.dxi), ibm_box_max(radfp, g.dyi), ibm_box_max(radfp, g.dzi)};
    for (int it = 0; it < trials; it++) {
      double c[3];
      for (int d = 0; d < 3; d++) {
        const double L = d == 0 ? g.lx : d == 1 ? g.ly : g.lz, h = d == 0 ? g.dx : d == 1 ? g.dy : g.dz;
        double v;
        const int mode = it % 4;
        if (mode == 0) {
          v = U(rng) * L;
        } else {
          // an extreme marker at (or an ulp from) an integer or half-integer cell coordinate
          const double k = std::floor(U(rng) * (L / h)) + (mode == 2 ? 0.5 : 0.0);
          v = (k + (U(rng) - 0.5) * 1e-13 * (it % 3)) * h + (U(rng) < 0.5 ? radfp : -radfp);
          if (mode == 3) v = std::nextafter(v, U(rng) < 0.5 ? -1e9 : 1e9);
        }
        if (d == 2) v = std::min(std::max(v, radfp + 0.05 * h), L - radfp - 0.05 * h);
        else { v = std::fmod(v, L); if (v < 0) v += L; }
        c[d] = v;
      }
      const Box r = reach_box(g, radfp, c[0], c[1], c[2]);
      if (r.bx > bm[0] + 1 || r.by > bm[1] + 1) bad++;
      for (int comp = 0; comp < 3; comp++) {
        const Box b = comp_box(g, radfp, c[0], c[1], c[2], comp);
        if (b.bx > bm[0] || b.by > bm[1] || b.bz > bm[2]) bad++;
        if (b.bi0 < r.bi0 || b.bi0 + b.bx > r.bi0 + r.bx || b.bj0 < r.bj0 || b.bj0 + b.by > r.bj0 + r.by ||
            b.bk0 < r.bk0 || b.bk0 + b.bz > r.bk0 + r.bz)
          bad++;
        for (int l = 0; l < nl; l++) {
          double xl, yl, zl;
          if (!marker_pos(g, c[0], c[1], c[2], mk[l], mk[nl + l], mk[2 * nl + l], xl, yl, zl)) continue;
          CompStencil s;
          comp_stencil<false>(g, xl, yl, zl, comp, s);
          tested++;
          // the kernels' test: first cell in the box, the (i+2, j+2) corner of its first plane too
          if (box_at(b, s.i0, s.j0, s.k0) < 0 || box_at(b, s.i0 + 2, s.j0 + 2, s.k0) < 0) {
            if (bad < 10)
              std::printf("miss n=%d comp=%d c=(%.17g %.17g %.17g) l=%d stencil=(%d %d %d) box=(%d+%d %d+%d %d+%d)\n", g.nx,
                          comp, c[0], c[1], c[2], l, s.i0, s.j0, s.k0, b.bi0, b.bx, b.bj0, b.by, b.bk0, b.bz);
            bad++;
          }
        }
      }
    }
  }
  std::printf("tested %ld bad %ld\n", tested, bad);
  return bad != 0;
}
"""


def test_every_stencil_lies_in_its_box(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    src = tmp_path / "boxes.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "boxes"
    # the library is built with -ffp-contract=off; so is this host copy
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        f"-I{ROOT}", str(src), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe), "3000"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-2000:]
    assert "bad 0" in r.stdout


def test_periodic_shift_needs_no_division_inside_the_domain():
    """comp_stencil takes the periodic shift floor((x - dx/2) / L) as 0 for 0 <=
    x - dx/2 < L without dividing: there the correctly rounded quotient stays
    below 1.  Checked at the quotient's worst case (the largest double below
    L), at random points, for random and power-of-two L, and at L itself,
    where the shift is 1 (outside the range, as for negative x - dx/2, the
    division is taken)."""
    import numpy as np
    rng = np.random.default_rng(7)
    lens = np.concatenate([rng.uniform(0.05, 50.0, 4000), 2.0 ** np.arange(-6, 7), [1.0, 0.75, 1.5, 3.0, np.pi]])
    for L in lens:
        below = np.nextafter(L, 0.0)
        a = np.concatenate([[0.0, below, np.nextafter(below, 0.0)], rng.uniform(0.0, L, 64), L * (1.0 - 2.0 ** -np.arange(40, 54))])
        a = a[(a >= 0.0) & (a < L)]
        assert np.all(np.floor(a / L) == 0.0), L
        assert np.floor(L / L) == 1.0
