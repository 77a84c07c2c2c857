"""The oracle's pins have power: each plausible mistake below, planted in a
copy of oracle/ipg_oracle.c built into a temporary library, makes at least one
pin of tests/test_oracle_pins.py, tests/test_oracle_sequence_pins.py or
tests/test_oracle_phase_pins.py fail.
The copies live in a temporary directory and are never loaded by anything
else; the committed oracle is untouched."""
import os
import subprocess
import sys
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "oracle", "ipg_oracle.c")

# (what the slip is, text in the oracle, its replacement)
MUTATIONS = [
    ("floor for nint in the face-direction stencil base (src/interp_spread.f90:98)",
     "t->ilows = nint_f(cx) - 1;", "t->ilows = (int)floor(cx) - 1;"),
    ("no half-cell shift for the cell-centred stencil weights (src/interp_spread.f90:124-126)",
     "t->phiyh[q] = orc_kernel((((double)t->jlow + (double)q) - 0.5) - cy);",
     "t->phiyh[q] = orc_kernel((((double)t->jlow + (double)q) - 0.0) - cy);"),
    ("a wrong RK3 zeta coefficient at stage 3 (src/param.f90:110-112)",
     "f3 = dt * (25.0 / 60.0); }", "f3 = dt * (15.0 / 60.0); }"),
    ("the gather skipping the top-wall k1 plane (src/interp_spread.f90:100)",
     "aw = aw + s->dwdtf[I2(s, t.ilow + i, t.jlow + j, t.klows + k)]",
     "aw = aw + (t.klows + k > s->nz ? 0.0 : 1.0) * s->dwdtf[I2(s, t.ilow + i, t.jlow + j, t.klows + k)]"),
    ("the sign of the z term of fillps (src/fillps.f90:43-46)",
     "s->p[q] = (s->dwdt[q] - s->dwdt[I1(s, i, j, k - 1)]) * dtidzi",
     "s->p[q] = (s->dwdt[q] - s->dwdt[I1(s, i, j, k - 1)]) * (-dtidzi)"),
    ("the no-slip ghost rule as a copy (src/bound.f90:43-51)",
     "if (s->bc_uvw_bot == 0) { u[I1(s, i, j, 0)] = -u[I1(s, i, j, 1)];",
     "if (s->bc_uvw_bot == 0) { u[I1(s, i, j, 0)] = u[I1(s, i, j, 1)];"),
    ("the phase indicator's image in slot nb, not nb + 4 (src/phase_indicator.f90:255-258)",
     "      if (nb > 0) { nbrecv = nb + 4; if (nbrecv > 8) nbrecv -= 8; }\n      double X = s->px[p], Y = s->py[p], Z = s->pz[p];\n      image_shift(s, nbrecv, &X, &Y);\n      double cxc = X - 0.0, cyc = Y - 0.0, czc = Z;\n      int ilow = nint_f((cxc - R) * s->dxi - 2.0), ihigh = nint_f((cxc + R) * s->dxi + 2.0);\n      int jlow = nint_f((cyc - R) * s->dyi - 2.0), jhigh = nint_f((cyc + R) * s->dyi + 2.0);\n      int klow = nint_f((czc - R) * s->dzi - 2.0), khigh = nint_f((czc + R) * s->dzi + 2.0);\n      if (ilow < 1) ilow = 1; if (jlow < 1) jlow = 1; if (klow < 1) klow = 1;\n      if (ihigh > s->nx) ihigh = s->nx; if (jhigh > s->ny) jhigh = s->ny; if (khigh > s->nz) khigh = s->nz;\n      for (int k = klow; k <= khigh; k++) {\n        double zmin = (double)(k - 1) * dz",
     "      (void)0;\n      double X = s->px[p], Y = s->py[p], Z = s->pz[p];\n      image_shift(s, nbrecv, &X, &Y);\n      double cxc = X - 0.0, cyc = Y - 0.0, czc = Z;\n      int ilow = nint_f((cxc - R) * s->dxi - 2.0), ihigh = nint_f((cxc + R) * s->dxi + 2.0);\n      int jlow = nint_f((cyc - R) * s->dyi - 2.0), jhigh = nint_f((cyc + R) * s->dyi + 2.0);\n      int klow = nint_f((czc - R) * s->dzi - 2.0), khigh = nint_f((czc + R) * s->dzi + 2.0);\n      if (ilow < 1) ilow = 1; if (jlow < 1) jlow = 1; if (klow < 1) klow = 1;\n      if (ihigh > s->nx) ihigh = s->nx; if (jhigh > s->ny) jhigh = s->ny; if (khigh > s->nz) khigh = s->nz;\n      for (int k = klow; k <= khigh; k++) {\n        double zmin = (double)(k - 1) * dz"),
    ("the phase indicator's denominator without the absolute values (src/phase_indicator.f90:454-456)",
     "for (int cp = 0; cp < 8; cp++) { if (sg[cp] > 0.0) s1 = s1 + sg[cp]; s2 = s2 + fabs(sg[cp]); }\n              /* gamp",
     "for (int cp = 0; cp < 8; cp++) { if (sg[cp] > 0.0) s1 = s1 + sg[cp]; s2 = s2 + sg[cp]; }\n              /* gamp"),
    ("omega x r taken as r x omega in the marker velocity (src/intgr_nwtn_eulr.f90:1474-1479)",
     "s->ul[m] = s->pu[p] + s->pomy[p] * (zl - s->pz[p]) - s->pomz[p] * (yl - s->py[p]);",
     "s->ul[m] = s->pu[p] - s->pomy[p] * (zl - s->pz[p]) + s->pomz[p] * (yl - s->py[p]);"),
]


@pytest.mark.parametrize("what,old,new", MUTATIONS, ids=[m[0].split(" (")[0] for m in MUTATIONS])
def test_planted_mistake_fails_a_pin(what, old, new):
    src = open(SRC).read()
    assert src.count(old) == 1, f"the oracle text for this mutation has changed: {old}"
    with tempfile.TemporaryDirectory(prefix="ipg_mut_") as d:
        c = os.path.join(d, "ipg_oracle.c")
        so = os.path.join(d, "liboracle.so")
        open(c, "w").write(src.replace(old, new))
        subprocess.run(["gcc", "-O2", "-fPIC", "-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared",
                        "-o", so, c, "-lm"], check=True)
        env = dict(os.environ, IPG_ORACLE_LIB=so)
        r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                            os.path.join(ROOT, "tests", "test_oracle_sequence_pins.py"),
                            os.path.join(ROOT, "tests", "test_oracle_phase_pins.py"),
                            os.path.join(ROOT, "tests", "test_oracle_pins.py")],
                           cwd=ROOT, env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=600)
        tail = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else ""
        print(what, "->", tail)
        assert r.returncode != 0, f"no pin failed for: {what}\n{tail}"
