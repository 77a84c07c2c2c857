"""The index emulation of tests/test_ibm_index_bounds.py has power: each
plausible slip below, planted in a copy of the IBM index header
(interparts_mpich_amd/csrc/ipg_ibm_geom.h) in a temporary tree, makes the
emulation report a failure on the small cases.  The committed header is
untouched; the copies are built and run only here, on the host."""
import os
import shutil
import subprocess

import pytest

from test_ibm_index_bounds import DRIVER, ROOT, _cases

HDR = os.path.join(ROOT, "interparts_mpich_amd", "csrc", "ipg_ibm_geom.h")

# (what the slip is, text in the header, its replacement)
MUTATIONS = [
    ("box walk carry off by one", "if (li >= bx) { li -= bx; row++; }", "if (li > bx) { li -= bx; row++; }"),
    ("the two cell masks overlap in LDS", "L.taken = L.flag + (size_t)L.nflagmax * sizeof(unsigned);", "L.taken = L.flag;"),
    ("one z plane too many counted inside the box", "nzin = imin(3, b.bk0 + b.bz - k0);",
     "nzin = imin(3, b.bk0 + b.bz - k0 + 1);"),
    ("k1-plane index without the upper wrap", "  r -= r >= g.nx ? g.nx : 0;\n", "\n"),
    ("box rows without the periodic wrap", "g.sx * wrapany(b.bj0 + lj, g.ny)", "g.sx * (b.bj0 + lj)"),
    ("row offsets with the columns' extent", "const int lk = row / b.by, lj = row - lk * b.by;",
     "const int lk = row / b.bx, lj = row - lk * b.bx;"),
    ("a neighbour's image shifted the wrong way", "  c.i0 += sxs * g.nx;\n  c.j0 += sys * g.ny;\n  return !(",
     "  c.i0 -= sxs * g.nx;\n  c.j0 += sys * g.ny;\n  return !("),
]


@pytest.mark.parametrize("what,old,new", MUTATIONS, ids=[m[0] for m in MUTATIONS])
def test_planted_index_slip_is_caught(tmp_path, what, old, new):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    src = open(HDR).read()
    assert src.count(old) == 1, f"the header text for this mutation has changed: {old}"
    csrc = tmp_path / "interparts_mpich_amd" / "csrc"
    csrc.mkdir(parents=True)
    (tmp_path / "include").mkdir()
    for h in ("ipg_internal.h", "ipg_batch.h"):
        shutil.copy(os.path.join(ROOT, "interparts_mpich_amd", "csrc", h), csrc)
    shutil.copy(os.path.join(ROOT, "include", "ipg.h"), tmp_path / "include")
    (csrc / "ipg_ibm_geom.h").write_text(src.replace(old, new))
    (tmp_path / "drv.cpp").write_text(DRIVER)
    exe = tmp_path / "drv"
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                        f"-I{tmp_path}", str(tmp_path / "drv.cpp"), "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    data = tmp_path / "cases.bin"
    # the small cases (the 256^3 one is left to the full check)
    data.write_bytes(b"".join(_cases()[1:]))
    r = subprocess.run([str(exe), str(data), "512", "2"], capture_output=True, text=True, timeout=600)
    tail = [l for l in r.stdout.splitlines() if l.startswith("FAIL") or l.startswith("checks")]
    print(what, "->", tail[:2], tail[-1:])
    assert r.returncode != 0 and "failures 0" not in r.stdout, f"not caught: {what}"
