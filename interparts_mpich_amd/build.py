"""Build steps: libipg.so (HIP, gfx950), the Fortran host driver, and the
test-side oracle libraries.

    python -m interparts_mpich_amd.build
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
HIP_SOURCES = ["ipg_api.hip", "ipg_fields.hip", "ipg_poisson.hip", "ipg_ibm.hip", "ipg_particles.hip", "ipg_ensemble.hip"]
LIB = os.path.join(HERE, "libipg.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-shared", "-std=c++17",
             # FP64 parity: no contraction of a*b+c into fma, the reference's
             # rounding sequence is kept
             "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "unknown target feature" not in l and "packed-fp32" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def build_hip():
    """Each translation unit compiled on its own (in parallel), then linked."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    cflags = [f for f in HIP_FLAGS if f != "-shared"]
    jobs = []
    for s in HIP_SOURCES:
        obj = os.path.join(objdir, s.replace(".hip", ".o"))
        jobs.append(([HIPCC] + cflags + ["-c", os.path.join(CSRC, s), "-o", obj], obj))
    with ThreadPoolExecutor(max_workers=min(len(jobs), os.cpu_count() or 1)) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _, o in jobs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """The Fortran host driver (fortran/) against libipg.so, with amdflang."""
    fdir = os.path.join(HERE, "fortran")
    if not os.path.exists(os.path.join(fdir, "Makefile")):
        return None
    _run(["make", "-s", "-B", "-C", fdir])
    return os.path.join(fdir, "interpt")


def build_oracle():
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-B", "-C", odir, "liboracle.so"])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-B", "-C", odir, "ref"])


def build_all():
    build_hip()
    build_fortran()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)
