"""The IBM kernels build without scratch (no register spills) and keep the
register count that lets two 512-thread workgroups share a CU, as reported by
the compiler for gfx950 (-Rpass-analysis=kernel-resource-usage) on
interparts_mpich_amd/csrc/ipg_ibm.hip with the library's flags; the sphere
integrals (ipg_particles.hip) likewise, with three 512-thread blocks per CU.

Both IBM anomalies on record (a NaN in one GPU run, a fault of a 384-thread
variant) came from builds of k_ibm that spilled to scratch; builds that
spilled as much ran cleanly, so spilling is not shown to be their cause, but
the library has had no spilling IBM build since and this test keeps it so."""
import os
import re
import subprocess

import pytest

from interparts_mpich_amd import build

SRC = os.path.join(build.CSRC, "ipg_ibm.hip")


def _usage(tmp_path, src=SRC):
    flags = [f for f in build.HIP_FLAGS if f != "-shared"]
    r = subprocess.run([build.HIPCC] + flags + ["-c", src, "-o", str(tmp_path / "k.o"),
                                                "-Rpass-analysis=kernel-resource-usage"],
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=900)
    assert r.returncode == 0, r.stdout[-3000:]
    kernels, name = {}, None
    for line in r.stdout.splitlines():
        m = re.search(r"Function Name: (\S+)", line)
        if m:
            name = m.group(1)
            kernels[name] = {}
            continue
        m = re.search(r"remark: +([A-Za-z ]+?)(?: \[[^\]]*\])?: (\d+) \[-Rpass", line)
        if m and name:
            kernels[name][m.group(1).strip()] = int(m.group(2))
    return kernels


def test_ibm_kernels_do_not_spill(tmp_path):
    if not os.path.exists(build.HIPCC):
        pytest.skip("needs hipcc")
    kernels = _usage(tmp_path)
    # the single-context kernels and the merged launches of a batch
    # (k_batch_wpe<IbmK<..>>, k_batch_persist<IbmK<..>>) running the same bodies
    ibm = {k: v for k, v in kernels.items() if "k_ibm" in k or "IbmK" in k}
    assert len(ibm) >= 20, sorted(kernels)
    for k, v in sorted(ibm.items()):
        print(k, v)
        assert v.get("ScratchSize", -1) == 0, f"{k} spills: {v}"
    # the three FP64 passes: at most 128 VGPRs, so that the 16 waves of two
    # 512-thread workgroups (4 per SIMD) fit one CU
    for k, v in ibm.items():
        if re.search(r"k_ibmILi[012]EdE", k) or re.search(r"k_batch_wpeI.*IbmKILi[012]EdE", k):
            assert v["VGPRs"] <= 128, f"{k}: {v}"


def test_sphere_integrals_do_not_spill(tmp_path):
    """k_sphere_int and the batch's k_batch_wpe<SphereIntK<type>>: no scratch,
    at most 80 VGPRs (6 wavefronts per SIMD: three 512-thread blocks per CU);
    the persistent k_sphere_int_persist, launched with one block per CU: no
    scratch."""
    if not os.path.exists(build.HIPCC):
        pytest.skip("needs hipcc")
    kernels = _usage(tmp_path, os.path.join(build.CSRC, "ipg_particles.hip"))
    sph = {k: v for k, v in kernels.items() if "k_sphere_int" in k or re.search(r"k_batch_wpeI.*SphereIntK", k)}
    assert len(sph) >= 5, sorted(kernels)
    for k, v in sorted(sph.items()):
        print(k, v)
        assert v.get("ScratchSize", -1) == 0, f"{k} spills: {v}"
        if "persist" not in k:
            assert v["VGPRs"] <= 80, f"{k}: {v}"
