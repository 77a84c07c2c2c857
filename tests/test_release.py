"""A context's release frees each device allocation once (a host build of
ipg_ctx_allocations, interparts_mpich_amd/csrc/ipg_internal.h, the list
release() in ipg_api.hip frees).  Regression test for a release that freed the
particle arrays twice, which an ipg_init failure reached: every failure path
of ipg_init calls release() on the partly built context."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

DRIVER = r"""
#include "interparts_mpich_amd/csrc/ipg_internal.h"
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

int main() {
  static ipg_ctx c;
  std::memset((void *)&c, 0, sizeof c);
  int bad = 0;
  void *out[IPG_MAX_ALLOCS];
  // an empty context (ipg_init failing at its first allocation) frees nothing
  if (ipg_ctx_allocations(&c, out) != 0) { std::printf("empty context lists allocations\n"); bad++; }
  // every allocation ipg_init and ipg_ibm_setup make, each a distinct address
  uintptr_t next = 0x1000;
  auto give = [&](auto *&p) { p = reinterpret_cast<std::remove_reference_t<decltype(p)>>(next); next += 0x100; };
  give(c.slab); give(c.mk); give(c.fl); give(c.fl2); give(c.nbl); give(c.nbc); give(c.nml); give(c.nmc);
  give(c.ibm_err); give(c.ibm_flg); give(c.ibm_psum); give(c.op); give(c.motion2); give(c.red); give(c.dscal); give(c.ps.nbf);
  for (auto &k : c.K1) give(k);
  give(c.pbuf); give(c.stamps);
  for (auto &q : c.ps.q) give(q);
  const int want = (int)((next - 0x1000) / 0x100);
  const int n = ipg_ctx_allocations(&c, out);
  std::vector<void *> v(out, out + n);
  std::sort(v.begin(), v.end());
  if (std::adjacent_find(v.begin(), v.end()) != v.end()) { std::printf("an allocation is freed twice\n"); bad++; }
  if (n != want) { std::printf("%d of %d allocations listed\n", n, want); bad++; }
  for (void *p : v)
    if ((uintptr_t)p < 0x1000 || (uintptr_t)p >= next) { std::printf("a pointer that was never allocated\n"); bad++; }
  // a context whose particle arrays alone were allocated (ipg_init failing just after them)
  std::memset((void *)&c, 0, sizeof c);
  next = 0x1000;
  for (auto &q : c.ps.q) give(q);
  const int m = ipg_ctx_allocations(&c, out);
  if (m != IPG_NPART) { std::printf("particle arrays listed %d times for %d\n", m, IPG_NPART); bad++; }
  std::printf("listed %d of %d, bad %d\n", n, want, bad);
  return bad != 0;
}
"""


def test_release_frees_each_allocation_once(tmp_path):
    if shutil.which("g++") is None or not os.path.isdir("/opt/rocm/include"):
        pytest.skip("needs g++ and the HIP headers")
    src = tmp_path / "release.cpp"
    src.write_text(DRIVER)
    exe = tmp_path / "release"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{ROOT}", str(src),
                        "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
