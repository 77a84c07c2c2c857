"""Writes tests/golden/ref_vectors.json from oracle/_ref/libref.so, the
reference's own src/kernel.f90 and src/libs/fft/fft.f compiled in place with
amdflang -fdefault-real-8 (oracle/Makefile, target `ref`).  The vectors let
the CPU tests pin the oracle against the reference's outputs where the
reference tree (and so oracle/_ref) is absent.

    python tests/golden/make_ref_vectors.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle.pyoracle import ref_lib  # noqa: E402


def vrfft(lib, n, x, forward):
    w = np.zeros(2 * n + 15)
    N, M = C.c_int(n), C.c_int(1)
    lib.vrffti_(C.byref(N), w.ctypes.data_as(C.c_void_p))
    r = np.array(x, dtype=np.float64)
    rt = np.zeros(n)
    fn = lib.vrfftf_ if forward else lib.vrfftb_
    fn(C.byref(M), C.byref(N), r.ctypes.data_as(C.c_void_p), rt.ctypes.data_as(C.c_void_p), C.byref(M),
       w.ctypes.data_as(C.c_void_p))
    return r


def main():
    lib = ref_lib()
    if lib is None:
        raise SystemExit("oracle/_ref/libref.so is not built (make -C oracle ref)")
    rs = np.linspace(-1.75, 1.75, 141)
    kern = [lib._QMmod_kernelPkernel(C.byref(C.c_double(float(r)))) for r in rs]
    s = C.c_double()
    lib._QMmod_kernelPkerneltest(C.byref(s))
    fft = {}
    for n in (8, 12, 16, 24, 64, 96):
        x = np.cos(0.37 * np.arange(n) ** 1.3) + 0.25 * np.sin(1.7 * np.arange(n))
        fft[str(n)] = {"x": x.tolist(), "forward": vrfft(lib, n, x, True).tolist(),
                       "backward": vrfft(lib, n, x, False).tolist()}
    out = {"source": "src/kernel.f90 and src/libs/fft/fft.f compiled with amdflang -O0 -fdefault-real-8",
           "kernel_r": rs.tolist(), "kernel": kern, "kerneltest": s.value, "vfftpack": fft}
    with open(os.path.join(HERE, "ref_vectors.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "ref_vectors.json"))


if __name__ == "__main__":
    main()
