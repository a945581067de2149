"""Build recipe for the CPU baseline (oracle/ampcor_cpu.cpp -> oracle/libampcor_cpu.so).

Test infrastructure: the library restates mroipac/ampcor (ampcor.F) on the CPU
and is loaded only by tests and by the CPU-baseline leg of bench.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

HERE = Path(__file__).resolve().parent
SRC = HERE / "ampcor_cpu.cpp"
LIB = HERE / "libampcor_cpu.so"
# x86-64-v3 (AVX2 + FMA) runs on the GPU hosts as well as on this build host
FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17", "-ffp-contract=fast"]


def build() -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    subprocess.run(["g++", *FLAGS, str(SRC), "-o", str(tmp)], check=True)
    os.replace(tmp, LIB)
    return LIB


def load():
    if not LIB.exists():
        build()
    lib = C.CDLL(str(LIB))
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    lib.ampcor_cpu_run.argtypes = [vp, i64, vp, i64, vp, vp, vp, vp, i32, i32, i32, i32, vp, vp, vp]
    lib.ampcor_cpu_run.restype = C.c_int
    lib.ampcor_cpu_max_threads.restype = C.c_int
    return lib


def run(ref, sec, ref_d, ref_a, sec_d, sec_a, window, hs, threads=0):
    """Run the CPU ampcor restatement on numpy complex64 images; returns (offset, snr, cov, accepted)."""
    import numpy as np
    lib = load()
    ref = np.ascontiguousarray(ref, np.complex64)
    sec = np.ascontiguousarray(sec, np.complex64)
    arrs = [np.ascontiguousarray(a, np.int32) for a in (ref_d, ref_a, sec_d, sec_a)]
    n = len(arrs[0])
    off = np.zeros((n, 2), np.float32)
    snr = np.zeros(n, np.float32)
    cov = np.zeros((n, 3), np.float32)
    p = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    acc = lib.ampcor_cpu_run(p(ref), ref.shape[1], p(sec), sec.shape[1], *map(p, arrs), n, window, hs, threads,
                             p(off), p(snr), p(cov))
    return off, snr, cov, acc


if __name__ == "__main__":
    print(build())
