"""ctypes binding of the CPU oracle (oracle/oracle.c).  TEST INFRASTRUCTURE:
imported only by tests/, the smoke check and the benchmark's CPU-baseline
leg, as the checker -- never by the product path."""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
FORWARD, INVERSE = 1, 0

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = C.CDLL(LIB)
        D = C.POINTER(C.c_double)
        I64 = C.c_int64
        lib.oracle_dft1d.argtypes = [D, D, I64, C.c_int]
        lib.oracle_dftnd.argtypes = [D, D, I64, I64, I64, C.c_int]
        lib.oracle_dft_bins.argtypes = [D, I64, I64, I64, C.POINTER(I64), I64, C.c_int, D]
        for n in ("refcpu_fft1d", "refcpu_fft2d", "refcpu_fft3d"):
            getattr(lib, n).restype = C.c_int
        lib.refcpu_fft1d.argtypes = [D, I64, C.c_int]
        lib.refcpu_fft2d.argtypes = [D, I64, I64, C.c_int]
        lib.refcpu_fft3d.argtypes = [D, I64, I64, I64, C.c_int]
        lib.refcpu_radix.argtypes = [I64]
        lib.refcpu_radix.restype = C.c_int
        lib.oracle_threads.restype = C.c_int
        _lib = lib
    return _lib


def _c(x):
    x = np.ascontiguousarray(x, dtype=np.complex128).reshape(-1)
    return x


def _p(a):
    return a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _dims3(dims):
    d = list(dims) + [1] * (3 - len(dims))
    return d[0], d[1], d[2]


def dft(x, dims, direction=FORWARD):
    """Naive separable DFT with exact phase reduction (long double)."""
    x = _c(x)
    out = np.empty_like(x)
    nx, ny, nz = _dims3(dims)
    if ny == 1 and nz == 1:
        load().oracle_dft1d(_p(x), _p(out), nx, direction)
    else:
        load().oracle_dftnd(_p(x), _p(out), nx, ny, nz, direction)
    return out


def dft_bins(x, dims, bins, direction=FORWARD):
    """Direct DFT at flat output indices `bins`."""
    x = _c(x)
    bins = np.ascontiguousarray(bins, dtype=np.int64)
    out = np.empty(len(bins), dtype=np.complex128)
    nx, ny, nz = _dims3(dims)
    load().oracle_dft_bins(_p(x), nx, ny, nz, bins.ctypes.data_as(C.POINTER(C.c_int64)), len(bins),
                           direction, _p(out))
    return out


def refcpu(x, dims, direction=FORWARD, inplace=False):
    """CPU restatement of the reference algorithm (in place on a copy)."""
    y = _c(x) if inplace else _c(x).copy()
    nx, ny, nz = _dims3(dims)
    lib = load()
    if len(dims) == 1:
        rc = lib.refcpu_fft1d(_p(y), nx, direction)
    elif len(dims) == 2:
        rc = lib.refcpu_fft2d(_p(y), nx, ny, direction)
    else:
        rc = lib.refcpu_fft3d(_p(y), nx, ny, nz, direction)
    if rc:
        raise ValueError(f"refcpu: unsupported size {dims}")
    return y


def radix(n):
    return load().refcpu_radix(n)


def threads():
    return load().oracle_threads()
