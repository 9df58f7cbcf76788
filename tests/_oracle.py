"""ctypes access to oracle/liboracle.so (the FP64 direct-sweep oracle) and helpers shared by tests.

Test infrastructure only; see oracle/oracle.c for what it restates from the reference.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF1D_SO = os.path.join(ROOT, "oracle", "_ref", "libref1d.so")
REF2D_SO = os.path.join(ROOT, "oracle", "_ref", "libref2d.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"{ORACLE_SO} missing: run make -C oracle")
        L = ctypes.CDLL(ORACLE_SO)
        D = ctypes.POINTER(ctypes.c_double)
        L.orc_sweep.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.c_int, D, ctypes.c_int,
                                ctypes.c_int64, D, D, ctypes.c_int]
        L.orc_sweep.restype = ctypes.c_int
        L.orc_sweep_fast.argtypes = L.orc_sweep.argtypes
        L.orc_sweep_fast.restype = ctypes.c_int
        L.orc_fill_uniform.argtypes = [ctypes.c_uint64, D, ctypes.c_int64]
        L.orc_fill_uniform.restype = None
        L.orc_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def sweep(x, radius, weights, steps, convention=0, threads=0):
    """FP64 direct periodic sweep of `steps` steps on grid x (any of 1..3 dims)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).ravel())
    assert w.size == (2 * radius + 1) ** x.ndim
    out = np.empty_like(x)
    sizes = (ctypes.c_int64 * 3)(*(list(x.shape) + [1] * (3 - x.ndim)))
    rc = lib().orc_sweep(x.ndim, sizes, radius, _p(w), convention, steps, _p(x), _p(out), threads)
    assert rc == 0
    return out


def sweep_fast(x, radius, weights, steps, convention=0, threads=0):
    """The same sweep evaluated for speed (oracle/sweep_fast.c: register-blocked FMAs); the CPU baseline."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    w = np.ascontiguousarray(np.asarray(weights, dtype=np.float64).ravel())
    assert w.size == (2 * radius + 1) ** x.ndim
    out = np.empty_like(x)
    sizes = (ctypes.c_int64 * 3)(*(list(x.shape) + [1] * (3 - x.ndim)))
    rc = lib().orc_sweep_fast(x.ndim, sizes, radius, _p(w), convention, steps, _p(x), _p(out), threads)
    assert rc == 0
    return out


def uniform(seed, shape):
    out = np.empty(int(np.prod(shape)), dtype=np.float64)
    lib().orc_fill_uniform(seed, _p(out), out.size)
    return out.reshape(shape)


def errors(got, ref):
    """(max-abs, max-rel = max|d|/max|ref|, rel-L2)."""
    d = np.abs(np.asarray(got) - np.asarray(ref))
    mref = np.max(np.abs(ref))
    return float(d.max()), float(d.max() / mref), float(np.linalg.norm(d) / np.linalg.norm(ref))


def ref_lib(which):
    path = REF1D_SO if which == 1 else REF2D_SO
    if not os.path.exists(path):
        return None
    L = ctypes.CDLL(path)
    D = ctypes.POINTER(ctypes.c_double)
    for name in ("ref_stencil1d", "ref_step1d_driver", "ref_step2d_driver", "ref_stencil2d_raw"):
        if hasattr(L, name):
            fn = getattr(L, name)
            fn.argtypes = [D, ctypes.c_int, D, ctypes.c_int, D]
            fn.restype = ctypes.c_int
    return L
