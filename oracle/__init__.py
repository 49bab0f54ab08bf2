"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (oracle/liboracle.so).

Used by tests/, by __graft_entry__.smoke() as the checker, and by the CPU
baseline leg of bench.py.  The HMVM library never imports this module.

The oracle restates the reference HMVM (src_hmvm/src_c/hmvm_seq.c and
hmvm_omp.c) with its loop and accumulation order; see hmvm_oracle.h for what
pins each function.  When the reference sources are present (this container
only), ``ref()`` loads the reference itself compiled by oracle/Makefile into
oracle/_ref/ for cross-checks.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(_HERE, "liboracle.so")
REF_LIB = os.path.join(_HERE, "_ref", "libhacapk_ref.so")
REF_CPP_LIB = os.path.join(_HERE, "_ref", "libhacapk_ref_cpp.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_float_p = ctypes.POINTER(ctypes.c_float)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            raise ImportError("oracle/liboracle.so missing: run `make -C oracle`")
        L = ctypes.CDLL(ORACLE_LIB)
        for name in ("oracle_hmvm_seq_1", "oracle_hmvm_seq_1t", "oracle_hmvm_omp_1", "oracle_hmvm_omp_1t"):
            getattr(L, name).argtypes = [c_double_p, ctypes.c_void_p, c_double_p]
            getattr(L, name).restype = None
        for name in ("oracle_hmvm_seq_2", "oracle_hmvm_seq_2t", "oracle_hmvm_omp_2", "oracle_hmvm_omp_2t",
                     "oracle_hmvm_seq_2_64", "oracle_hmvm_omp_2_64"):
            getattr(L, name).argtypes = [c_double_p, ctypes.c_void_p, c_double_p]
            getattr(L, name).restype = None
        L.oracle_hmvm_seq_2_f32.argtypes = [c_float_p, ctypes.c_void_p, c_float_p, c_float_p]
        L.oracle_hmvm_seq_2_f32.restype = None
        L.oracle_convert_d2f.argtypes = [c_double_p, c_float_p, ctypes.c_longlong]
        L.oracle_convert_d2f.restype = None
        L.oracle_omp_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(c_double_p)


def _run(name, h, b, use_mat2, v=None):
    n = h.nd
    if v is None:
        v = np.zeros(n)
    m = ctypes.byref(h.mat2) if use_mat2 else ctypes.byref(h.mat)
    getattr(lib(), name)(_p(v), ctypes.cast(m, ctypes.c_void_p), _p(np.ascontiguousarray(b, dtype=np.float64)))
    return v


def seq_1(h, b, v=None):
    """v += A b  (hmvm_seq.c:9-58)"""
    return _run("oracle_hmvm_seq_1", h, b, False, v)


def seq_1t(h, b, v=None):
    return _run("oracle_hmvm_seq_1t", h, b, False, v)


def seq_2(h, b, v=None):
    """v += A b over rowmat (hmvm_seq.c:113-166): the FP64 oracle."""
    return _run("oracle_hmvm_seq_2", h, b, True, v)


def seq_2t(h, b, v=None):
    return _run("oracle_hmvm_seq_2t", h, b, True, v)


def omp_1(h, b):
    """zau = A zu with OpenMP (hmvm_omp.c:9-75); zeroes its output."""
    return _run("oracle_hmvm_omp_1", h, b, False)


def omp_1t(h, b):
    return _run("oracle_hmvm_omp_1t", h, b, False)


def omp_2(h, b):
    return _run("oracle_hmvm_omp_2", h, b, True)


def omp_2t(h, b):
    return _run("oracle_hmvm_omp_2t", h, b, True)


def seq_2_64(h, b, v=None):
    """oracle_hmvm_seq_2 over an HMatrix64 (int64 offsets)."""
    return _run("oracle_hmvm_seq_2_64", h, b, True, v)


def omp_2_64(h, b):
    return _run("oracle_hmvm_omp_2_64", h, b, True)


def seq_2_f32(h, b32, rowmat32):
    """FP32 HMVM in the loop order of src_cpp/hmvm_seq.cpp:154-229 (a2i=0)."""
    v = np.zeros(h.nd, dtype=np.float32)
    lib().oracle_hmvm_seq_2_f32(v.ctypes.data_as(c_float_p), ctypes.cast(ctypes.byref(h.mat2), ctypes.c_void_p),
                                rowmat32.ctypes.data_as(c_float_p), b32.ctypes.data_as(c_float_p))
    return v


def convert_d2f(a):
    """Element-wise FP64 -> FP32 (round to nearest), as convertD2F of
    src_cpp/loadmatrix.cpp:252-342 converts every factor."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    out = np.empty(a.shape, dtype=np.float32)
    lib().oracle_convert_d2f(_p(a), out.ctypes.data_as(c_float_p), a.size)
    return out


def omp_threads():
    return lib().oracle_omp_threads()


# ---- the reference itself, compiled from /root/reference (this container only) ----------
def ref_available():
    return os.path.exists(REF_LIB)


_ref = None


def ref():
    """The reference C HMVM (src_c) built by `make -C oracle ref`; None when absent."""
    global _ref
    if _ref is None and ref_available():
        L = ctypes.CDLL(REF_LIB)
        for name in ("hmvm_seq_1", "hmvm_seq_1t", "hmvm_seq_2", "hmvm_seq_2t",
                     "hmvm_omp_1", "hmvm_omp_1t", "hmvm_omp_2", "hmvm_omp_2t"):
            getattr(L, name).argtypes = [c_double_p, ctypes.c_void_p, c_double_p]
            getattr(L, name).restype = None
        L.loadHmatrix.argtypes = [ctypes.c_char_p, ctypes.c_void_p, ctypes.c_void_p]
        L.loadHmatrix.restype = ctypes.c_int
        L.loadHmatrix2.argtypes = [ctypes.c_char_p, ctypes.c_void_p]
        L.loadHmatrix2.restype = ctypes.c_int
        L.dummymatrix.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.dummymatrix.restype = ctypes.c_int
        _ref = L
    return _ref


_ref_cpp = None


def ref_cpp():
    """The reference C++ templates (src_cpp/hmvm_seq.cpp, loadmatrix.cpp); None when absent."""
    global _ref_cpp
    if _ref_cpp is None and os.path.exists(REF_CPP_LIB):
        _ref_cpp = ctypes.CDLL(REF_CPP_LIB)
    return _ref_cpp
