"""The CPU backend (CPUImpl, the C++ API built with REAPERS_NOGPU) through
libreapers_cpu.so's C functions (include/reapers_cpu.h).

This is a separate backend that a caller selects, as the reference's
REAPERS_NOGPU switch does (inc/reapers.h:96-100); the GPU entry points of
reapers_amd never fall back to it.  Vectors are numpy complex64 (fp=32) or
complex128 (fp=64) arrays of 2^len amplitudes.
"""
import ctypes as _c
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CPU_LIB = os.path.join(HERE, "libreapers_cpu.so")

_vp = _c.c_void_p
_u64 = _c.c_uint64
_dbl = _c.c_double
_lib = None


class CpuError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(CPU_LIB):
            raise RuntimeError("libreapers_cpu.so is not built: run __graft_entry__.build()")
        L = _c.CDLL(CPU_LIB)
        L.rpc_last_error.restype = _c.c_char_p
        sig = {
            "rpc_apply": [_c.c_int, _c.c_int, _vp, _vp, _u64, _vp, _vp, _c.c_int],
            "rpc_evolve": [_c.c_int, _c.c_int, _vp, _vp, _u64, _vp, _dbl, _dbl, _c.c_int, _dbl,
                           _c.POINTER(_c.c_int64)],
            "rpc_ed_evolve": [_c.c_int, _c.c_int, _vp, _vp, _u64, _vp, _dbl, _dbl],
            "rpc_ground_state": [_c.c_int, _c.c_int, _vp, _vp, _u64, _vp, _c.c_int, _dbl,
                                 _c.POINTER(_dbl)],
            "rpc_eigensystem": [_c.c_int, _vp, _vp, _u64, _vp, _vp],
            "rpc_random_state": [_c.c_int, _c.c_int, _u64, _vp],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = _c.c_int
        _lib = L
    return _lib


def _check(st):
    if st != 0:
        raise CpuError("%d: %s" % (st, lib().rpc_last_error().decode(errors="replace")))


def _terms(bits, coeff):
    bits = np.ascontiguousarray(bits, np.uint64)
    c = np.asarray(coeff, np.complex128)
    co = np.ascontiguousarray(np.stack([c.real, c.imag], 1).ravel())
    return bits, co


def _dtype(fp):
    if fp not in (32, 64):
        raise ValueError("fp must be 32 or 64")
    return np.complex64 if fp == 32 else np.complex128


def _p(a):
    return a.ctypes.data_as(_vp)


def apply(length, bits, coeff, psi, fp=64, out=None):
    """out (+)= H psi (out=None: H psi)."""
    dt = _dtype(fp)
    b, co = _terms(bits, coeff)
    vin = np.ascontiguousarray(psi, dt)
    acc = out is not None
    vout = np.ascontiguousarray(out, dt).copy() if acc else np.empty_like(vin)
    _check(lib().rpc_apply(fp, length, _p(b), _p(co), len(b), _p(vin), _p(vout), int(acc)))
    return vout


def evolve(length, bits, coeff, psi, t, beta=0.0, krydim=5, tol=0.0, fp=64):
    """exp(-(beta + i t) H) psi by Krylov<CPUImpl>; returns (state, matvecs)."""
    dt = _dtype(fp)
    b, co = _terms(bits, coeff)
    v = np.ascontiguousarray(psi, dt).copy()
    mv = _c.c_int64()
    _check(lib().rpc_evolve(fp, length, _p(b), _p(co), len(b), _p(v), t, beta, krydim, tol,
                            _c.byref(mv)))
    return v, mv.value


def ed_evolve(length, bits, coeff, psi, t, beta=0.0, fp=64):
    dt = _dtype(fp)
    b, co = _terms(bits, coeff)
    v = np.ascontiguousarray(psi, dt).copy()
    _check(lib().rpc_ed_evolve(fp, length, _p(b), _p(co), len(b), _p(v), t, beta))
    return v


def ground_state(length, bits, coeff, psi, krydim=5, eps=1e-10, fp=64):
    dt = _dtype(fp)
    b, co = _terms(bits, coeff)
    v = np.ascontiguousarray(psi, dt).copy()
    e = _dbl()
    _check(lib().rpc_ground_state(fp, length, _p(b), _p(co), len(b), _p(v), krydim, eps,
                                  _c.byref(e)))
    return e.value, v


def eigensystem(length, bits, coeff):
    b, co = _terms(bits, coeff)
    n = 1 << length
    w = np.empty(n)
    V = np.empty(n * n, np.complex128)
    _check(lib().rpc_eigensystem(length, _p(b), _p(co), len(b), _p(w), _p(V)))
    return w, V.reshape(n, n).T.copy()


def random_state(length, seed, fp=64):
    v = np.empty(1 << length, _dtype(fp))
    _check(lib().rpc_random_state(fp, length, seed, _p(v)))
    return v
