"""Thin helpers over the C ABI (include/reapers_hip.h) for GPU tests: device
buffers and direct calls of rp_* entry points."""
import ctypes

import numpy as np

import reapers_amd as ra

_vp = ctypes.c_void_p


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _vp()
        ra.check(ra.hip().rp_malloc(ctypes.byref(self.ptr), self.nbytes))

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        ra.check(ra.hip().rp_memcpy(b.ptr, a.ctypes.data_as(_vp), a.nbytes, ra.RP_H2D))
        return b

    def to_array(self, dtype, count):
        a = np.empty(count, dtype=dtype)
        ra.check(ra.hip().rp_memcpy(a.ctypes.data_as(_vp), self.ptr, a.nbytes, ra.RP_D2H))
        return a

    def free(self):
        if self.ptr:
            ra.check(ra.hip().rp_free(self.ptr))
            self.ptr = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def prec_of(dtype):
    return ra.RP_C64 if np.dtype(dtype) == np.complex64 else ra.RP_C128


class Terms:
    """rp_terms_t handle from a (bits, coeff) term list; flags as for
    rp_terms_upload_ex (RP_TERMS_AUTOTUNE, RP_TERMS_ACC_GROUP, ...)."""

    def __init__(self, length, bits, coeff, prec, flags=0):
        bits = np.ascontiguousarray(bits, np.uint64)
        co = np.empty(2 * len(bits))
        co[0::2] = np.real(coeff)
        co[1::2] = np.imag(coeff)
        self.h = _vp()
        args = (ctypes.byref(self.h), prec, length, bits.ctypes.data_as(_vp),
                co.ctypes.data_as(_vp), len(bits))
        if flags:
            ra.check(ra.hip().rp_terms_upload_ex(*args, flags))
        else:
            ra.check(ra.hip().rp_terms_upload(*args))
        self.length = length
        self.prec = prec

    def info(self):
        i = ra.TermsInfo()
        ra.check(ra.hip().rp_terms_info(self.h, ctypes.byref(i)))
        return i

    def free(self):
        if self.h:
            ra.check(ra.hip().rp_terms_free(self.h))
            self.h = _vp()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def gpu_apply(length, bits, coeff, psi, accumulate_into=None):
    """H psi (or accumulate_into + H psi) on the GPU through rp_apply_ops."""
    dtype = psi.dtype
    prec = prec_of(dtype)
    t = Terms(length, bits, coeff, prec)
    vin = DeviceBuffer.from_array(psi)
    if accumulate_into is None:
        vout = DeviceBuffer(psi.nbytes)
        ra.check(ra.hip().rp_apply_ops(t.h, vout.ptr, vin.ptr, 0))
    else:
        vout = DeviceBuffer.from_array(np.ascontiguousarray(accumulate_into, dtype))
        ra.check(ra.hip().rp_apply_ops(t.h, vout.ptr, vin.ptr, 1))
    out = vout.to_array(dtype, len(psi))
    vin.free(); vout.free(); t.free()
    return out
