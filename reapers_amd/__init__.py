"""reapers_amd: MI355X backend for the REAPERS Pauli-sum apply and the
restarted Krylov exp(-(it+beta)H)psi built on it.

The native parts are
  libreapers_hip.so   gfx950 HIP kernels behind the C ABI in include/reapers_hip.h
  libreapers_host.so  the header-only C++ API (reapers_amd/include/reapers/)
                      instantiated for ctypes callers
This module only loads them and wraps a handful of calls for Python tests and
the benchmark.  There is no Python or CPU fallback: if a library is missing
the loaders raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(HERE, "libreapers_hip.so")
HOST_LIB = os.path.join(HERE, "libreapers_host.so")

RP_C64, RP_C128 = 0, 1
RP_ROUND_NEAREST, RP_ROUND_TOWARD_ZERO = 0, 1
RP_H2D, RP_D2H, RP_D2D = 0, 1, 2
RP_KERNEL_APPLY, RP_KERNEL_GS, RP_KERNEL_OTHER = 0, 1, 2

_c = ctypes
_vp = _c.c_void_p
_u64 = _c.c_uint64
_i64 = _c.c_int64
_dbl = _c.c_double


class RpError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("status %d: %s" % (code, msg))
        self.code = code


class TermsInfo(_c.Structure):
    _fields_ = [("prec", _c.c_int), ("len", _c.c_int), ("tile_bits", _c.c_int),
                ("nterms", _u64), ("nflips", _u64), ("ngroups", _u64),
                ("bytes_per_apply", _u64)]


class EvolveStats(_c.Structure):
    _fields_ = [("matvecs", _i64), ("restarts", _i64), ("gs_passes", _i64),
                ("breakdown", _c.c_int32), ("substeps", _c.c_int32),
                ("last_nrm", _dbl), ("host_ms", _dbl), ("wall_ms", _dbl)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_hip = None
_host = None


def hip():
    """The C ABI library (libreapers_hip.so)."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise RuntimeError("libreapers_hip.so is not built: run __graft_entry__.build()")
        lib = _c.CDLL(HIP_LIB)
        lib.rp_last_error.restype = _c.c_char_p
        sig = {
            "rp_init": [_c.c_int], "rp_shutdown": [], "rp_synchronize": [],
            "rp_malloc": [_c.POINTER(_vp), _c.c_size_t], "rp_free": [_vp],
            "rp_memcpy": [_vp, _vp, _c.c_size_t, _c.c_int],
            "rp_zero_vec": [_c.c_int, _vp, _u64],
            "rp_copy_vec": [_c.c_int, _vp, _c.c_int, _vp, _u64, _c.c_int],
            "rp_add_vec": [_c.c_int, _vp, _vp, _u64],
            "rp_add_vec3": [_c.c_int, _vp, _vp, _vp, _u64],
            "rp_axpy": [_c.c_int, _vp, _dbl, _dbl, _vp, _u64],
            "rp_scale_vec": [_c.c_int, _vp, _dbl, _dbl, _u64],
            "rp_vec_prod": [_c.c_int, _vp, _vp, _u64, _c.POINTER(_dbl), _c.POINTER(_dbl)],
            "rp_vec_norm": [_c.c_int, _vp, _u64, _c.POINTER(_dbl)],
            "rp_init_random": [_c.c_int, _vp, _u64, _u64],
            "rp_terms_upload": [_c.POINTER(_vp), _c.c_int, _c.c_int, _vp, _vp, _u64],
            "rp_terms_free": [_vp],
            "rp_terms_info": [_vp, _c.POINTER(TermsInfo)],
            "rp_apply_ops": [_vp, _vp, _vp, _c.c_int],
            "rp_gs_pass": [_c.c_int, _vp, _c.c_int, _vp, _u64, _vp, _c.POINTER(_dbl)],
            "rp_lincomb": [_c.c_int, _vp, _vp, _c.c_int, _vp, _u64],
            "rp_prof_enable": [_c.c_int], "rp_prof_reset": [],
            "rp_prof_query": [_c.c_int, _c.POINTER(_u64), _c.POINTER(_dbl)],
            "rp_mem_get_info": [_c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)],
            "rp_device_props": [_c.POINTER(_c.c_int), _c.c_char_p, _c.c_size_t],
        }
        for name, args in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = _c.c_int
        _hip = lib
    return _hip


def host():
    """The C++ API library (libreapers_host.so)."""
    global _host
    if _host is None:
        if not os.path.exists(HOST_LIB):
            raise RuntimeError("libreapers_host.so is not built: run __graft_entry__.build()")
        hip()
        lib = _c.CDLL(HOST_LIB)
        lib.rph_last_error.restype = _c.c_char_p
        sig = {
            "rph_syk_gen": ([_c.c_int, _c.c_int, _dbl, _dbl, _c.c_int, _c.c_uint32, _c.c_int,
                             _vp, _vp, _i64, _c.POINTER(_i64)], _c.c_int),
            "rph_majorana": ([_c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp, _vp], _c.c_int),
            "rph_ops_mul": ([_c.c_int, _vp, _vp, _i64, _vp, _vp, _i64, _vp, _vp, _i64,
                             _c.POINTER(_i64)], _c.c_int),
            "rph_random_stream": ([_c.c_int, _c.c_int, _c.c_uint32, _dbl, _dbl, _vp, _i64],
                                  _c.c_int),
            "rph_init": ([_c.c_int], _c.c_int),
            "rph_ham_new": ([_c.c_int, _c.c_int, _vp, _vp, _i64], _vp),
            "rph_ham_info": ([_vp, _c.POINTER(TermsInfo)], _c.c_int),
            "rph_ham_free": ([_vp], None),
            "rph_state_new": ([_c.c_int, _c.c_int], _vp),
            "rph_state_free": ([_vp], None),
            "rph_state_load": ([_vp, _vp], _c.c_int),
            "rph_state_store": ([_vp, _vp], _c.c_int),
            "rph_state_random": ([_vp, _u64], _c.c_int),
            "rph_state_copy": ([_vp, _vp, _c.c_int], _c.c_int),
            "rph_state_apply": ([_vp, _vp], _c.c_int),
            "rph_state_evolve": ([_vp, _vp, _dbl, _dbl, _c.c_int, _dbl, _c.c_int, _c.c_int,
                                  _dbl, _c.POINTER(EvolveStats)], _c.c_int),
            "rph_state_inner": ([_vp, _vp, _c.POINTER(_dbl), _c.POINTER(_dbl)], _c.c_int),
            "rph_state_norm": ([_vp, _c.POINTER(_dbl)], _c.c_int),
            "rph_state_ground": ([_vp, _vp, _c.c_int, _dbl, _c.c_int, _c.POINTER(_dbl)], _c.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(lib, name)
            f.argtypes = args
            f.restype = res
        _host = lib
    return _host


def check(status):
    if status != 0:
        raise RpError(status, hip().rp_last_error().decode(errors="replace"))


def hcheck(status):
    if status != 0:
        raise RpError(status, host().rph_last_error().decode(errors="replace"))


def _ptr(a):
    return a.ctypes.data_as(_vp)


# ---------------------------------------------------------------------------
# host-side term generation through the C++ API (no GPU needed)
# ---------------------------------------------------------------------------
def syk_terms(N, sp=0.0, J=1.0, seed=0, fp=32, std_gamma=True, regularize=False):
    """Term list of SYK(N, sp, J).gen_ham(std::mt19937(seed)): (bits u64, coeff
    complex128 holding the fp-precision values exactly)."""
    cap = 1 << 12
    while True:
        bits = np.zeros(cap, np.uint64)
        co = np.zeros(2 * cap, np.float64)
        n = _i64()
        hcheck(host().rph_syk_gen(fp, N, sp, J, int(std_gamma), seed, int(regularize),
                                  _ptr(bits), _ptr(co), cap, _c.byref(n)))
        if n.value >= 0:
            return bits[:n.value].copy(), (co[0:2 * n.value:2] + 1j * co[1:2 * n.value:2])
        cap = -n.value


def majorana(N, n, fp=32, std_gamma=True):
    bits = np.zeros(1, np.uint64)
    co = np.zeros(2, np.float64)
    hcheck(host().rph_majorana(fp, N, n, int(std_gamma), _ptr(bits), _ptr(co)))
    return int(bits[0]), complex(co[0], co[1])


def random_stream(kind, n, seed, fp=64, a=0.0, b=0.0):
    """n draws from std::mt19937(seed) through libstdc++ <random>:
    kind 'normal' (normal_distribution<fp>(0,1)), 'bernoulli' (p=a),
    'uniform_int' (a..b) or 'canonical' (generate_canonical<fp>)."""
    k = {"normal": 0, "bernoulli": 1, "uniform_int": 2, "canonical": 3}[kind]
    out = np.zeros(n)
    hcheck(host().rph_random_stream(k, fp, seed, a, b, _ptr(out), n))
    return out


def ops_mul(a, b, fp=64):
    """Product of two term lists (bits, complex coeff) via HostSumOps::operator*."""
    (b0, c0), (b1, c1) = a, b
    b0 = np.ascontiguousarray(b0, np.uint64)
    b1 = np.ascontiguousarray(b1, np.uint64)
    cc0 = np.empty(2 * len(b0)); cc0[0::2] = np.real(c0); cc0[1::2] = np.imag(c0)
    cc1 = np.empty(2 * len(b1)); cc1[0::2] = np.real(c1); cc1[1::2] = np.imag(c1)
    cap = len(b0) * len(b1) + 1
    bo = np.zeros(cap, np.uint64)
    co = np.zeros(2 * cap)
    n = _i64()
    hcheck(host().rph_ops_mul(fp, _ptr(b0), _ptr(cc0), len(b0), _ptr(b1), _ptr(cc1), len(b1),
                              _ptr(bo), _ptr(co), cap, _c.byref(n)))
    return bo[:n.value].copy(), co[0:2 * n.value:2] + 1j * co[1:2 * n.value:2]


# ---------------------------------------------------------------------------
# device objects (GPU required)
# ---------------------------------------------------------------------------
def init(device=0):
    check(hip().rp_init(device))


class Hamiltonian:
    """A REAPERS SumOps<Fp> with its device layout (C++ API handle)."""

    def __init__(self, length, bits, coeff, fp=32):
        self.len = int(length)
        self.fp = fp
        bits = np.ascontiguousarray(bits, np.uint64)
        co = np.empty(2 * len(bits))
        co[0::2] = np.real(coeff)
        co[1::2] = np.imag(coeff)
        self._h = host().rph_ham_new(fp, self.len, _ptr(bits), _ptr(co), len(bits))
        if not self._h:
            raise RpError(-1, host().rph_last_error().decode())

    def info(self):
        i = TermsInfo()
        hcheck(host().rph_ham_info(self._h, _c.byref(i)))
        return i

    def close(self):
        if self._h:
            host().rph_ham_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceState:
    """A REAPERS State<Fp, GPUImpl> (C++ API handle)."""

    def __init__(self, length, fp=32):
        self.len = int(length)
        self.fp = fp
        self.dim = 1 << self.len
        self._s = host().rph_state_new(fp, self.len)
        if not self._s:
            raise RpError(-1, host().rph_last_error().decode())

    @property
    def dtype(self):
        return np.complex64 if self.fp == 32 else np.complex128

    def load(self, psi):
        a = np.ascontiguousarray(psi, dtype=self.dtype)
        assert a.shape == (self.dim,)
        hcheck(host().rph_state_load(self._s, _ptr(a)))

    def store(self):
        a = np.empty(self.dim, dtype=self.dtype)
        hcheck(host().rph_state_store(self._s, _ptr(a)))
        return a

    def random(self, seed):
        hcheck(host().rph_state_random(self._s, seed))

    def copy_from(self, other, rounding=RP_ROUND_NEAREST):
        hcheck(host().rph_state_copy(self._s, other._s, rounding))

    def apply(self, ham):
        hcheck(host().rph_state_apply(self._s, ham._h))

    def evolve(self, ham, t, beta=0.0, krydim=5, tol=0.0, max_iters=0, reorth=0, substep=0.0):
        st = EvolveStats()
        hcheck(host().rph_state_evolve(self._s, ham._h, t, beta, krydim, tol, max_iters, reorth,
                                       substep, _c.byref(st)))
        return st

    def inner(self, other):
        re, im = _dbl(), _dbl()
        hcheck(host().rph_state_inner(self._s, other._s, _c.byref(re), _c.byref(im)))
        return complex(re.value, im.value)

    def norm(self):
        n = _dbl()
        hcheck(host().rph_state_norm(self._s, _c.byref(n)))
        return n.value

    def ground_state(self, ham, krydim=5, eps=1e-10, maxiter=0):
        e = _dbl()
        hcheck(host().rph_state_ground(self._s, ham._h, krydim, eps, maxiter, _c.byref(e)))
        return e.value

    def close(self):
        if self._s:
            host().rph_state_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
