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
RP_OK, RP_ERR_INVALID_ARGUMENT, RP_ERR_OUT_OF_MEMORY, RP_ERR_RUNTIME, RP_ERR_NO_DEVICE = 0, 1, 2, 3, 4
RP_ROUND_NEAREST, RP_ROUND_TOWARD_ZERO = 0, 1
RP_H2D, RP_D2H, RP_D2D = 0, 1, 2
RP_KERNEL_APPLY, RP_KERNEL_GS, RP_KERNEL_OTHER = 0, 1, 2
RP_APPLY_SMALL, RP_APPLY_TILED, RP_APPLY_STREAM = 0, 1, 2
RP_TERMS_ACC_NATIVE, RP_TERMS_AUTOTUNE, RP_TERMS_ACC_GROUP = 1, 2, 4

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
                ("bytes_per_apply", _u64), ("threads", _c.c_int), ("depth", _c.c_int),
                ("kernel", _c.c_int), ("apply_passes", _dbl), ("tile_map", _c.c_int)]


class EvolveStats(_c.Structure):
    _fields_ = [("matvecs", _i64), ("restarts", _i64), ("gs_passes", _i64),
                ("breakdown", _c.c_int32), ("substeps", _c.c_int32),
                ("last_nrm", _dbl), ("host_ms", _dbl), ("wall_ms", _dbl),
                ("rejected", _c.c_int32), ("reserved", _c.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchInfo(_c.Structure):
    _fields_ = [("prec", _c.c_int), ("nprob", _c.c_int), ("nrhs", _c.c_int), ("kmax", _c.c_int),
                ("tile_bits", _c.c_int), ("threads", _c.c_int), ("grid", _c.c_int),
                ("planes", _c.c_int), ("nbasis", _u64), ("vec_bytes", _u64), ("nterms", _u64),
                ("ngroups", _u64), ("bytes_per_apply", _u64)]


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
            "rp_terms_upload_ex": [_c.POINTER(_vp), _c.c_int, _c.c_int, _vp, _vp, _u64, _c.c_int],
            "rp_terms_free": [_vp],
            "rp_terms_info": [_vp, _c.POINTER(TermsInfo)],
            "rp_terms_set_tiling": [_vp, _c.c_int, _c.c_int, _c.c_int],
            "rp_batch_groups": [_vp, _vp],
            "rp_apply_ops": [_vp, _vp, _vp, _c.c_int],
            "rp_apply_gs": [_vp, _vp, _vp, _vp, _c.c_int, _vp, _c.POINTER(_dbl), _dbl,
                            _c.POINTER(_c.c_int)],
            "rp_gs_pass": [_c.c_int, _vp, _c.c_int, _vp, _u64, _vp, _c.POINTER(_dbl)],
            "rp_gs_pass_fused": [_c.c_int, _vp, _c.c_int, _vp, _u64, _vp, _c.POINTER(_dbl), _dbl,
                                 _c.POINTER(_c.c_int)],
            "rp_gs_pass_fused_noalloc": [_c.c_int, _vp, _c.c_int, _vp, _u64, _vp,
                                         _c.POINTER(_dbl), _dbl, _c.POINTER(_c.c_int)],
            "rp_lincomb": [_c.c_int, _vp, _vp, _c.c_int, _vp, _u64],
            "rp_prof_enable": [_c.c_int], "rp_prof_reset": [],
            "rp_prof_query": [_c.c_int, _c.POINTER(_u64), _c.POINTER(_dbl)],
            "rp_mem_get_info": [_c.POINTER(_c.c_size_t), _c.POINTER(_c.c_size_t)],
            "rp_device_props": [_c.POINTER(_c.c_int), _c.c_char_p, _c.c_size_t],
            "rp_mlp_forward": [_vp, _vp, _i64, _vp],
            "rp_mlp_loss_grad": [_vp, _vp, _vp, _vp, _i64, _vp, _c.POINTER(_dbl)],
            "rp_mlp_adam": [_vp, _vp, _vp, _vp, _i64, _dbl, _dbl, _dbl, _dbl, _dbl, _i64],
            "rp_mlp_infer": [_vp, _vp, _c.c_int, _vp],
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
            "rph_xxz_gen": ([_c.c_int, _c.c_int, _dbl, _dbl, _dbl, _c.c_uint32, _vp, _vp, _i64,
                             _c.POINTER(_i64)], _c.c_int),
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
            "rph_state_evolve_list": ([_vp, _vp, _dbl, _dbl, _vp, _c.c_int, _c.c_int, _dbl,
                                       _c.c_int, _c.c_int, _c.POINTER(EvolveStats)], _c.c_int),
            "rph_sched_list_new": ([_vp, _c.c_int, _c.c_int], _vp),
            "rph_sched_mlp_new": ([_vp, _c.c_int, _c.c_int, _c.c_int], _vp),
            "rph_sched_free": ([_vp], None),
            "rph_sched_set_tracing": ([_vp, _c.c_int], _c.c_int),
            "rph_sched_trace": ([_vp, _vp, _vp, _vp, _c.c_int, _c.POINTER(_c.c_int)], _c.c_int),
            "rph_state_evolve_sched": ([_vp, _vp, _vp, _dbl, _dbl, _c.c_int, _dbl, _c.c_int,
                                        _c.c_int, _c.POINTER(EvolveStats)], _c.c_int),
            "rph_otoc": ([_vp, _vp, _vp, _vp, _dbl, _vp, _c.c_int, _c.c_int, _dbl, _vp,
                          _c.POINTER(EvolveStats), _vp], _c.c_int),
            "rph_batch_new": ([_c.c_int, _c.c_int, _vp, _vp, _vp, _vp, _c.c_int, _c.c_int,
                               _c.c_int], _vp),
            "rph_batch_free": ([_vp], None),
            "rph_batch_info": ([_vp, _c.POINTER(BatchInfo), _vp], _c.c_int),
            "rph_batch_load": ([_vp, _vp], _c.c_int),
            "rph_batch_store": ([_vp, _vp], _c.c_int),
            "rph_batch_apply": ([_vp], _c.c_int),
            "rph_batch_evolve": ([_vp, _dbl, _dbl, _c.c_int, _dbl, _c.c_int, _c.c_int, _vp,
                                  _c.POINTER(EvolveStats), _vp], _c.c_int),
            "rph_set_upload_flags": ([_c.c_int], _c.c_int),
            "rph_state_inner": ([_vp, _vp, _c.POINTER(_dbl), _c.POINTER(_dbl)], _c.c_int),
            "rph_state_norm": ([_vp, _c.POINTER(_dbl)], _c.c_int),
            "rph_state_ground": ([_vp, _vp, _c.c_int, _dbl, _c.c_int, _c.POINTER(_dbl)], _c.c_int),
            "rph_syk_block_gen": ([_c.c_int, _c.c_int, _dbl, _dbl, _c.c_int, _c.c_uint32,
                                   _c.c_int, _c.c_int, _vp, _vp, _i64, _c.POINTER(_i64)],
                                  _c.c_int),
            "rph_majorana_block": ([_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp, _vp,
                                    _i64, _c.POINTER(_i64)], _c.c_int),
            "rph_block_anticomm": ([_c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int, _c.c_int,
                                    _vp, _vp, _i64, _c.POINTER(_i64), _c.POINTER(_c.c_int)],
                                   _c.c_int),
            "rph_block_state_new": ([_c.c_int, _c.c_int], _vp),
            "rph_block_state_free": ([_vp], None),
            "rph_block_state_load": ([_vp, _vp, _vp], _c.c_int),
            "rph_block_state_store": ([_vp, _vp, _vp, _c.POINTER(_c.c_int),
                                       _c.POINTER(_c.c_int)], _c.c_int),
            "rph_block_state_random": ([_vp, _u64], _c.c_int),
            "rph_block_state_get": ([_vp, _u64, _c.POINTER(_dbl), _c.POINTER(_dbl)], _c.c_int),
            "rph_block_state_norm": ([_vp, _c.POINTER(_dbl)], _c.c_int),
            "rph_block_state_apply": ([_vp, _vp, _vp, _vp, _vp], _c.c_int),
            "rph_block_state_evolve": ([_vp, _vp, _vp, _dbl, _dbl, _c.c_int, _dbl,
                                        _c.POINTER(EvolveStats)], _c.c_int),
            "rph_block_state_ground": ([_vp, _vp, _vp, _c.c_int, _dbl, _c.c_int,
                                        _c.POINTER(_dbl), _c.POINTER(_c.c_int)], _c.c_int),
            "rph_syk_block_npoint": ([_c.c_int, _c.c_int, _c.c_int, _dbl, _dbl, _c.c_uint32,
                                      _c.c_int, _c.c_int, _vp, _u64, _dbl, _dbl, _c.c_int,
                                      _c.c_int, _dbl, _vp, _c.POINTER(_dbl),
                                      _c.POINTER(EvolveStats), _vp, _c.c_int], _c.c_int),
            "rph_state_evolve_ed": ([_vp, _vp, _dbl, _dbl], _c.c_int),
            "rph_state_ground_ed": ([_vp, _vp, _c.POINTER(_dbl)], _c.c_int),
            "rph_ham_dense": ([_vp, _c.c_int, _dbl, _dbl, _vp, _vp], _c.c_int),
            "rph_ham_matexp_entries": ([_vp, _c.POINTER(_i64)], _c.c_int),
            "rph_state_set": ([_vp, _u64, _dbl, _dbl], _c.c_int),
            "rph_set_dense_apply": ([_c.c_int], _c.c_int),
            "rph_set_fused_first_pass": ([_c.c_int], _c.c_int),
            "rph_cache_budget": ([_i64, _c.POINTER(_i64), _c.POINTER(_i64)], _c.c_int),
            "rph_ham_to_string": ([_c.c_int, _vp, _vp, _i64, _c.c_char_p, _i64,
                                   _c.POINTER(_i64)], _c.c_int),
            "rph_syk_block_dump": ([_c.c_int, _c.c_int, _dbl, _dbl, _c.c_uint32, _c.c_int,
                                    _c.c_char_p, _i64, _c.POINTER(_i64)], _c.c_int),
            "rph_format_results": ([_dbl, _vp, _i64, _c.c_char_p, _i64, _c.POINTER(_i64)],
                                   _c.c_int),
            "rph_cross_backend_evolve": ([_c.c_int, _c.c_int, _vp, _vp, _i64, _vp, _dbl, _dbl,
                                          _c.c_int, _dbl, _vp, _vp, _vp], _c.c_int),
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


def xxz_terms(L, J=1.0, Delta=1.0, W=1.0, seed=0, fp=32):
    """Term list of the random-field XXZ chain XXZ(L, J, Delta, W).gen_ham(
    std::mt19937(seed))."""
    cap = 4 * L
    bits = np.zeros(cap, np.uint64)
    co = np.zeros(2 * cap, np.float64)
    n = _i64()
    hcheck(host().rph_xxz_gen(fp, L, J, Delta, W, seed, _ptr(bits), _ptr(co), cap, _c.byref(n)))
    return bits[:n.value].copy(), (co[0:2 * n.value:2] + 1j * co[1:2 * n.value:2])


def majorana(N, n, fp=32, std_gamma=True):
    bits = np.zeros(1, np.uint64)
    co = np.zeros(2, np.float64)
    hcheck(host().rph_majorana(fp, N, n, int(std_gamma), _ptr(bits), _ptr(co)))
    return int(bits[0]), complex(co[0], co[1])


def _terms_out(fn, cap=1 << 12):
    """Call fn(bits, coeff, cap, &n) until the term list fits; (bits, complex)."""
    while True:
        bits = np.zeros(cap, np.uint64)
        co = np.zeros(2 * cap, np.float64)
        n = _i64()
        fn(_ptr(bits), _ptr(co), cap, _c.byref(n))
        if n.value >= 0:
            return bits[:n.value].copy(), (co[0:2 * n.value:2] + 1j * co[1:2 * n.value:2])
        cap = -n.value


def syk_block_terms(N, which, sp=0.0, J=1.0, seed=0, fp=32, std_gamma=True, regularize=False):
    """Block which (0 = LL, 1 = RR) of SYKBlockForm(N, sp, J).gen_ham(
    std::mt19937(seed)): a term list on N/2-1 sites."""
    return _terms_out(lambda b, c, cap, n: hcheck(host().rph_syk_block_gen(
        fp, N, sp, J, int(std_gamma), seed, int(regularize), int(which), b, c, cap, n)))


def majorana_block(N, n, which, fp=32, std_gamma=True):
    """Block which (0 = LR, 1 = RL) of FermionOpBlockForm(N, n): (bits, coeff)."""
    b, c = _terms_out(lambda bb, cc, cap, k: hcheck(host().rph_majorana_block(
        fp, N, n, int(std_gamma), int(which), bb, cc, cap, k)), cap=4)
    assert len(b) == 1
    return int(b[0]), complex(c[0])


def block_anticomm(N, i, j, which, fp=64, std_gamma=True):
    """Block which (0 = LL, 1 = RR) of chi_i chi_j + chi_j chi_i in block
    form: (bits, coeff, null)."""
    null = _c.c_int()
    b, c = _terms_out(lambda bb, cc, cap, k: hcheck(host().rph_block_anticomm(
        fp, N, i, j, int(std_gamma), int(which), bb, cc, cap, k, _c.byref(null))), cap=8)
    return b, c, bool(null.value)


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
def set_upload_flags(flags):
    """Flags of every later Hamiltonian upload (RP_TERMS_AUTOTUNE,
    RP_TERMS_ACC_NATIVE)."""
    hcheck(host().rph_set_upload_flags(flags))


def init(device=0):
    check(hip().rp_init(device))


def set_dense_apply(on):
    """Apply through the dense matrix when the term count reaches 2^len (the
    reference's small-problem switch); returns the previous setting."""
    return bool(host().rph_set_dense_apply(int(bool(on))))


def set_fused_first_pass(on):
    """Have the Krylov driver's Arnoldi steps form their first Gram-Schmidt
    pass's coefficients in the apply kernel (rp_apply_gs); returns the
    previous setting."""
    return bool(host().rph_set_fused_first_pass(int(bool(on))))


def matexp_cache_budget(max_bytes=None):
    """(max, used) bytes of the matexp cache budget; max_bytes sets the
    maximum first (GPUImpl::set_max_cache_size)."""
    mx, used = _i64(), _i64()
    hcheck(host().rph_cache_budget(-1 if max_bytes is None else int(max_bytes), _c.byref(mx),
                                   _c.byref(used)))
    return mx.value, used.value


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

    def _dense(self, what, c=0j, with_vals=False):
        n = 1 << self.len
        out = np.zeros(2 * n * n)
        vals = np.zeros(n) if with_vals else None
        hcheck(host().rph_ham_dense(self._h, what, c.real, c.imag, _ptr(out),
                                    None if vals is None else _ptr(vals)))
        m = (out[0::2] + 1j * out[1::2]).reshape(n, n).T  # column-major on the device
        return m, vals

    def matrix(self):
        """The dense form built on the device (get_matrix)."""
        return self._dense(0)[0]

    def matexp(self, c):
        """exp(c H) by exact diagonalization on the device (matexp, cached)."""
        return self._dense(1, complex(c))[0]

    def eigensystem(self):
        """(eigenvalues ascending, column eigenvectors) from rocSOLVER heevd."""
        m, vals = self._dense(2, with_vals=True)
        return vals, m

    def matexp_cache_entries(self):
        n = _i64()
        hcheck(host().rph_ham_matexp_entries(self._h, _c.byref(n)))
        return n.value

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

    def __setitem__(self, i, value):
        """One amplitude written on the device (State's ElemRefType)."""
        value = complex(value)
        hcheck(host().rph_state_set(self._s, int(i), value.real, value.imag))

    def copy_from(self, other, rounding=RP_ROUND_NEAREST):
        hcheck(host().rph_state_copy(self._s, other._s, rounding))

    def apply(self, ham):
        hcheck(host().rph_state_apply(self._s, ham._h))

    def evolve(self, ham, t, beta=0.0, krydim=5, tol=0.0, max_iters=0, reorth=0, substep=0.0):
        st = EvolveStats()
        hcheck(host().rph_state_evolve(self._s, ham._h, t, beta, krydim, tol, max_iters, reorth,
                                       substep, _c.byref(st)))
        return st

    def evolve_schedule(self, ham, t, dims, nsub=1, beta=0.0, tol=0.0, max_iters=0, reorth=0):
        """Evolve with Krylov dimension dims[min(r, len-1)] at restart r and
        nsub equal sub-steps."""
        d = (_c.c_int * len(dims))(*dims)
        st = EvolveStats()
        hcheck(host().rph_state_evolve_list(self._s, ham._h, t, beta, d, len(dims), nsub, tol,
                                            max_iters, reorth, _c.byref(st)))
        return st

    def evolve_with(self, ham, t, sched, beta=0.0, krydim=20, tol=0.0, max_iters=0, reorth=0):
        """Evolve with a Schedule; krydim is the fixed dimension a rejected
        proposal falls back to."""
        st = EvolveStats()
        hcheck(host().rph_state_evolve_sched(self._s, ham._h, sched._h, t, beta, krydim, tol,
                                             max_iters, reorth, _c.byref(st)))
        return st

    def otoc(self, ham, A, B, times, beta=0.0, krydim=12, tol=0.0, schedule=None):
        """F(t) = <i|A(t) B A(t) B|i> for each t, |i> this state (which
        becomes exp(-beta H/8)|psi> when beta != 0).  A Schedule chooses each
        restart's Krylov dimension (krydim is then the fallback).  Returns
        (F, stats)."""
        ts = np.ascontiguousarray(times, dtype=np.float64)
        out = np.zeros(2 * len(ts))
        st = EvolveStats()
        hcheck(host().rph_otoc(ham._h, A._h, B._h, self._s, beta, _ptr(ts), len(ts), krydim,
                               tol, _ptr(out), _c.byref(st),
                               None if schedule is None else schedule._h))
        return out[0::2] + 1j * out[1::2], st

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

    def evolve_ed(self, ham, t, beta=0.0):
        """exp(-(beta + it) H) by exact diagonalization (ExactDiagonalization)."""
        hcheck(host().rph_state_evolve_ed(self._s, ham._h, t, beta))

    def ground_state_ed(self, ham):
        e = _dbl()
        hcheck(host().rph_state_ground_ed(self._s, ham._h, _c.byref(e)))
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


class BlockState:
    """A REAPERS BlockState<Fp, GPUImpl>: the L and R parity sectors of
    2^length amplitudes each (full vector: L first)."""

    def __init__(self, length, fp=32):
        self.len = int(length)
        self.fp = fp
        self.half = 1 << self.len
        self._s = host().rph_block_state_new(fp, self.len)
        if not self._s:
            raise RpError(-1, host().rph_last_error().decode())

    @property
    def dtype(self):
        return np.complex64 if self.fp == 32 else np.complex128

    def load(self, L, R):
        """Load the sectors; None makes a sector null."""
        arrs = []
        for v in (L, R):
            if v is None:
                arrs.append(None)
            else:
                a = np.ascontiguousarray(v, dtype=self.dtype)
                assert a.shape == (self.half,)
                arrs.append(a)
        hcheck(host().rph_block_state_load(self._s, *[None if a is None else _ptr(a)
                                                      for a in arrs]))

    def store(self):
        """(L, R), None for a null sector."""
        L = np.empty(self.half, self.dtype)
        R = np.empty(self.half, self.dtype)
        nl, nr = _c.c_int(), _c.c_int()
        hcheck(host().rph_block_state_store(self._s, _ptr(L), _ptr(R), _c.byref(nl),
                                            _c.byref(nr)))
        return (None if nl.value else L), (None if nr.value else R)

    def random(self, seed):
        hcheck(host().rph_block_state_random(self._s, seed))

    def __getitem__(self, i):
        re, im = _dbl(), _dbl()
        hcheck(host().rph_block_state_get(self._s, int(i), _c.byref(re), _c.byref(im)))
        return complex(re.value, im.value)

    def norm(self):
        n = _dbl()
        hcheck(host().rph_block_state_norm(self._s, _c.byref(n)))
        return n.value

    def apply(self, LL=None, LR=None, RL=None, RR=None):
        """state *= the block operator with these blocks (Hamiltonian
        handles; None = null block)."""
        h = [None if x is None else x._h for x in (LL, LR, RL, RR)]
        hcheck(host().rph_block_state_apply(self._s, *h))

    def evolve(self, LL, RR, t, beta=0.0, krydim=5, tol=0.0):
        st = EvolveStats()
        hcheck(host().rph_block_state_evolve(self._s, LL._h, RR._h, t, beta, krydim, tol,
                                             _c.byref(st)))
        return st

    def ground_state(self, LL, RR, krydim=5, eps=1e-10, maxiter=0):
        """(energy, sector kept: 0 = L, 1 = R)."""
        e, k = _dbl(), _c.c_int()
        hcheck(host().rph_block_state_ground(self._s, LL._h, RR._h, krydim, eps, maxiter,
                                             _c.byref(e), _c.byref(k)))
        return e.value, k.value

    def close(self):
        if self._s:
            host().rph_block_state_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def cross_backend_evolve(length, bits, coeff, psi, t, beta=0.0, krydim=10, tol=0.0, fp=64):
    """psi as a State on the CPU backend, converted to a GPU State, both
    evolved by exp(-(beta + i t) H): returns (cpu result, gpu result, the GPU
    state converted back before the evolution)."""
    dt = np.complex64 if fp == 32 else np.complex128
    bits = np.ascontiguousarray(bits, np.uint64)
    c = np.asarray(coeff, np.complex128)
    co = np.ascontiguousarray(np.stack([c.real, c.imag], 1).ravel())
    v = np.ascontiguousarray(psi, dt)
    outs = [np.empty_like(v) for _ in range(3)]
    hcheck(host().rph_cross_backend_evolve(fp, length, _ptr(bits), _ptr(co), len(bits), _ptr(v),
                                           t, beta, krydim, tol, *[_ptr(o) for o in outs]))
    return tuple(outs)


def syk_block_npoint(N, kind="otoc", sp=0.0, J=1.0, seed=0, regularize=False, otoc_idx=-1,
                     init=None, state_seed=1, beta=0.0, tmax=1.0, nsteps=10, krylov_dim=5,
                     tol=0.0, fp=32, schedule=None, promote=False):
    """One disorder realization of the parity-block SYK OTOC or Green's
    function: returns (t, v, e0, stats) with v_k / v_0 at t_k = k tmax/nsteps
    and e0 the ground-state energy the Hamiltonian was shifted by.  kind
    'otoc-trace' / 'green-trace' use the trace definitions by exact
    diagonalization (no state; e0 = 0).  A Schedule chooses each restart's
    Krylov dimension (krylov_dim stays the fallback and the Lanczos size).
    promote (fp=64) generates H and the random state in FP32 and converts
    them: the FP64 counterpart of an FP32 run on the same couplings."""
    k = {"otoc": 0, "green": 1, "otoc-trace": 2, "green-trace": 3}[kind]
    out = np.zeros(2 * (nsteps + 1))
    e0 = _dbl()
    st = EvolveStats()
    a = None
    if init is not None:
        a = np.ascontiguousarray(init, np.complex64 if fp == 32 else np.complex128)
        assert a.shape == (1 << (N // 2 - 1),)
    hcheck(host().rph_syk_block_npoint(k, fp, N, sp, J, seed, int(regularize), otoc_idx,
                                       None if a is None else _ptr(a), state_seed, beta, tmax,
                                       nsteps, krylov_dim, tol, _ptr(out), _c.byref(e0),
                                       _c.byref(st), None if schedule is None else schedule._h,
                                       int(promote)))
    t = np.arange(nsteps + 1) * (tmax / nsteps if nsteps else 0.0)
    return t, out[0::2] + 1j * out[1::2], e0.value, st


# ---------------------------------------------------------------------------
# Krylov schedules and the proposer network
# ---------------------------------------------------------------------------
MLP_IN, MLP_OUT, MLP_NPARAMS, MLP_MAX_INFER = 12, 2, 5122, 64


class Schedule:
    """A C++ KrylovSchedule handle (ListSchedule or MLPSchedule) that can
    record every decision with the proposer's 12 inputs."""

    def __init__(self, handle):
        if not handle:
            raise RpError(-1, host().rph_last_error().decode())
        self._h = handle

    @classmethod
    def from_list(cls, dims, nsub=1):
        d = (_c.c_int * len(dims))(*dims)
        return cls(host().rph_sched_list_new(d, len(dims), nsub))

    @classmethod
    def from_mlp(cls, params, min_dim=3, max_dim=40, max_sub=8):
        p = np.ascontiguousarray(params, dtype=np.float32)
        assert p.size == MLP_NPARAMS
        return cls(host().rph_sched_mlp_new(_ptr(p), min_dim, max_dim, max_sub))

    def tracing(self, on=True):
        hcheck(host().rph_sched_set_tracing(self._h, 1 if on else 0))

    def trace(self):
        """(features[n,12] float32, meta[n,4] int32 (kind, value, restart,
        sub-step), previous update norms[n])."""
        cnt = _c.c_int()
        hcheck(host().rph_sched_trace(self._h, None, None, None, 0, _c.byref(cnt)))
        n = cnt.value
        f = np.zeros((n, MLP_IN), np.float32)
        m = np.zeros((n, 4), np.int32)
        u = np.zeros(n)
        hcheck(host().rph_sched_trace(self._h, _ptr(f), _ptr(m), _ptr(u), n, _c.byref(cnt)))
        return f, m, u

    def close(self):
        if self._h:
            host().rph_sched_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceArray:
    """A float32 device array from rp_malloc (for the proposer's training)."""

    def __init__(self, host_array):
        a = np.ascontiguousarray(host_array, dtype=np.float32)
        self.shape, self.nbytes = a.shape, a.nbytes
        self.ptr = _vp()
        check(hip().rp_malloc(_c.byref(self.ptr), max(4, a.nbytes)))
        if a.nbytes:
            check(hip().rp_memcpy(self.ptr, _ptr(a), a.nbytes, RP_H2D))

    @classmethod
    def zeros(cls, shape):
        return cls(np.zeros(shape, np.float32))

    def get(self):
        a = np.empty(self.shape, np.float32)
        check(hip().rp_memcpy(_ptr(a), self.ptr, self.nbytes, RP_D2H))
        return a

    def set(self, host_array):
        a = np.ascontiguousarray(host_array, dtype=np.float32)
        assert a.nbytes == self.nbytes
        check(hip().rp_memcpy(self.ptr, _ptr(a), a.nbytes, RP_H2D))

    def close(self):
        if self.ptr:
            hip().rp_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def mlp_forward(params, X):
    """Y = net(X) by the HIP forward kernel; params and X are DeviceArrays."""
    n = X.shape[0]
    Y = DeviceArray.zeros((n, MLP_OUT))
    check(hip().rp_mlp_forward(params.ptr, X.ptr, n, Y.ptr))
    return Y


def mlp_loss_grad(params, X, T, W, grad):
    """loss of sum W (net(X) - T)^2 / n; grad (a DeviceArray) <- its gradient."""
    loss = _dbl()
    check(hip().rp_mlp_loss_grad(params.ptr, X.ptr, T.ptr, None if W is None else W.ptr,
                                 X.shape[0], grad.ptr, _c.byref(loss)))
    return loss.value


def mlp_adam(params, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8, wd=0.0):
    check(hip().rp_mlp_adam(params.ptr, grad.ptr, m.ptr, v.ptr, MLP_NPARAMS, lr, b1, b2, eps,
                            wd, step))


def mlp_infer(params, x):
    """net(x) for up to MLP_MAX_INFER host samples, through rp_mlp_infer."""
    x = np.ascontiguousarray(x, dtype=np.float32).reshape(-1, MLP_IN)
    y = np.empty((x.shape[0], MLP_OUT), np.float32)
    check(hip().rp_mlp_infer(params.ptr, _ptr(x), x.shape[0], _ptr(y)))
    return y


# ---------------------------------------------------------------------------
# batches of right-hand sides
# ---------------------------------------------------------------------------
class Batch:
    """BatchedStates<Fp>: problems (lens[p], bits[p], coeff[p]) with nrhs
    states each in the block-state layout."""

    def __init__(self, lens, terms, nrhs=4, kmax=20, fp=32, autotune=True, tile_bits=0,
                 acc_group=False, wide=False, planes=False):
        self.lens = [int(x) for x in lens]
        self.nprob = len(self.lens)
        self.nrhs = nrhs
        self.fp = fp
        bits = np.ascontiguousarray(np.concatenate([np.asarray(b, np.uint64) for b, _ in terms]))
        coeff = np.concatenate([np.asarray(c, np.complex128) for _, c in terms])
        co = np.ascontiguousarray(np.stack([coeff.real, coeff.imag], 1).ravel())
        toff = np.zeros(self.nprob + 1, np.int64)
        toff[1:] = np.cumsum([len(b) for b, _ in terms])
        la = np.array(self.lens, np.int32)
        self._h = host().rph_batch_new(fp, self.nprob, _ptr(la), _ptr(bits), _ptr(co), _ptr(toff),
                                       nrhs, kmax, ((tile_bits << 8) if tile_bits
                                                    else (1 if autotune else 0))
                                       | (2 if acc_group else 0) | (4 if wide else 0)
                                       | (8 if planes else 0))
        if not self._h:
            raise RpError(-1, host().rph_last_error().decode())
        self.info = BatchInfo()
        self.offsets = np.zeros(self.nprob + 1, np.uint64)
        hcheck(host().rph_batch_info(self._h, _c.byref(self.info), _ptr(self.offsets)))

    @property
    def dtype(self):
        return np.complex64 if self.fp == 32 else np.complex128

    def pack(self, states):
        """states[p]: array (2^lens[p], nrhs) -> the batch vector (host)."""
        out = np.zeros((int(self.info.nbasis), self.nrhs), self.dtype)
        for p, x in enumerate(states):
            o = int(self.offsets[p])
            out[o:o + (1 << self.lens[p])] = x
        return out.ravel()

    def unpack(self, v):
        v = v.reshape(-1, self.nrhs)
        return [v[int(self.offsets[p]):int(self.offsets[p]) + (1 << self.lens[p])].copy()
                for p in range(self.nprob)]

    def load(self, states):
        a = np.ascontiguousarray(self.pack(states))
        hcheck(host().rph_batch_load(self._h, _ptr(a)))

    def store(self):
        a = np.empty(int(self.info.nbasis) * self.nrhs, self.dtype)
        hcheck(host().rph_batch_store(self._h, _ptr(a)))
        return self.unpack(a)

    def apply(self):
        hcheck(host().rph_batch_apply(self._h))

    def evolve(self, t, m=12, beta=0.0, tol=0.0, max_iters=0, graph=True, schedule=None):
        """Evolve every state; schedule: a Schedule (Krylov dimension per
        restart, at most kmax) or None for the fixed m."""
        st = EvolveStats()
        rs = np.zeros(self.nprob * self.nrhs, np.int32)
        hcheck(host().rph_batch_evolve(self._h, t, beta, m, tol, max_iters, 1 if graph else 0,
                                       schedule._h if schedule is not None else None,
                                       _c.byref(st), _ptr(rs)))
        return st, rs.reshape(self.nprob, self.nrhs)

    def close(self):
        if self._h:
            host().rph_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


from .formats import (format_result_lines, ham_to_string, parse_ham_string,  # noqa: E402
                      read_result_file, service_request, syk_block_dump, write_result_file)
