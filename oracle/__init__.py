"""ctypes wrapper of the CPU oracle (liboracle.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  Nothing in reapers_amd imports
it.  See oracle.h for the reference file:line each function restates.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_c = ctypes
_vp = _c.c_void_p
_lib = None
_variants = {}


class KrylovStats(_c.Structure):
    _fields_ = [("matvecs", _c.c_int64), ("restarts", _c.c_int64), ("gs_passes", _c.c_int64),
                ("breakdown", _c.c_int), ("last_nrm", _c.c_double)]


def _bind(L):
    """Declare the C signatures on a loaded oracle library."""
    if True:
        L.or_mt19937_kat.restype = _c.c_uint32
        L.or_mt_stream.argtypes = [_c.c_uint32, _vp, _c.c_int64]
        L.or_normal_stream_f32.argtypes = [_c.c_uint32, _vp, _c.c_int64]
        L.or_normal_stream_f64.argtypes = [_c.c_uint32, _vp, _c.c_int64]
        L.or_pauli_mul.restype = _c.c_uint64
        L.or_pauli_mul.argtypes = [_c.c_uint64, _c.c_uint64, _c.POINTER(_c.c_int)]
        L.or_pauli_apply.argtypes = [_c.c_uint64, _c.c_uint32, _c.POINTER(_c.c_uint32),
                                     _c.POINTER(_c.c_int)]
        L.or_count_sy.argtypes = [_c.c_uint64]
        L.or_pauli_sparse_matrix.argtypes = [_c.c_uint64, _c.c_int, _vp, _vp]
        L.or_majorana_f32.argtypes = [_c.c_int, _c.c_int, _c.c_int, _vp, _vp]
        L.or_majorana_f64.argtypes = [_c.c_int, _c.c_int, _c.c_int, _vp, _vp]
        L.or_syk_gen_f32.restype = _c.c_int64
        L.or_syk_gen_f32.argtypes = [_c.c_int, _c.c_float, _c.c_float, _c.c_int, _c.c_uint32,
                                     _c.c_int, _vp, _vp, _c.c_int64]
        L.or_syk_gen_f64.restype = _c.c_int64
        L.or_syk_gen_f64.argtypes = [_c.c_int, _c.c_double, _c.c_double, _c.c_int, _c.c_uint32,
                                     _c.c_int, _vp, _vp, _c.c_int64]
        L.or_majorana_block_f32.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp, _vp]
        L.or_majorana_block_f64.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp, _vp]
        L.or_syk_block_gen_f32.restype = _c.c_int64
        L.or_syk_block_gen_f32.argtypes = [_c.c_int, _c.c_float, _c.c_float, _c.c_int,
                                           _c.c_uint32, _c.c_int, _c.c_int, _vp, _vp, _c.c_int64]
        L.or_syk_block_gen_f64.restype = _c.c_int64
        L.or_syk_block_gen_f64.argtypes = [_c.c_int, _c.c_double, _c.c_double, _c.c_int,
                                           _c.c_uint32, _c.c_int, _c.c_int, _vp, _vp, _c.c_int64]
        L.or_majorana_anticomm_f64.restype = _c.c_int64
        L.or_majorana_anticomm_f64.argtypes = [_c.c_int, _c.c_int, _c.c_int, _c.c_int, _vp, _vp,
                                               _c.c_int64]
        for n in ("or_apply_f64", "or_apply_f32"):
            getattr(L, n).argtypes = [_c.c_int, _vp, _vp, _c.c_int64, _vp, _vp]
        for n in ("or_apply_rows_f64", "or_apply_rows_f32", "or_apply_window_f64",
                  "or_apply_window_f32"):
            getattr(L, n).argtypes = [_c.c_int, _vp, _vp, _c.c_int64, _vp, _vp, _c.c_int64,
                                      _c.c_int64]
        L.or_dense_matrix_f64.argtypes = [_c.c_int, _vp, _vp, _c.c_int64, _vp]
        L.or_expm_f64.argtypes = [_c.c_int, _vp, _vp]
        L.or_expm_f32.argtypes = [_c.c_int, _vp, _vp]
        L.or_krylov_evolve_f64.argtypes = [_c.c_int, _vp, _vp, _c.c_int64, _vp, _c.c_double,
                                           _c.c_double, _c.c_int, _c.c_double, _c.c_int,
                                           _c.POINTER(KrylovStats)]
        L.or_krylov_evolve_f32.argtypes = [_c.c_int, _vp, _vp, _c.c_int64, _vp, _c.c_float,
                                           _c.c_float, _c.c_int, _c.c_float, _c.c_int,
                                           _c.POINTER(KrylovStats)]
        L.or_lanczos_ground_state_f64.restype = _c.c_double
        L.or_lanczos_ground_state_f64.argtypes = [_c.c_int, _vp, _vp, _c.c_int64, _vp, _c.c_int,
                                                  _c.c_double, _c.c_int, _c.POINTER(_c.c_int)]
        L.or_num_threads.restype = _c.c_int
        L.or_bernoulli_stream.argtypes = [_c.c_uint32, _c.c_double, _vp, _c.c_int64]
        L.or_uniform_int_stream.argtypes = [_c.c_uint32, _c.c_int, _c.c_int, _vp, _c.c_int64]
        L.or_canonical_stream.argtypes = [_c.c_uint32, _c.c_int, _vp, _c.c_int64]
        L.or_mlp_nparams.restype = _c.c_int
        L.or_mlp_forward_f64.argtypes = [_vp, _vp, _c.c_int64, _vp]
        L.or_mlp_loss_grad_f64.restype = _c.c_double
        L.or_mlp_loss_grad_f64.argtypes = [_vp, _vp, _vp, _vp, _c.c_int64, _vp]
        L.or_mlp_adam_f64.argtypes = [_vp, _vp, _vp, _vp, _c.c_int64, _c.c_double, _c.c_double,
                                      _c.c_double, _c.c_double, _c.c_double, _c.c_int64]
    return L


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        _lib = _bind(_c.CDLL(LIB))
    return _lib


def fast():
    """The timing-only build (liboracle_fast.so: -O3 with AVX2/FMA, FMA
    contraction allowed), for bench.py's CPU baseline; never the checker."""
    if "fast" not in _variants:
        so = os.path.join(HERE, "liboracle_fast.so")
        if not os.path.exists(so):
            raise RuntimeError("oracle/liboracle_fast.so is not built (make -C oracle)")
        _variants["fast"] = _bind(_c.CDLL(so))
    return _variants["fast"]


def planted(n):
    """A deliberately wrong build of the oracle (OR_PLANT=n, see oracle.c),
    built on demand; used only to show that the pinning tests catch it."""
    if n not in _variants:
        import subprocess
        so = os.path.join(HERE, "liboracle_plant%d.so" % n)
        subprocess.check_call(["make", "-s", "-C", HERE, "liboracle_plant%d.so" % n])
        _variants[n] = _bind(_c.CDLL(so))
    return _variants[n]


class use:
    """Context manager: route this module's wrappers to another loaded
    oracle library (a planted variant)."""

    def __init__(self, L):
        self.L = L

    def __enter__(self):
        global _lib
        self.prev = lib()
        _lib = self.L
        return self

    def __exit__(self, *a):
        global _lib
        _lib = self.prev


def _p(a):
    return a.ctypes.data_as(_vp)


def _split(coeff, dt):
    c = np.empty(2 * len(coeff), dt)
    c[0::2] = np.real(coeff)
    c[1::2] = np.imag(coeff)
    return c


def _join(c):
    return c[0::2] + 1j * c[1::2]


def mt_stream(seed, n):
    out = np.zeros(n, np.uint32)
    lib().or_mt_stream(seed, _p(out), n)
    return out


def normal_stream(seed, n, fp=64):
    out = np.zeros(n, np.float32 if fp == 32 else np.float64)
    (lib().or_normal_stream_f32 if fp == 32 else lib().or_normal_stream_f64)(seed, _p(out), n)
    return out


def bernoulli_stream(seed, p, n):
    out = np.zeros(n)
    lib().or_bernoulli_stream(seed, p, _p(out), n)
    return out


def uniform_int_stream(seed, a, b, n):
    out = np.zeros(n)
    lib().or_uniform_int_stream(seed, a, b, _p(out), n)
    return out


def canonical_stream(seed, n, fp=64):
    out = np.zeros(n)
    lib().or_canonical_stream(seed, fp, _p(out), n)
    return out


def pauli_mul(b0, b1):
    m = _c.c_int()
    b = lib().or_pauli_mul(b0, b1, _c.byref(m))
    return int(b), bool(m.value)


def pauli_apply(bits, n):
    r = _c.c_uint32()
    m = _c.c_int()
    lib().or_pauli_apply(bits, n, _c.byref(r), _c.byref(m))
    return int(r.value), bool(m.value)


def count_sy(bits):
    return lib().or_count_sy(bits)


def sparse_matrix(bits, length):
    dim = 1 << length
    cols = np.zeros(max(dim, 2), np.uint32)
    neg = np.zeros(max(dim, 2), np.uint8)
    lib().or_pauli_sparse_matrix(bits, length, _p(cols), _p(neg))
    return cols[:dim], neg[:dim].astype(bool)


def majorana(N, n, fp=64, std_gamma=True):
    bits = np.zeros(1, np.uint64)
    if fp == 32:
        c = np.zeros(2, np.float32)
        lib().or_majorana_f32(N, n, int(std_gamma), _p(bits), _p(c))
    else:
        c = np.zeros(2, np.float64)
        lib().or_majorana_f64(N, n, int(std_gamma), _p(bits), _p(c))
    return int(bits[0]), complex(c[0], c[1])


def syk_terms(N, sp=0.0, J=1.0, seed=0, fp=64, std_gamma=True, regularize=False):
    cap = 1 << 12
    while True:
        bits = np.zeros(cap, np.uint64)
        if fp == 32:
            c = np.zeros(2 * cap, np.float32)
            n = lib().or_syk_gen_f32(N, sp, J, int(std_gamma), seed, int(regularize), _p(bits),
                                     _p(c), cap)
        else:
            c = np.zeros(2 * cap, np.float64)
            n = lib().or_syk_gen_f64(N, sp, J, int(std_gamma), seed, int(regularize), _p(bits),
                                     _p(c), cap)
        if n >= 0:
            return bits[:n].copy(), _join(c[:2 * n].astype(np.float64))
        cap = -n


def majorana_block(N, n, which, fp=64, std_gamma=True):
    """Block which (0 = LR, 1 = RL) of the parity-block Majorana n: (bits, coeff)."""
    bits = np.zeros(1, np.uint64)
    if fp == 32:
        c = np.zeros(2, np.float32)
        lib().or_majorana_block_f32(N, n, int(std_gamma), int(which), _p(bits), _p(c))
    else:
        c = np.zeros(2, np.float64)
        lib().or_majorana_block_f64(N, n, int(std_gamma), int(which), _p(bits), _p(c))
    return int(bits[0]), complex(c[0], c[1])


def syk_block_terms(N, which, sp=0.0, J=1.0, seed=0, fp=64, std_gamma=True, regularize=False):
    """Block which (0 = LL, 1 = RR) of the parity-block SYK Hamiltonian."""
    cap = 1 << 12
    while True:
        bits = np.zeros(cap, np.uint64)
        if fp == 32:
            c = np.zeros(2 * cap, np.float32)
            n = lib().or_syk_block_gen_f32(N, sp, J, int(std_gamma), seed, int(regularize),
                                           int(which), _p(bits), _p(c), cap)
        else:
            c = np.zeros(2 * cap, np.float64)
            n = lib().or_syk_block_gen_f64(N, sp, J, int(std_gamma), seed, int(regularize),
                                           int(which), _p(bits), _p(c), cap)
        if n >= 0:
            return bits[:n].copy(), _join(c[:2 * n].astype(np.float64))
        cap = -n


def majorana_anticomm(N, i, j, std_gamma=True):
    bits = np.zeros(8, np.uint64)
    c = np.zeros(16)
    n = lib().or_majorana_anticomm_f64(N, i, j, int(std_gamma), _p(bits), _p(c), 8)
    return bits[:n].copy(), _join(c[:2 * n])


def apply(length, bits, coeff, psi, fp=64, out=None):
    """out += H psi (matrix-free, reference term order)."""
    bits = np.ascontiguousarray(bits, np.uint64)
    dt = np.float32 if fp == 32 else np.float64
    c = _split(coeff, dt)
    vin = _split(psi, dt)
    vout = np.zeros(2 * (1 << length), dt) if out is None else _split(out, dt)
    f = lib().or_apply_f32 if fp == 32 else lib().or_apply_f64
    f(length, _p(bits), _p(c), len(bits), _p(vin), _p(vout))
    return _join(vout)


def apply_rows(length, bits, coeff, psi, row_begin, row_end, fp=64):
    """Rows [row_begin, row_end) of H psi (the reference's CPU apply,
    inc/cpuimpl.h:737-754, restricted to a window of output rows).  A psi
    already of the fp's complex dtype is read in place (no copy of a large
    state)."""
    bits = np.ascontiguousarray(bits, np.uint64)
    dt = np.float32 if fp == 32 else np.float64
    ct = np.complex64 if fp == 32 else np.complex128
    c = _split(coeff, dt)
    if psi.dtype == ct and psi.flags.c_contiguous:
        vin = psi.view(dt)
    else:
        vin = _split(psi, dt)
    vout = np.zeros(2 * (row_end - row_begin), dt)
    f = lib().or_apply_window_f32 if fp == 32 else lib().or_apply_window_f64
    f(length, _p(bits), _p(c), len(bits), _p(vin), _p(vout), row_begin, row_end)
    return _join(vout)


def apply_rows_timed(length, bits, coeff, psi, row_end, fp=32):
    """Time rows [0, row_end) of out += H psi (the reference's CPU apply) and
    return (seconds, out rows).  psi must already have the fp's dtype layout
    (complex64 for fp=32)."""
    import time
    bits = np.ascontiguousarray(bits, np.uint64)
    dt = np.float32 if fp == 32 else np.float64
    c = _split(coeff, dt)
    vin = np.ascontiguousarray(psi).view(dt)
    vout = np.zeros(2 * row_end, dt)
    vout.fill(0)                # map the output pages before the timed call
    f = lib().or_apply_rows_f32 if fp == 32 else lib().or_apply_rows_f64
    t0 = time.perf_counter()
    f(length, _p(bits), _p(c), len(bits), _p(vin), _p(vout), 0, row_end)
    return time.perf_counter() - t0, _join(vout)


def dense_matrix(length, bits, coeff):
    dim = 1 << length
    bits = np.ascontiguousarray(bits, np.uint64)
    c = _split(coeff, np.float64)
    m = np.zeros(2 * dim * dim)
    lib().or_dense_matrix_f64(length, _p(bits), _p(c), len(bits), _p(m))
    return _join(m).reshape(dim, dim).T.copy()


def expm(a, fp=64):
    a = np.asarray(a)
    n = a.shape[0]
    dt = np.float32 if fp == 32 else np.float64
    flat = _split(a.T.reshape(-1), dt)
    out = np.zeros_like(flat)
    (lib().or_expm_f32 if fp == 32 else lib().or_expm_f64)(n, _p(flat), _p(out))
    return _join(out.astype(np.float64)).reshape(n, n).T.copy()


def krylov_evolve(length, bits, coeff, psi, t, beta=0.0, krydim=5, tol=None, max_iters=0, fp=64):
    """exp(-(beta + i t) H) psi by the reference's restarted Krylov method."""
    bits = np.ascontiguousarray(bits, np.uint64)
    dt = np.float32 if fp == 32 else np.float64
    if tol is None:
        tol = float(np.finfo(dt).eps)
    c = _split(coeff, dt)
    v = _split(psi, dt)
    st = KrylovStats()
    f = lib().or_krylov_evolve_f32 if fp == 32 else lib().or_krylov_evolve_f64
    rc = f(length, _p(bits), _p(c), len(bits), _p(v), t, beta, krydim, tol, max_iters,
           _c.byref(st))
    if rc != 0:
        raise RuntimeError("oracle krylov failed: %d" % rc)
    return _join(v.astype(np.float64)), st


def lanczos_ground_state(length, bits, coeff, psi, krydim=5, eps=1e-10, maxiter=0):
    bits = np.ascontiguousarray(bits, np.uint64)
    c = _split(coeff, np.float64)
    v = _split(psi, np.float64)
    it = _c.c_int()
    e = lib().or_lanczos_ground_state_f64(length, _p(bits), _p(c), len(bits), _p(v), krydim,
                                          eps, maxiter, _c.byref(it))
    return e, _join(v), it.value


def num_threads():
    return lib().or_num_threads()


# ---- Krylov-schedule proposer network ------------------------------------
MLP_NPARAMS = 5122


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def mlp_forward(params, X):
    """Y[n][2] of the 12-64-64-2 tanh network, in double precision."""
    p, X = _f64(params), _f64(X).reshape(-1, 12)
    Y = np.empty((X.shape[0], 2))
    lib().or_mlp_forward_f64(_p(p), _p(X), X.shape[0], _p(Y))
    return Y


def mlp_loss_grad(params, X, T, W=None):
    """(loss, gradient) of sum w (y - t)^2 / n."""
    p, X, T = _f64(params), _f64(X).reshape(-1, 12), _f64(T).reshape(-1, 2)
    Wc = None if W is None else _f64(W).reshape(-1, 2)
    g = np.empty(MLP_NPARAMS)
    loss = lib().or_mlp_loss_grad_f64(_p(p), _p(X), _p(T), None if Wc is None else _p(Wc),
                                      X.shape[0], _p(g))
    return loss, g


def mlp_adam(params, grad, m, v, lr, b1, b2, eps, wd, step):
    """One AdamW step in place on float64 arrays params, m, v."""
    lib().or_mlp_adam_f64(_p(params), _p(_f64(grad)), _p(m), _p(v), params.size, lr, b1, b2,
                          eps, wd, step)
