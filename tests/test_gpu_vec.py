"""GPU parity of the vector operations behind the C ABI (GPUImpl's BLAS-1,
inc/gpuimpl.h:806-1062, and the fused Gram-Schmidt / linear-combination
sweeps) against numpy in FP64.  Reductions accumulate in FP64 for both
precisions, so FP32 results are checked to FP32 input rounding (1e-6) and
FP64 results to 1e-13."""
import ctypes

import numpy as np
import pytest

import reapers_amd as ra
from gpu_util import DeviceBuffer, prec_of

pytestmark = pytest.mark.gpu


def _rand(rng, n, dtype):
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(dtype)


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_blas1(gpu, dtype):
    rng = np.random.default_rng(0)
    n = 100003          # odd, not a power of two
    a, b = _rand(rng, n, dtype), _rand(rng, n, dtype)
    p = prec_of(dtype)
    h = gpu.hip()
    da, db = DeviceBuffer.from_array(a), DeviceBuffer.from_array(b)
    re, im = ctypes.c_double(), ctypes.c_double()
    ra.check(h.rp_vec_prod(p, da.ptr, db.ptr, n, ctypes.byref(re), ctypes.byref(im)))
    ref = np.vdot(a.astype(complex), b.astype(complex))
    assert abs(complex(re.value, im.value) - ref) <= 1e-12 * np.sqrt(n) * 4
    nrm = ctypes.c_double()
    ra.check(h.rp_vec_norm(p, da.ptr, n, ctypes.byref(nrm)))
    assert abs(nrm.value - np.linalg.norm(a.astype(complex))) <= 1e-12 * nrm.value
    tol = 1e-6 if dtype == np.complex64 else 1e-14
    s = 0.3 - 1.7j
    ra.check(h.rp_axpy(p, da.ptr, s.real, s.imag, db.ptr, n))
    got = da.to_array(dtype, n)
    assert np.abs(got - (a.astype(complex) + s * b.astype(complex))).max() <= tol * 10
    ra.check(h.rp_scale_vec(p, da.ptr, 0.5, 0.25, n))
    got2 = da.to_array(dtype, n)
    assert np.abs(got2 - got.astype(complex) * (0.5 + 0.25j)).max() <= tol * 10
    ra.check(h.rp_add_vec(p, da.ptr, db.ptr, n))
    got3 = da.to_array(dtype, n)
    assert np.abs(got3 - (got2.astype(complex) + b)).max() <= tol * 10
    dc = DeviceBuffer(a.nbytes)
    ra.check(h.rp_add_vec3(p, dc.ptr, da.ptr, db.ptr, n))
    assert np.abs(dc.to_array(dtype, n) - (got3.astype(complex) + b)).max() <= tol * 10
    ra.check(h.rp_zero_vec(p, dc.ptr, n))
    assert np.all(dc.to_array(dtype, n) == 0)


def test_precision_conversion_rounding(gpu):
    rng = np.random.default_rng(1)
    n = 4097
    x = rng.standard_normal(n) * 1e3 + 1j * rng.standard_normal(n)
    x = x.astype(np.complex128)
    dx = DeviceBuffer.from_array(x)
    df = DeviceBuffer(n * 8)
    h = gpu.hip()
    ra.check(h.rp_copy_vec(ra.RP_C64, df.ptr, ra.RP_C128, dx.ptr, n, ra.RP_ROUND_NEAREST))
    assert np.array_equal(df.to_array(np.complex64, n), x.astype(np.complex64))
    ra.check(h.rp_copy_vec(ra.RP_C64, df.ptr, ra.RP_C128, dx.ptr, n, ra.RP_ROUND_TOWARD_ZERO))
    got = df.to_array(np.complex64, n)
    for part in (np.real, np.imag):
        g, r = part(got).astype(np.float64), part(x)
        assert np.all(np.abs(g) <= np.abs(r))                 # never rounds away from zero
        nxt = np.nextafter(part(got), np.sign(r).astype(np.float32) * np.inf).astype(np.float64)
        assert np.all(np.abs(nxt) > np.abs(r) - 0) & True     # truncation is within one ulp
        assert np.all(np.abs(g - r) < np.abs(nxt - g) + 1e-300)
    dd = DeviceBuffer(n * 16)
    ra.check(h.rp_copy_vec(ra.RP_C128, dd.ptr, ra.RP_C64, df.ptr, n, 0))
    assert np.array_equal(dd.to_array(np.complex128, n), got.astype(np.complex128))


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("k", [0, 1, 7, 8, 9, 16, 17, 20, 33, 64])
def test_gs_pass_and_lincomb(gpu, dtype, k):
    rng = np.random.default_rng(100 + k)
    n = 65536 + 3
    V = [_rand(rng, n, dtype) for _ in range(k)]
    w = _rand(rng, n, dtype)
    p = prec_of(dtype)
    dV = [DeviceBuffer.from_array(v) for v in V]
    dw = DeviceBuffer.from_array(w)
    ptrs = (ctypes.c_void_p * max(k, 1))(*[d.ptr.value for d in dV])
    hbuf = np.zeros(2 * max(k, 1))
    nrm = ctypes.c_double()
    ra.check(gpu.hip().rp_gs_pass(p, ptrs, k, dw.ptr, n, hbuf.ctypes.data_as(ctypes.c_void_p),
                                  ctypes.byref(nrm)))
    Vd = [v.astype(complex) for v in V]
    wd = w.astype(complex)
    href = np.array([np.vdot(v, wd) for v in Vd]) if k else np.zeros(0, complex)
    hgot = hbuf[0:2 * k:2] + 1j * hbuf[1:2 * k:2]
    if k:
        assert np.abs(hgot - href).max() <= 1e-12 * np.sqrt(n) * 8
    wref = wd - sum((href[i] * Vd[i] for i in range(k)), np.zeros(n, complex))
    wgot = dw.to_array(dtype, n)
    tol = 2e-6 * (1 + np.abs(href).sum()) if dtype == np.complex64 else 1e-12 * (1 + np.abs(href).sum())
    assert np.abs(wgot - wref).max() <= tol
    assert abs(nrm.value - np.linalg.norm(wgot.astype(complex))) <= 1e-12 * nrm.value
    # lincomb: y += sum c_i V_i
    if k:
        c = rng.standard_normal(k) + 1j * rng.standard_normal(k)
        cc = np.empty(2 * k); cc[0::2] = c.real; cc[1::2] = c.imag
        y = _rand(rng, n, dtype)
        dy = DeviceBuffer.from_array(y)
        ra.check(gpu.hip().rp_lincomb(p, dy.ptr, ptrs, k, cc.ctypes.data_as(ctypes.c_void_p), n))
        yref = y.astype(complex) + sum(c[i] * Vd[i] for i in range(k))
        tol = 2e-6 * (1 + np.abs(c).sum()) if dtype == np.complex64 else 1e-12 * (1 + np.abs(c).sum())
        assert np.abs(dy.to_array(dtype, n) - yref).max() <= tol


def test_init_random_is_normalized_and_seeded(gpu):
    n = 1 << 16
    h = gpu.hip()
    d1, d2 = DeviceBuffer(n * 16), DeviceBuffer(n * 16)
    ra.check(h.rp_init_random(ra.RP_C128, d1.ptr, n, 42))
    ra.check(h.rp_init_random(ra.RP_C128, d2.ptr, n, 42))
    a, b = d1.to_array(np.complex128, n), d2.to_array(np.complex128, n)
    assert np.array_equal(a, b)
    assert abs(np.linalg.norm(a) - 1) < 1e-14
    # Gaussian moments of the unnormalized components: mean 0, variance 1/(2n)
    comp = np.concatenate([a.real, a.imag]) * np.sqrt(2 * n)
    assert abs(comp.mean()) < 0.02 and abs(comp.var() - 1) < 0.03


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
@pytest.mark.parametrize("k", [1, 7, 17, 40])
def test_gs_pass_fused_normalization(gpu, dtype, k):
    """rp_gs_pass_fused: with an orthonormal basis and a w mostly outside its
    span, the pass writes (w - V h) / beta with beta = sqrt(||w||^2 - |h|^2);
    with fuse_min above beta, or with heavy cancellation, it is the plain pass."""
    rng = np.random.default_rng(7 * k)
    n = 65536 + 5
    A = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    Q, _ = np.linalg.qr(A)
    V = [Q[:, i].astype(dtype) for i in range(k)]
    p = prec_of(dtype)
    dV = [DeviceBuffer.from_array(v) for v in V]
    ptrs = (ctypes.c_void_p * k)(*[d.ptr.value for d in dV])
    Vd = [v.astype(complex) for v in V]
    for case in ("fused", "too_small", "cancellation"):
        if case == "cancellation":
            # nearly all of w inside the span: beta^2 << ||w||^2 / 2
            inside = sum(Vd[i] * (1 + abs(rng.standard_normal())) for i in range(k))
            w = (inside * 10 * np.sqrt(n) + _rand(rng, n, dtype).astype(complex) * 0.1).astype(dtype)
        else:
            w = _rand(rng, n, dtype)
        dw = DeviceBuffer.from_array(w)
        hbuf = np.zeros(2 * k)
        nrm, nz = ctypes.c_double(), ctypes.c_int()
        wd = w.astype(complex)
        href = np.array([np.vdot(v, wd) for v in Vd])
        wref = wd - sum(href[i] * Vd[i] for i in range(k))
        beta = np.linalg.norm(wref)
        fuse_min = 1e-3 if case != "too_small" else 2 * beta
        ra.check(gpu.hip().rp_gs_pass_fused(p, ptrs, k, dw.ptr, n,
                                            hbuf.ctypes.data_as(ctypes.c_void_p),
                                            ctypes.byref(nrm), fuse_min, ctypes.byref(nz)))
        wgot = dw.to_array(dtype, n).astype(complex)
        tol = 2e-6 if dtype == np.complex64 else 1e-12
        if case == "fused":
            assert nz.value == 1
            assert abs(nrm.value - beta) <= tol * beta
            assert np.abs(wgot - wref / beta).max() <= tol * np.abs(wref / beta).max() * 4
        else:
            assert nz.value == 0
            assert abs(nrm.value - np.linalg.norm(wgot)) <= 1e-12 * nrm.value
            assert np.abs(wgot - wref).max() <= tol * (1 + np.abs(href).sum())
        dw.free()


def test_freed_blocks_are_reused(gpu):
    """rp_free keeps a block for the next rp_malloc of the same size (no
    driver clear of freed memory in between), counts it as free memory, and
    hands out fresh memory for other sizes; data written through the reused
    block reads back intact."""
    hip = gpu.hip()
    nbytes = (48 << 20) + 8 * 1237      # a size no other test allocates
    p = ctypes.c_void_p()
    gpu.check(hip.rp_malloc(ctypes.byref(p), nbytes))
    first = p.value
    free0, total = ctypes.c_size_t(), ctypes.c_size_t()
    gpu.check(hip.rp_mem_get_info(ctypes.byref(free0), ctypes.byref(total)))
    gpu.check(hip.rp_free(p))
    free1 = ctypes.c_size_t()
    gpu.check(hip.rp_mem_get_info(ctypes.byref(free1), ctypes.byref(total)))
    # the kept block counts as free again
    assert free1.value >= free0.value + nbytes - (2 << 20)
    q = ctypes.c_void_p()
    gpu.check(hip.rp_malloc(ctypes.byref(q), nbytes))
    assert q.value == first
    r = ctypes.c_void_p()
    gpu.check(hip.rp_malloc(ctypes.byref(r), nbytes + 4096))
    assert r.value not in (0, first)
    a = np.arange(nbytes // 8, dtype=np.float64)
    gpu.check(hip.rp_memcpy(q, a.ctypes.data_as(ctypes.c_void_p), nbytes, ra.RP_H2D))
    b = np.empty_like(a)
    gpu.check(hip.rp_memcpy(b.ctypes.data_as(ctypes.c_void_p), q, nbytes, ra.RP_D2H))
    assert np.array_equal(a, b)
    gpu.check(hip.rp_free(q))
    gpu.check(hip.rp_free(r))


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_lincomb_beyond_64_vectors(gpu, dtype):
    """rp_lincomb takes any number of vectors (16 per sweep): k = 80, the
    Lanczos ground state's combination at a Krylov dimension of 80."""
    import reapers_amd as ra
    rng = np.random.default_rng(80)
    n, k = 3001, 80
    p = gpu.RP_C64 if dtype == np.complex64 else gpu.RP_C128
    Vh = [_rand(rng, n, dtype) for _ in range(k)]
    Vd = [DeviceBuffer.from_array(v) for v in Vh]
    ptrs = (ctypes.c_void_p * k)(*[b.ptr.value for b in Vd])
    c = rng.standard_normal(k) + 1j * rng.standard_normal(k)
    cc = np.empty(2 * k); cc[0::2] = c.real; cc[1::2] = c.imag
    y = _rand(rng, n, dtype)
    dy = DeviceBuffer.from_array(y)
    ra.check(gpu.hip().rp_lincomb(p, dy.ptr, ptrs, k, cc.ctypes.data_as(ctypes.c_void_p), n))
    yref = y.astype(complex) + sum(c[i] * Vh[i].astype(complex) for i in range(k))
    tol = 2e-6 * (1 + np.abs(c).sum()) if dtype == np.complex64 else 1e-12 * (1 + np.abs(c).sum())
    assert np.abs(dy.to_array(dtype, n) - yref).max() <= tol
    for b in Vd + [dy]:
        b.free()
