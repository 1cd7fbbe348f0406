"""rp_apply_gs (the Arnoldi step's apply with its first Gram-Schmidt pass's
coefficients formed in the apply kernel's epilogue) against the separate
calls it replaces (rp_apply_ops then rp_gs_pass_fused) and numpy, through the
C ABI: h = V^H (H v), the updated vector, its norm and the fused
normalization, for basis sizes across the epilogue's chunks of 8 vectors, with
the fused kernel (complex<float>, 2^12 tiles, FP64 or group accumulation)
and with the fallback (complex<double>, other tilings).

Tolerances: the two paths differ only in how the FP64 partial sums are
grouped, so h and the norms agree to 1e-12 relative and the vectors to
1e-6 (FP32) of their largest entry; against numpy 1e-5 relative (FP32)."""
import ctypes

import numpy as np
import pytest

from gpu_util import DeviceBuffer, Terms

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _orthonormal(rng, n, k, dtype):
    a = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    q, _ = np.linalg.qr(a)
    return q.astype(dtype)


def _call(gpu, fused, terms, vin, V, k, n, dtype, fuse_min):
    hip = gpu.hip()
    out = DeviceBuffer(n * np.dtype(dtype).itemsize)
    arr = (_vp * max(k, 1))(*([b.ptr.value for b in V[:k]] or [None]))
    h = np.zeros(2 * max(k, 1))
    norm, nz = ctypes.c_double(), ctypes.c_int()
    if fused:
        gpu.check(hip.rp_apply_gs(terms.h, out.ptr, vin.ptr, arr, k, h.ctypes.data_as(_vp),
                                  ctypes.byref(norm), fuse_min, ctypes.byref(nz)))
    else:
        prec = gpu.RP_C64 if dtype == np.complex64 else gpu.RP_C128
        gpu.check(hip.rp_apply_ops(terms.h, out.ptr, vin.ptr, 0))
        gpu.check(hip.rp_gs_pass_fused(prec, arr, k, out.ptr, n, h.ctypes.data_as(_vp),
                                       ctypes.byref(norm), fuse_min, ctypes.byref(nz)))
    w = out.to_array(dtype, n)
    out.free()
    return h[:2 * k].view(np.complex128), norm.value, nz.value, w


@pytest.mark.parametrize("fp,flags", [(32, 4), (32, 0), (64, 0)])
@pytest.mark.parametrize("k,fuse_min", [(1, 0.0), (8, 0.5), (9, 0.0), (17, 0.5)])
def test_apply_gs_matches_separate_calls(gpu, orc, fp, flags, k, fuse_min):
    N, L = 36, 18
    n = 1 << L
    dtype = np.complex64 if fp == 32 else np.complex128
    rng = np.random.default_rng(100 * k + fp)
    bits, coeff = orc.syk_terms(N, 8.0, seed=91, fp=fp)
    if fp == 32:
        coeff = coeff.astype(np.complex64).astype(np.complex128)
    terms = Terms(L, bits, coeff, gpu.RP_C64 if fp == 32 else gpu.RP_C128, flags=flags)
    info = terms.info()
    Q = _orthonormal(rng, n, k + 1, dtype)
    V = [DeviceBuffer.from_array(np.ascontiguousarray(Q[:, q])) for q in range(k)]
    # v = the last basis vector, as in an Arnoldi step (v is V[k-1]'s successor)
    vin = V[k - 1]
    hf, nf, zf, wf = _call(gpu, True, terms, vin, V, k, n, dtype, fuse_min)
    hs, ns, zs, ws = _call(gpu, False, terms, vin, V, k, n, dtype, fuse_min)
    print("fp%d flags %d tiling (%d,%d,%d) k=%d: |dh| %.1e, norms %.12g / %.12g, normalized %d/%d"
          % (fp, flags, info.tile_bits, info.threads, info.depth, k, np.abs(hf - hs).max(), nf, ns,
             zf, zs), flush=True)
    assert zf == zs
    assert np.abs(hf - hs).max() <= 1e-12 * np.abs(hs).max()
    assert abs(nf - ns) <= 1e-12 * ns
    tol_w = 1e-6 if fp == 32 else 1e-13
    assert np.abs(wf.astype(np.complex128) - ws).max() <= tol_w * np.abs(ws).max()
    # numpy: h = V^H H v, w = (H v - V h) (/ beta)
    Q64 = Q[:, :k].astype(np.complex128)
    hv = orc.apply(L, bits, coeff, Q[:, k - 1].astype(np.complex128))
    h_ref = Q64.conj().T @ hv
    r = hv - Q64 @ h_ref
    if zf:
        r /= np.linalg.norm(r)
    tol = 1e-5 if fp == 32 else 1e-11
    assert np.abs(hf - h_ref).max() <= tol * np.abs(h_ref).max()
    assert np.abs(wf.astype(np.complex128) - r).max() <= tol * np.abs(r).max()
    for b in V:
        b.free()
    terms.free()


def test_krylov_with_fused_first_pass_vs_oracle(gpu, orc):
    """A whole FP32 evolution through Krylov<GPUImpl> with the fused first
    pass switched on (its Arnoldi steps then use rp_apply_gs on the
    group-accumulating upload), against the oracle's FP64 evolution of the
    same couplings."""
    N, L = 40, 20
    bits, coeff = gpu.syk_terms(N, 8.0, seed=404, fp=32)
    gpu.set_upload_flags(gpu.RP_TERMS_ACC_GROUP)
    try:
        H = gpu.Hamiltonian(L, bits, coeff, fp=32)
    finally:
        gpu.set_upload_flags(0)
    s = gpu.DeviceState(L, fp=32)
    s.random(405)
    psi = s.store()
    prev = gpu.set_fused_first_pass(True)
    try:
        st = s.evolve(H, 4.0, 0.0, krydim=20, tol=1e-7)
    finally:
        gpu.set_fused_first_pass(prev)
    got = s.store().astype(np.complex128)
    ref, _ = orc.krylov_evolve(L, bits, coeff, psi.astype(np.complex128), 4.0, 0.0, krydim=20,
                               tol=1e-12)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("N=40 FP32 evolution with the fused pass: %d matvecs, rel err %.2e" % (st.matvecs, err),
          flush=True)
    s.close()
    H.close()
    assert err < 1e-5
