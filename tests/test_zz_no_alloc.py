"""rp_gs_pass_fused_noalloc (the Gram-Schmidt pass whose partials live in the
library's fixed device array instead of its allocated workspace) against
rp_gs_pass_fused through the C ABI: identical h, norm, normalized flag and
updated vector, bit for bit (same kernels, same launch shapes, same
reduction order), on two calls in a row, the second smaller; both also
against numpy; the device's free memory is unchanged by the call; k beyond
the array's 64 vectors is rejected."""
import ctypes

import numpy as np
import pytest

from gpu_util import DeviceBuffer, prec_of

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p


def _basis(rng, n, k, dtype):
    """k orthonormal columns of length n and a vector w with a component
    outside their span."""
    a = rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k))
    q, _ = np.linalg.qr(a)
    w = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    return q.astype(dtype), w.astype(dtype)


def _free_bytes(gpu):
    free, tot = ctypes.c_size_t(), ctypes.c_size_t()
    gpu.check(gpu.hip().rp_mem_get_info(ctypes.byref(free), ctypes.byref(tot)))
    return free.value


def _run(gpu, fn, prec, vbufs, k, w_host, fuse_min):
    """One pass on a device copy of w_host; also the device's free bytes
    just before and just after the call."""
    hip_arr = (_vp * max(k, 1))(*[b.ptr.value for b in vbufs[:k]] or [None])
    wb = DeviceBuffer.from_array(w_host)
    h = np.zeros(2 * max(k, 1))
    norm = ctypes.c_double()
    nz = ctypes.c_int()
    gpu.check(gpu.hip().rp_synchronize())
    free0 = _free_bytes(gpu)
    st = fn(prec, hip_arr, k, wb.ptr, len(w_host), h.ctypes.data_as(_vp), ctypes.byref(norm),
            fuse_min, ctypes.byref(nz))
    free1 = _free_bytes(gpu)
    w_out = wb.to_array(w_host.dtype, len(w_host))
    wb.free()
    return st, h[:2 * k].view(np.complex128), norm.value, nz.value, w_out, free1 - free0


@pytest.mark.parametrize("dtype", [np.complex64, np.complex128])
def test_noalloc_gs_pass_matches_allocating_pass(gpu, dtype):
    hip = gpu.hip()
    prec = prec_of(dtype)
    rng = np.random.default_rng(707)
    for n, k, fuse_min in ((1_000_003, 9, 0.5), (4099, 3, 0.0)):
        V, w = _basis(rng, n, k, dtype)
        vbufs = [DeviceBuffer.from_array(np.ascontiguousarray(V[:, q])) for q in range(k)]
        # the allocating pass first: it also has the runtime load the sweep
        # kernels, whose first launch takes device memory of its own
        st_a, h_a, nrm_a, nz_a, w_a, _ = _run(gpu, hip.rp_gs_pass_fused, prec, vbufs, k, w,
                                              fuse_min)
        st_n, h_n, nrm_n, nz_n, w_n, dfree = _run(gpu, hip.rp_gs_pass_fused_noalloc, prec, vbufs,
                                                  k, w, fuse_min)
        for b in vbufs:
            b.free()
        assert st_n == 0 and st_a == 0
        print("n=%d k=%d fuse_min=%g: norm %.15g (normalized %d), free bytes changed by %d"
              % (n, k, fuse_min, nrm_n, nz_n, dfree), flush=True)
        assert dfree == 0
        assert np.array_equal(h_n, h_a)
        assert nrm_n == nrm_a and nz_n == nz_a
        assert np.array_equal(w_n, w_a)
        # and against numpy
        V64, w64 = V.astype(np.complex128), w.astype(np.complex128)
        h_ref = V64.conj().T @ w64
        r = w64 - V64 @ h_ref
        tol = 1e-12 if dtype == np.complex128 else 1e-5
        assert np.abs(h_n - h_ref).max() <= tol * np.abs(h_ref).max()
        if nz_n:
            r = r / np.linalg.norm(r)
        assert np.abs(w_n.astype(np.complex128) - r).max() <= tol * max(np.abs(r).max(), 1e-300)


def test_noalloc_rejects_more_vectors_than_its_scratch(gpu):
    hip = gpu.hip()
    b = DeviceBuffer.from_array(np.zeros(64, np.complex128))
    arr = (_vp * 65)(*([b.ptr.value] * 65))
    h = np.zeros(130)
    norm, nz = ctypes.c_double(), ctypes.c_int()
    st = hip.rp_gs_pass_fused_noalloc(gpu.RP_C128, arr, 65, b.ptr, 64, h.ctypes.data_as(_vp),
                                      ctypes.byref(norm), 0.0, ctypes.byref(nz))
    b.free()
    assert st == gpu.RP_ERR_INVALID_ARGUMENT
    assert b"64" in hip.rp_last_error()
