"""The benchmark's own size: sparse SYK k=8 at N=48 (24 qubits, 2^24
amplitudes), where one oracle apply of the whole state takes too long for a
test.  Parity on windows of output rows against the oracle's row-window
restatement of CPUImpl::apply_ops (inc/cpuimpl.h:737-754) -- the first and
last tiles, a window across the middle and random windows -- for FP64, FP32
with FP64 sums and the benchmark's FP32 upload (autotuned tiling, group
accumulation); then size-independent properties of the whole state:
linearity of the apply, Hermiticity <phi|H psi> = conj(<psi|H phi>), the
norm kept by exp(-iHt) and its time reversal.  And the OTOC workload's
size, N=60 (30 qubits), on row windows against the oracle's FP32 apply.

Tolerances: rows FP64 1e-11, FP32 1e-5 (FP64 sums) and 4e-7 (group
accumulation) of the window's largest amplitude, the bounds of the smaller
apply tests; linearity and Hermiticity 1e-12 relative (FP64); norm 1e-6 and
time reversal 1e-5 for FP32 states (tol 1e-7 per evolution)."""
import ctypes

import numpy as np
import pytest

from gpu_util import DeviceBuffer, Terms, prec_of

pytestmark = pytest.mark.gpu

N, L = 48, 24
DIM = 1 << L


def _state(rng, dtype):
    psi = rng.standard_normal(DIM) + 1j * rng.standard_normal(DIM)
    return (psi / np.linalg.norm(psi)).astype(dtype)


def _windows(rng):
    w = 4096
    starts = [0, DIM // 2 - w // 2, DIM - w] + [int(s) for s in rng.integers(0, DIM - w, 2)]
    return [(s, s + w) for s in starts]


def _gpu_apply(gpu, terms, psi):
    vin = DeviceBuffer.from_array(psi)
    vout = DeviceBuffer(psi.nbytes)
    gpu.check(gpu.hip().rp_apply_ops(terms.h, vout.ptr, vin.ptr, 0))
    out = vout.to_array(psi.dtype, len(psi))
    vin.free()
    vout.free()
    return out


@pytest.mark.parametrize("mode", ["fp64", "fp32", "fp32-group"])
def test_fullsize_row_windows_match_oracle(gpu, orc, mode):
    fp = 64 if mode == "fp64" else 32
    dtype = np.complex128 if fp == 64 else np.complex64
    rng = np.random.default_rng(48)
    bits, coeff = orc.syk_terms(N, 8.0, seed=2000, fp=fp)
    if fp == 32:
        coeff = coeff.astype(np.complex64).astype(np.complex128)
    psi = _state(rng, dtype)
    if mode == "fp32-group":
        terms = Terms(L, bits, coeff, gpu.RP_C64,
                      flags=gpu.RP_TERMS_AUTOTUNE | gpu.RP_TERMS_ACC_GROUP)
        tol = 4e-7
    else:
        terms = Terms(L, bits, coeff, prec_of(dtype))
        tol = 1e-11 if fp == 64 else 1e-5
    info = terms.info()
    got = _gpu_apply(gpu, terms, psi)
    terms.free()
    psi64 = psi.astype(np.complex128)
    for lo, hi in _windows(rng):
        ref = orc.apply_rows(L, bits, coeff, psi64, lo, hi)
        err = np.abs(got[lo:hi].astype(np.complex128) - ref).max() / np.abs(ref).max()
        print("%s tiling (%d,%d,%d) rows [%d,%d): %.2e" % (mode, info.tile_bits, info.threads,
                                                          info.depth, lo, hi, err), flush=True)
        assert err <= tol, (lo, err)


def test_fullsize_linearity_and_hermiticity(gpu, orc):
    rng = np.random.default_rng(4848)
    bits, coeff = orc.syk_terms(N, 8.0, seed=2000, fp=64)
    terms = Terms(L, bits, coeff, prec_of(np.complex128))
    psi = _state(rng, np.complex128)
    phi = _state(rng, np.complex128)
    a, b = 0.3 - 1.1j, -0.7 + 0.2j
    hpsi = _gpu_apply(gpu, terms, psi)
    hphi = _gpu_apply(gpu, terms, phi)
    hmix = _gpu_apply(gpu, terms, a * psi + b * phi)
    lin = np.abs(hmix - (a * hpsi + b * hphi)).max() / np.abs(hmix).max()
    # H is Hermitian (SYK: i^2 chi chi chi chi with real couplings)
    x, y = np.vdot(phi, hpsi), np.conj(np.vdot(psi, hphi))
    herm = abs(x - y) / max(np.linalg.norm(hpsi), 1e-300)
    print("linearity %.2e, Hermiticity %.2e" % (lin, herm), flush=True)
    terms.free()
    assert lin < 1e-12
    assert herm < 1e-12


def test_fullsize_evolution_norm_and_time_reversal(gpu, ra, orc):
    """exp(-iHt) of the benchmark's FP32 states: the norm stays 1 to FP32
    accuracy after t = 10, and evolving back by -t returns the state."""
    bits, coeff = ra.syk_terms(N, 8.0, seed=2000, fp=32)
    ra.set_upload_flags(ra.RP_TERMS_AUTOTUNE | ra.RP_TERMS_ACC_GROUP)
    try:
        H = ra.Hamiltonian(L, bits, coeff, fp=32)
    finally:
        ra.set_upload_flags(0)
    s = ra.DeviceState(L, fp=32)
    s.random(2001)
    psi0 = s.store()
    st = s.evolve(H, 10.0, krydim=20, tol=1e-7)
    n1 = s.norm()
    s.evolve(H, -10.0, krydim=20, tol=1e-7)
    back = s.store().astype(np.complex128)
    dev = np.linalg.norm(back - psi0.astype(np.complex128))
    print("norm after t=10: %.9f (%d applies); |psi(-t)(t) - psi| = %.2e"
          % (n1, st.matvecs, dev), flush=True)
    s.close()
    H.close()
    assert abs(n1 - 1.0) < 1e-6
    assert dev < 1e-5


@pytest.mark.parametrize("NM,seed", [(60, 4000), (62, 6200), (64, 6400)])
def test_otoc_size_row_windows_match_oracle(gpu, orc, NM, seed):
    """30 qubits (sparse SYK N=60, the OTOC workload's 8.6 GB complex<float>
    state, indices up to 2^30), 31 (N=62, 17 GB, indices past the signed
    32-bit range) and 32, the largest chain the API takes (N=64, 34 GB):
    H psi with the workload's group-accumulating upload,
    compared on windows of rows (one at the top of the index range) with the
    oracle's FP32 apply, the reference's FP32 CPU arithmetic (an FP64 copy
    of the state would double the host memory)."""
    L60 = NM // 2
    dim = 1 << L60
    bits, coeff = orc.syk_terms(NM, 8.0, seed=seed, fp=32)
    coeff = coeff.astype(np.complex64).astype(np.complex128)
    hip = gpu.hip()
    vin = DeviceBuffer(8 * dim)
    vout = DeviceBuffer(8 * dim)
    gpu.check(hip.rp_init_random(gpu.RP_C64, vin.ptr, dim, seed + 1))
    psi = vin.to_array(np.complex64, dim)
    terms = Terms(L60, bits, coeff, gpu.RP_C64, flags=gpu.RP_TERMS_ACC_GROUP)
    gpu.check(hip.rp_apply_ops(terms.h, vout.ptr, vin.ptr, 0))
    terms.free()
    vin.free()
    rng = np.random.default_rng(NM)
    w = 4096
    starts = [0, dim // 2 - w // 2, dim - w] + [int(s) for s in rng.integers(0, dim - w, 2)]
    for lo in starts:
        got = np.empty(w, np.complex64)
        gpu.check(hip.rp_memcpy(got.ctypes.data_as(ctypes.c_void_p),
                                ctypes.c_void_p(vout.ptr.value + 8 * lo), 8 * w, gpu.RP_D2H))
        ref = orc.apply_rows(L60, bits, coeff, psi, lo, lo + w, fp=32)
        err = np.abs(got.astype(np.complex128) - ref).max() / np.abs(ref).max()
        print("N=%d rows [%d,%d): %.2e" % (NM, lo, lo + w, err), flush=True)
        assert err <= 1e-5, (lo, err)
    vout.free()
