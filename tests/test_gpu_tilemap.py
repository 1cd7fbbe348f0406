"""Every instantiated tiling of the tiled apply at 22 qubits, the size from
which each of them launches its workgroups through the output-tile map (at
least p + 3 tile-index bits for its 2^p workgroups resident per XCD) and
walks the groups in the map's class order: complex<float> with FP64 and
with group accumulation, complex<double>, a plain and an accumulating apply,
against the oracle's FP64 apply.

Tolerances as the other apply tests: 2e-6 of the largest amplitude in FP32
(FP32 inputs and outputs; group accumulation rounds partial sums of four
tile groups to FP32), 1e-13 in FP64."""
import ctypes

import numpy as np
import pytest

from conftest import random_pauli_sum, random_state

pytestmark = pytest.mark.gpu

L = 22
FP32_TILINGS = [(12, 512, 2), (12, 512, 3), (12, 512, 4), (12, 1024, 2), (13, 1024, 2),
                (10, 128, 2), (8, 128, 2)]
FP64_TILINGS = [(12, 512, 2), (11, 256, 4), (10, 128, 2), (8, 128, 2)]


@pytest.fixture(scope="module")
def case(orc):
    """Sparse SYK N=44 (22 qubits) plus 400 general terms on at most 8 sites
    (complex coefficients; with them every tiling's map beats the contiguous
    order, so every tiling keeps its map) and 20 terms sharing a flip mask
    with one of those; the oracle's apply for the FP64 and the FP32-rounded
    coefficients."""
    rng = np.random.default_rng(2222)
    b1, c1 = orc.syk_terms(2 * L, 8.0, seed=2222, fp=64)
    b2, c2 = random_pauli_sum(rng, L, 400, max_sites=8)
    b3 = b2[:20] ^ np.uint64(2 << 10)      # the Z bit of site 5 toggled: same x
    c3 = rng.standard_normal(20) + 1j * rng.standard_normal(20)
    bits = np.concatenate([b1, b2, b3]).astype(np.uint64)
    coeff = np.concatenate([c1, c2, c3])
    psi = random_state(rng, L)
    prior = random_state(rng, L)
    c32 = coeff.astype(np.complex64).astype(np.complex128)
    p32 = psi.astype(np.complex64).astype(np.complex128)
    refs = {64: orc.apply(L, bits, coeff, psi, fp=64), 32: orc.apply(L, bits, c32, p32, fp=64)}
    return bits, coeff, psi, prior, refs


def run(gpu, orc, case, fp, flags, tiling):
    from gpu_util import DeviceBuffer
    bits, coeff, psi, prior, refs = case
    dt = np.complex64 if fp == 32 else np.complex128
    if fp == 32:
        coeff = coeff.astype(np.complex64).astype(np.complex128)
    psi, prior = psi.astype(dt), prior.astype(dt)
    hip = gpu.hip()
    h = ctypes.c_void_p()
    b = np.ascontiguousarray(bits, np.uint64)
    co = np.empty(2 * len(b))
    co[0::2], co[1::2] = coeff.real, coeff.imag
    gpu.check(hip.rp_terms_upload_ex(ctypes.byref(h), gpu.RP_C64 if fp == 32 else gpu.RP_C128,
                                     L, b.ctypes.data_as(ctypes.c_void_p),
                                     co.ctypes.data_as(ctypes.c_void_p), len(b), flags))
    try:
        gpu.check(hip.rp_terms_set_tiling(h, *tiling))
        info = gpu.TermsInfo()
        gpu.check(hip.rp_terms_info(h, ctypes.byref(info)))
        assert info.tile_bits == tiling[0] and info.tile_map == L - tiling[0], info.tile_map
        ref = refs[fp]
        vin = DeviceBuffer.from_array(psi)
        vout = DeviceBuffer(psi.nbytes)
        gpu.check(hip.rp_apply_ops(h, vout.ptr, vin.ptr, 0))
        got = vout.to_array(dt, len(psi)).astype(np.complex128)
        err = np.abs(got - ref).max() / np.abs(ref).max()
        acc = DeviceBuffer.from_array(prior)
        gpu.check(hip.rp_apply_ops(h, acc.ptr, vin.ptr, 1))
        want = prior.astype(np.complex128) + ref
        got2 = acc.to_array(dt, len(psi)).astype(np.complex128)
        err2 = np.abs(got2 - want).max() / np.abs(want).max()
        return err, err2, info.kernel
    finally:
        gpu.check(hip.rp_terms_free(h))


@pytest.mark.parametrize("tiling", FP32_TILINGS)
@pytest.mark.parametrize("group_acc", [False, True])
def test_fp32_every_tiling_with_the_map(gpu, orc, case, tiling, group_acc):
    flags = gpu.RP_TERMS_ACC_GROUP if group_acc else 0
    err, err2, kernel = run(gpu, orc, case, 32, flags, tiling)
    print("fp32 %s group_acc=%d kernel %d: %.2e / %.2e" % (tiling, group_acc, kernel, err, err2),
          flush=True)
    assert err < 2e-6 and err2 < 2e-6


@pytest.mark.parametrize("tiling", FP64_TILINGS)
def test_fp64_every_tiling_with_the_map(gpu, orc, case, tiling):
    err, err2, kernel = run(gpu, orc, case, 64, 0, tiling)
    print("fp64 %s: %.2e / %.2e" % (tiling, err, err2), flush=True)
    assert err < 1e-13 and err2 < 1e-13
