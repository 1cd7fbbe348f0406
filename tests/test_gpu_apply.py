"""GPU parity of the Pauli-sum apply (rp_apply_ops, through the C ABI)
against the oracle's restatement of CPUImpl::apply_ops
(inc/cpuimpl.h:737-754).

Tolerances: FP64 max|diff| <= 1e-11 * max|ref| ; FP32 <= 1e-5 * max|ref|
(relative to the largest output amplitude; the oracle runs in FP64 on the
same FP32-representable inputs)."""
import numpy as np
import pytest

from conftest import random_pauli_sum, random_state
from gpu_util import gpu_apply

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-11, np.complex64: 1e-5}


def _check(orc, length, bits, coeff, psi, acc=None):
    dtype = psi.dtype.type
    coeff = coeff.astype(np.complex64).astype(np.complex128) if dtype is np.complex64 else coeff
    ref = orc.apply(length, bits, coeff, psi.astype(np.complex128),
                    out=None if acc is None else acc.astype(np.complex128))
    got = gpu_apply(length, bits, coeff, psi, accumulate_into=acc)
    scale = max(np.abs(ref).max(), 1e-300)
    err = np.abs(got.astype(np.complex128) - ref).max() / scale
    print("len=%d terms=%d dtype=%s relerr=%.3e" % (length, len(bits), dtype.__name__, err), flush=True)
    assert err <= TOL[dtype], err


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("length", [1, 3, 7, 8, 10, 13, 14, 17, 18, 20])
def test_random_pauli_sums(gpu, orc, length, dtype):
    rng = np.random.default_rng(1000 + length)
    bits, coeff = random_pauli_sum(rng, length, 37)
    psi = random_state(rng, length, dtype)
    _check(orc, length, bits, coeff, psi)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_accumulate_semantics(gpu, orc, dtype):
    """res += H vec (the reference's apply_ops semantics)."""
    rng = np.random.default_rng(7)
    length = 18
    bits, coeff = random_pauli_sum(rng, length, 20, max_sites=4)
    psi = random_state(rng, length, dtype)
    acc = random_state(rng, length, dtype)
    _check(orc, length, bits, coeff, psi, acc=acc)


@pytest.mark.parametrize("N,sp", [(24, 0.0), (32, 8.0), (40, 8.0), (44, 4.0)])
def test_syk_hamiltonians(gpu, orc, N, sp):
    """SYK term lists from the oracle's generator (dense SYK at N=24 is the
    reference's terms >= 2^len case, which it routes to a dense gemv)."""
    for dtype, fp in ((np.complex128, 64), (np.complex64, 32)):
        bits, coeff = orc.syk_terms(N, sp, seed=11, fp=fp)
        rng = np.random.default_rng(N)
        psi = random_state(rng, N // 2, dtype)
        _check(orc, N // 2, bits, coeff, psi)


def test_empty_and_identity(gpu, orc):
    rng = np.random.default_rng(3)
    psi = random_state(rng, 12, np.complex128)
    got = gpu_apply(12, np.zeros(0, np.uint64), np.zeros(0, complex), psi)
    assert np.all(got == 0)
    # identity term (bits 0) times 2.5: a pure scaling
    got = gpu_apply(12, np.array([0], np.uint64), np.array([2.5 + 0j]), psi)
    assert np.abs(got - 2.5 * psi).max() < 1e-15


def test_merged_flips_dense(gpu, orc):
    """Many terms sharing one flip mask (the per-x merge in the kernel)."""
    rng = np.random.default_rng(5)
    length = 16
    bits = []
    x = 0b1011001110100101
    for z in rng.integers(0, 1 << length, size=64):
        b = 0
        for s in range(length):
            b |= (((x >> s) & 1) | (((int(z) >> s) & 1) << 1)) << (2 * s)
        bits.append(b)
    bits = np.array(bits, np.uint64)
    coeff = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    psi = random_state(rng, length, np.complex128)
    _check(orc, length, bits, coeff, psi)


@pytest.mark.parametrize("fp,L", [(32, 16), (64, 14), (32, 21)])
def test_autotuned_upload_matches_oracle(gpu, orc, fp, L):
    """RP_TERMS_AUTOTUNE times the instantiated tilings at upload; whichever
    it keeps must give the oracle's result."""
    rng = np.random.default_rng(L + fp)
    bits, coeff = orc.syk_terms(2 * L, 8.0, seed=70 + L, fp=fp)
    psi = random_state(rng, L)
    gpu.set_upload_flags(gpu.RP_TERMS_AUTOTUNE)
    try:
        ham = gpu.Hamiltonian(L, bits, coeff, fp=fp)
        s = gpu.DeviceState(L, fp=fp)
        s.load(psi.astype(np.complex64 if fp == 32 else np.complex128))
        s.apply(ham)
        info = ham.info()
    finally:
        gpu.set_upload_flags(0)
    ref = orc.apply(L, bits, coeff, psi, fp=64)
    got = s.store().astype(np.complex128)
    print("autotuned L=%d fp%d: tile bits %d" % (L, fp, info.tile_bits), flush=True)
    assert (info.tile_bits, info.threads, info.depth) in [
        (0, 256, 1), (0, 256, 8), (0, 256, 32), (0, 256, 128),
        (8, 128, 2), (10, 128, 2), (11, 256, 4), (12, 512, 2), (12, 512, 3), (12, 512, 4),
        (12, 1024, 2), (13, 1024, 2)]
    assert np.abs(got - ref).max() < (1e-5 if fp == 32 else 1e-12) * np.abs(ref).max()


SMALL = [(0, 256, 1), (0, 256, 8), (0, 256, 32), (0, 256, 128)]   # terms split in slices
TILINGS = {32: SMALL + [(8, 128, 2), (10, 128, 2), (12, 512, 2), (12, 512, 3), (12, 512, 4),
                        (12, 1024, 2), (13, 1024, 2)],
           64: SMALL + [(8, 128, 2), (10, 128, 2), (11, 256, 4), (12, 512, 2)]}


@pytest.mark.parametrize("fp", [32, 64])
def test_every_tiling_matches_oracle(gpu, orc, fp):
    """Every instantiated (tile bits, threads, pipeline depth) of the tiled
    kernel and every term-slice count of the small kernel (tile bits 0), set
    through rp_terms_set_tiling, on SYK (real/imaginary subs, many groups)
    and a general sum (complex and multi-term subs), overwriting and
    accumulating."""
    import ctypes
    from gpu_util import DeviceBuffer, Terms, prec_of
    rng = np.random.default_rng(fp)
    dt = np.complex64 if fp == 32 else np.complex128
    cases = [(16, *orc.syk_terms(32, 8.0, seed=17, fp=fp))]
    b, c = random_pauli_sum(rng, 15, 60)
    cases.append((15, np.concatenate([b, b[:7]]), np.concatenate([c, rng.standard_normal(7)])))
    for L, bits, coeff in cases:
        psi = random_state(rng, L, dt)
        ref = orc.apply(L, bits, coeff, psi.astype(np.complex128), fp=64)
        t = Terms(L, bits, coeff, prec_of(dt))
        vin = DeviceBuffer.from_array(psi)
        vout = DeviceBuffer(psi.nbytes)
        for tb, nt, d in TILINGS[fp]:
            gpu.check(gpu.hip().rp_terms_set_tiling(t.h, tb, nt, d))
            info = t.info()
            assert (info.tile_bits, info.threads, info.depth) == (tb, nt, d)
            gpu.check(gpu.hip().rp_apply_ops(t.h, vout.ptr, vin.ptr, 0))
            got = vout.to_array(dt, len(psi)).astype(np.complex128)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            print("L=%d fp%d tiling %s: %.2e" % (L, fp, (tb, nt, d), err), flush=True)
            assert err < (1e-6 if fp == 32 else 1e-13)
            # accumulate: res += H vec on top of the previous result
            gpu.check(gpu.hip().rp_apply_ops(t.h, vout.ptr, vin.ptr, 1))
            got2 = vout.to_array(dt, len(psi)).astype(np.complex128)
            assert np.abs(got2 - 2 * ref).max() / np.abs(ref).max() < (2e-6 if fp == 32 else 1e-13)
        # a tiling that is not instantiated is an error, not a launch
        assert gpu.hip().rp_terms_set_tiling(t.h, 9, 128, 2) != 0
        vin.free(); vout.free(); t.free()


@pytest.mark.parametrize("tiling", [(12, 512), (12, 1024), (13, 1024), (10, 128)])
def test_group_accumulation_matches_oracle(gpu, orc, tiling):
    """RP_TERMS_ACC_GROUP (FP32 sums within a tile group, FP64 across
    groups) against the FP64 oracle: as accurate as the FP32 input allows."""
    from gpu_util import DeviceBuffer, prec_of
    import ctypes
    L = 18
    rng = np.random.default_rng(sum(tiling))
    bits, coeff = orc.syk_terms(36, 8.0, seed=3, fp=32)
    psi = random_state(rng, L, np.complex64)
    ref = orc.apply(L, bits, coeff, psi.astype(np.complex128), fp=64)
    b = np.ascontiguousarray(bits, np.uint64)
    co = np.empty(2 * len(b))
    co[0::2], co[1::2] = np.real(coeff), np.imag(coeff)
    h = ctypes.c_void_p()
    gpu.check(gpu.hip().rp_terms_upload_ex(ctypes.byref(h), gpu.RP_C64, L,
                                           b.ctypes.data_as(ctypes.c_void_p),
                                           co.ctypes.data_as(ctypes.c_void_p), len(b),
                                           gpu.RP_TERMS_ACC_GROUP))
    gpu.check(gpu.hip().rp_terms_set_tiling(h, tiling[0], tiling[1], 2))
    vin = DeviceBuffer.from_array(psi)
    vout = DeviceBuffer(psi.nbytes)
    gpu.check(gpu.hip().rp_apply_ops(h, vout.ptr, vin.ptr, 0))
    got = vout.to_array(np.complex64, len(psi)).astype(np.complex128)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print("group accumulation %s: %.2e" % (tiling, err), flush=True)
    # the FP32 partial sums hold ~11 terms (four tile groups) before they
    # join the FP64 sum: at most ~11 roundings of half an FP32 ulp (6e-8) of
    # the largest amplitude each; FP32 accumulation throughout gives ~1e-6
    assert err < 6e-7
    gpu.check(gpu.hip().rp_terms_free(h))


@pytest.mark.parametrize("fp", [32, 64])
def test_default_split_tiling_dense_syk24(gpu, orc, fp):
    """Dense SYK N=24 (10626 strings on 4096 amplitudes) gets the term-split
    small kernel by default; its result matches the oracle."""
    L = 12
    bits, coeff = orc.syk_terms(24, 0.0, seed=5, fp=fp)
    psi = random_state(np.random.default_rng(9), L)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=fp)
    info = ham.info()
    assert (info.tile_bits, info.threads) == (0, 256) and info.depth > 1
    s = gpu.DeviceState(L, fp=fp)
    s.load(psi.astype(np.complex64 if fp == 32 else np.complex128))
    s.apply(ham)
    ref = orc.apply(L, bits, coeff, psi, fp=64)
    err = np.abs(s.store().astype(np.complex128) - ref).max() / np.abs(ref).max()
    print("dense SYK N=24 fp%d, %d slices: %.2e" % (fp, info.depth, err), flush=True)
    assert err < (1e-6 if fp == 32 else 1e-13)


@pytest.mark.parametrize("tiling", [(12, 512, 2), (12, 1024, 2), (13, 1024, 2), (10, 128, 2),
                                    (8, 128, 2)])
def test_group_accumulation_every_sub_class(gpu, orc, tiling):
    """RP_TERMS_ACC_GROUP on a term list with every kind of sub the tiled
    tables store: single terms with real, imaginary and complex
    coefficients, with and without phase bits among the row bits, and flip
    masks shared by several terms; a plain and an accumulating apply
    against the FP64 oracle."""
    from gpu_util import DeviceBuffer
    import ctypes
    L = 18
    rng = np.random.default_rng(77 + tiling[0] + tiling[1] + tiling[2])
    bits, coeff = random_pauli_sum(rng, L, 240, max_sites=4)
    coeff[0::3] = coeff[0::3].real
    coeff[1::3] = 1j * coeff[1::3].imag
    # 60 terms that share a flip mask with an earlier one: the Z bit of one
    # of its sites toggled (X <-> Y, I <-> Z) keeps x and changes z
    extra = []
    for i in rng.choice(len(bits), 60, replace=False):
        s = int(rng.integers(0, L))
        extra.append(int(bits[i]) ^ (2 << (2 * s)))
    bits = np.concatenate([bits, np.array(extra, np.uint64)])
    coeff = np.concatenate([coeff, rng.standard_normal(60) + 1j * rng.standard_normal(60)])
    coeff = coeff.astype(np.complex64).astype(np.complex128)
    psi = random_state(rng, L, np.complex64)
    prior = random_state(rng, L, np.complex64)
    ref = orc.apply(L, bits, coeff, psi.astype(np.complex128), fp=64)
    b = np.ascontiguousarray(bits, np.uint64)
    co = np.empty(2 * len(b))
    co[0::2], co[1::2] = np.real(coeff), np.imag(coeff)
    h = ctypes.c_void_p()
    hip = gpu.hip()
    gpu.check(hip.rp_terms_upload_ex(ctypes.byref(h), gpu.RP_C64, L,
                                     b.ctypes.data_as(ctypes.c_void_p),
                                     co.ctypes.data_as(ctypes.c_void_p), len(b),
                                     gpu.RP_TERMS_ACC_GROUP))
    gpu.check(hip.rp_terms_set_tiling(h, *tiling))
    vin = DeviceBuffer.from_array(psi)
    vout = DeviceBuffer(psi.nbytes)
    gpu.check(hip.rp_apply_ops(h, vout.ptr, vin.ptr, 0))
    got = vout.to_array(np.complex64, len(psi)).astype(np.complex128)
    scale = np.abs(ref).max()
    err = np.abs(got - ref).max() / scale
    vacc = DeviceBuffer.from_array(prior)
    gpu.check(hip.rp_apply_ops(h, vacc.ptr, vin.ptr, 1))
    got_acc = vacc.to_array(np.complex64, len(psi)).astype(np.complex128)
    err_acc = np.abs(got_acc - (ref + prior.astype(np.complex128))).max() / scale
    print("every sub class, tiling %s: %d terms, rel err %.2e, accumulating %.2e"
          % (tiling, len(b), err, err_acc), flush=True)
    for buf in (vin, vout, vacc):
        buf.free()
    gpu.check(hip.rp_terms_free(h))
    # FP32 sums of up to a few dozen terms per group, FP64 across groups
    assert err < 2e-6
    assert err_acc < 2e-6

