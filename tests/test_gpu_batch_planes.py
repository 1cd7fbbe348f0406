"""The batched apply through state planes (RP_BATCH_PLANES): one problem of
20 or more qubits whose nrhs states are gathered into planes of their own,
the single-state apply run on each plane, the results scattered back into
the block layout.  Checked through the C ABI against the oracle's per-state
apply (both precisions, FP64 and group accumulation, with and without
accumulate), with an inactive problem left untouched; and the batched
evolution (graph-replayed restarts) against FP64 single-state evolutions
and against the same batch without the graph.

Tolerances: the apply within 2e-6 of the largest amplitude in FP32 (FP32
inputs and outputs; group accumulation rounds partial sums of four tile
groups to FP32) and 1e-13 in FP64; the FP32 evolution to t=2 within 1e-5
of the FP64 one (tol 1e-7 per restart, the batched path's own bound)."""
import ctypes as C

import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu

L = 20                  # the smallest problem the plane apply takes
RP_BATCH_ACC_GROUP, RP_BATCH_PLANES = 2, 8


def c_batch(gpu, fp, nrhs, bits, coeff, flags):
    hip = gpu.hip()
    vp = C.c_void_p
    b = vp()
    co = np.ascontiguousarray(np.stack([coeff.real, coeff.imag], 1).ravel())
    bits = np.ascontiguousarray(bits, np.uint64)
    lens = np.array([L], np.int32)
    toff = np.array([0, len(bits)], np.int64)
    prec = gpu.RP_C64 if fp == 32 else gpu.RP_C128
    gpu.check(hip.rp_batch_create(C.byref(b), prec, 1, lens.ctypes.data_as(vp),
                                  bits.ctypes.data_as(vp), co.ctypes.data_as(vp),
                                  toff.ctypes.data_as(vp), nrhs, 4, flags))
    info = gpu.BatchInfo()
    gpu.check(hip.rp_batch_info(b, C.byref(info)))
    return b, info


def to_dev(gpu, a):
    hip = gpu.hip()
    p = C.c_void_p()
    gpu.check(hip.rp_malloc(C.byref(p), C.c_size_t(a.nbytes)))
    gpu.check(hip.rp_memcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 0))
    return p


def from_dev(gpu, p, like):
    out = np.empty_like(like)
    gpu.check(gpu.hip().rp_memcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 1))
    return out


@pytest.mark.parametrize("fp,acc_group", [(32, False), (32, True), (64, False)])
@pytest.mark.parametrize("nrhs", [1, 2, 4])
def test_plane_apply_vs_oracle(gpu, orc, fp, acc_group, nrhs):
    rng = np.random.default_rng(100 * nrhs + fp + acc_group)
    bits, coeff = orc.syk_terms(2 * L, 8.0, seed=4100 + nrhs, fp=fp)
    dt = np.complex64 if fp == 32 else np.complex128
    flags = RP_BATCH_PLANES | (RP_BATCH_ACC_GROUP if acc_group else 0)
    b, info = c_batch(gpu, fp, nrhs, bits, coeff, flags)
    hip = gpu.hip()
    try:
        assert info.planes == 1 and info.tile_bits > 0
        psi = np.stack([random_state(rng, L, dt) for _ in range(nrhs)], 1)   # (2^L, nrhs)
        r0 = np.stack([random_state(rng, L, dt) for _ in range(nrhs)], 1)
        x, y = to_dev(gpu, np.ascontiguousarray(psi)), to_dev(gpu, np.ascontiguousarray(r0))
        gpu.check(hip.rp_batch_apply(b, y, x, 1))          # y = r0 + H psi
        got = from_dev(gpu, y, r0)
        worst = 0.0
        for s in range(nrhs):
            ref = orc.apply(L, bits, coeff, psi[:, s].astype(np.complex128), fp=64)
            want = r0[:, s].astype(np.complex128) + ref
            worst = max(worst, np.abs(got[:, s] - want).max() / np.abs(want).max())
        print("planes fp%d acc_group=%d nrhs=%d (tiling %d/%d): max rel err %.2e" % (
            fp, acc_group, nrhs, info.tile_bits, info.threads, worst), flush=True)
        assert worst < (2e-6 if fp == 32 else 1e-13)
        # an inactive problem: the apply writes nothing
        gpu.check(hip.rp_batch_set_active(b, np.array([0], np.int32).ctypes.data_as(C.c_void_p)))
        gpu.check(hip.rp_batch_apply(b, y, x, 0))
        assert np.array_equal(from_dev(gpu, y, r0), got)
        gpu.check(hip.rp_free(x))
        gpu.check(hip.rp_free(y))
    finally:
        gpu.check(hip.rp_batch_free(b))


@pytest.mark.parametrize("acc_group", [False, True])
def test_plane_batch_evolve(gpu, orc, acc_group):
    rng = np.random.default_rng(7 + acc_group)
    bits, coeff = orc.syk_terms(2 * L, 8.0, seed=4200, fp=32)
    nrhs, t, m = 4, 2.0, 12
    states = [np.stack([random_state(rng, L, np.complex64) for _ in range(nrhs)], 1)]
    out = []
    for graph in (True, False, True):
        b = gpu.Batch([L], [(bits, coeff)], nrhs=nrhs, kmax=m, fp=32, autotune=False,
                      acc_group=acc_group, planes=True)
        assert b.info.planes == 1
        b.load(states)
        st, rs = b.evolve(t, m=m, tol=1e-7, graph=graph)
        out.append(b.store()[0])
        b.close()
    # graph replays and eager launches run the same kernels in the same order
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])
    H64 = gpu.Hamiltonian(L, bits, coeff, fp=64)
    s64 = gpu.DeviceState(L, fp=64)
    worst = 0.0
    for q in range(nrhs):
        s64.load(states[0][:, q].astype(np.complex128))
        s64.evolve(H64, t, 0.0, krydim=20, tol=1e-13)
        ref = s64.store()
        worst = max(worst, np.linalg.norm(out[0][:, q] - ref))
    s64.close()
    H64.close()
    print("planes evolve acc_group=%d: max |err| %.2e, restarts %s, matvecs %d" % (
        acc_group, worst, rs.max(1).tolist(), st.matvecs), flush=True)
    assert worst < 1e-5


def test_plane_autotune_choice(gpu, orc):
    """With the autotuner a one-problem batch of 20 qubits times the plane
    apply against the batched kernel and keeps either; both report a
    tiling."""
    bits, coeff = orc.syk_terms(2 * L, 8.0, seed=4300, fp=32)
    b = gpu.Batch([L], [(bits, coeff)], nrhs=4, kmax=6, fp=32, autotune=True)
    assert b.info.planes in (0, 1)
    assert b.info.tile_bits > 0 and b.info.grid >= 1
    print("autotuned one-problem batch at %d qubits: planes=%d tiling %d/%d" % (
        L, b.info.planes, b.info.tile_bits, b.info.threads), flush=True)
    b.close()


def test_plane_apply_declines_an_empty_term_list(gpu):
    """H = 0 (its one term has a zero coefficient): the single-state apply
    would clear its output with a memset, which a captured restart must not
    hold, so the batch keeps its own kernel; the apply gives zero."""
    bits = np.array([0], np.uint64)
    coeff = np.array([0.0 + 0.0j])
    b = gpu.Batch([L], [(bits, coeff)], nrhs=2, kmax=4, fp=32, autotune=False, planes=True)
    assert b.info.planes == 0
    rng = np.random.default_rng(3)
    b.load([np.stack([random_state(rng, L, np.complex64) for _ in range(2)], 1)])
    b.apply()
    assert not np.any(b.store()[0])
    b.close()
