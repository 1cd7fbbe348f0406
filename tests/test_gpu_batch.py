"""Batched right-hand sides: problems of uneven size with 1, 2 or 4 states
each in the block-state layout.  The batched apply against the oracle's
per-state apply (every instantiated tiling), and the batched restarted
Krylov evolution (HIP-graph replayed restarts) against the oracle's
single-state evolution, the same evolution without the graph, and FP64."""
import numpy as np
import pytest

from conftest import random_pauli_sum, random_state

pytestmark = pytest.mark.gpu


def problems(orc, rng, lens, fp):
    out = []
    for i, L in enumerate(lens):
        if i % 2 == 0:
            bits, coeff = orc.syk_terms(2 * L, 8.0, seed=500 + L + i, fp=fp)
        else:
            bits, coeff = random_pauli_sum(rng, L, 40)
            # a Hermitian-free general sum exercises complex and multi subs
            bits = np.concatenate([bits, bits[:5]])
            coeff = np.concatenate([coeff, rng.standard_normal(5) * 1j])
            if fp == 32:
                coeff = coeff.astype(np.complex64).astype(np.complex128)
        out.append((bits, coeff))
    return out


def tile_bits(fp, nrhs, large):
    esz = (8 if fp == 32 else 16) * nrhs
    return int(np.log2(32768 // esz)) + large


@pytest.mark.parametrize("fp", [32, 64])
@pytest.mark.parametrize("nrhs", [1, 2, 4])
@pytest.mark.parametrize("large", [0, 1, 2])
def test_batch_apply_vs_oracle(gpu, orc, fp, nrhs, large):
    """large 0: 32 KiB tiles, 256 threads; 2: the same tiles with 512
    threads (RP_BATCH_WIDE); 1: 64 KiB tiles, 512 threads."""
    tb = tile_bits(fp, nrhs, 1 if large == 1 else 0)
    rng = np.random.default_rng(fp * 100 + nrhs * 10 + tb + large)
    lens = [14, 16, 15, 14] if tb > 12 else [12, 14, 13, 12]
    terms = problems(orc, rng, lens, fp)
    b = gpu.Batch(lens, terms, nrhs=nrhs, kmax=4, fp=fp, tile_bits=tb, wide=(large == 2))
    assert b.info.tile_bits == tb
    assert b.info.threads == (256 if large == 0 else 512)
    dt = np.complex64 if fp == 32 else np.complex128
    states = [np.stack([random_state(rng, L, dt) for _ in range(nrhs)], 1) for L in lens]
    b.load(states)
    b.apply()
    got = b.store()
    worst = 0.0
    for p, L in enumerate(lens):
        bits, coeff = terms[p]
        for s in range(nrhs):
            ref = orc.apply(L, bits, coeff, states[p][:, s].astype(np.complex128), fp=64)
            e = np.abs(got[p][:, s] - ref).max() / np.abs(ref).max()
            worst = max(worst, e)
    print("fp%d nrhs=%d tb=%d: max rel err %.2e" % (fp, nrhs, tb, worst), flush=True)
    assert worst < (2e-6 if fp == 32 else 1e-13)


def test_batch_autotune_picks_an_instantiated_tiling(gpu, orc):
    rng = np.random.default_rng(5)
    lens = [14, 16, 15]
    b = gpu.Batch(lens, problems(orc, rng, lens, 32), nrhs=4, kmax=4, fp=32, autotune=True)
    assert b.info.tile_bits in (tile_bits(32, 4, 0), tile_bits(32, 4, 1))
    assert b.info.grid >= 1


@pytest.mark.parametrize("fp", [64, 32])
def test_batch_evolve_vs_oracle(gpu, orc, fp):
    rng = np.random.default_rng(11 + fp)
    lens = [12, 14, 13, 12, 14]
    terms = []
    for i, L in enumerate(lens):
        terms.append(orc.syk_terms(2 * L, 8.0, seed=900 + i, fp=fp))
    nrhs, t, m = 4, 2.0, 10
    dt = np.complex64 if fp == 32 else np.complex128
    states = [np.stack([random_state(rng, L, dt) for _ in range(nrhs)], 1) for L in lens]
    b = gpu.Batch(lens, terms, nrhs=nrhs, kmax=m, fp=fp)
    b.load(states)
    tol = 1e-13 if fp == 64 else 1e-7
    st, rs = b.evolve(t, m=m, tol=tol)
    got = b.store()
    worst = 0.0
    for p, L in enumerate(lens):
        bits, coeff = terms[p]
        for s in range(nrhs):
            ref, _ = orc.krylov_evolve(L, bits, coeff, states[p][:, s].astype(np.complex128), t,
                                       0.0, krydim=20, tol=1e-14)
            worst = max(worst, np.linalg.norm(got[p][:, s] - ref))
    print("fp%d batch evolve: max |err| %.2e, restarts %s, matvecs %d, %.1f ms" % (
        fp, worst, rs.max(1).tolist(), st.matvecs, st.wall_ms), flush=True)
    assert worst < (1e-11 if fp == 64 else 1e-5)


def test_batch_graph_matches_eager(gpu, orc):
    rng = np.random.default_rng(3)
    lens = [13, 12, 14]
    terms = [orc.syk_terms(2 * L, 8.0, seed=40 + L, fp=32) for L in lens]
    states = [np.stack([random_state(rng, L, np.complex64) for _ in range(2)], 1) for L in lens]
    out = []
    for graph in (True, False, True):
        b = gpu.Batch(lens, terms, nrhs=2, kmax=8, fp=32, tile_bits=tile_bits(32, 2, 0))
        b.load(states)
        b.evolve(3.0, m=8, tol=1e-7, graph=graph)
        out.append(b.store())
        b.close()
    for p in range(len(lens)):
        # graph replays and eager launches run the same kernels in the same order
        assert np.array_equal(out[0][p], out[1][p])
        assert np.array_equal(out[0][p], out[2][p])


def test_batch_uneven_convergence(gpu, orc):
    # a problem with H = 0 converges after one restart; the others keep going
    rng = np.random.default_rng(8)
    lens = [12, 12]
    terms = [(np.array([0], np.uint64), np.array([0.5 + 0j])),
             orc.syk_terms(24, 8.0, seed=77, fp=64)]
    states = [np.stack([random_state(rng, 12) for _ in range(4)], 1) for _ in lens]
    b = gpu.Batch(lens, terms, nrhs=4, kmax=6, fp=64)
    b.load(states)
    st, rs = b.evolve(4.0, m=6, tol=1e-12)
    got = b.store()
    # H = 0.5 Id: exp(-0.5 i t) psi, exact after one restart
    assert np.abs(got[0] - np.exp(-2.0j) * states[0]).max() < 1e-13
    assert rs[0].max() == 1 and rs[1].min() > 1


def test_batch_with_proposer_schedule(gpu, orc):
    """The batch with the trained proposer choosing each restart's dimension
    (the largest the active states propose) reaches the oracle's result."""
    from reapers_amd import proposer
    rng = np.random.default_rng(21)
    lens = [12, 13, 14]
    terms = [orc.syk_terms(2 * L, 8.0, seed=60 + L, fp=64) for L in lens]
    states = [np.stack([random_state(rng, L) for _ in range(4)], 1) for L in lens]
    sched = gpu.Schedule.from_mlp(proposer.load_weights(), max_dim=16)
    b = gpu.Batch(lens, terms, nrhs=4, kmax=16, fp=64)
    b.load(states)
    t, tol = 4.0, 1e-10
    st, rs = b.evolve(t, m=12, tol=tol, schedule=sched)
    got = b.store()
    worst = 0.0
    for p, L in enumerate(lens):
        for s in range(4):
            ref, _ = orc.krylov_evolve(L, *terms[p], states[p][:, s], t, 0.0, krydim=20,
                                       tol=1e-14)
            worst = max(worst, np.linalg.norm(got[p][:, s] - ref))
    print("batch + proposer: err %.2e, matvecs %d, restarts %s, rejected %d" % (
        worst, st.matvecs, rs.max(1).tolist(), st.rejected), flush=True)
    assert worst < 10 * tol


@pytest.mark.parametrize("nrhs", [1, 2, 4])
@pytest.mark.parametrize("large", [0, 1])
def test_batch_group_accumulation_vs_oracle(gpu, orc, nrhs, large):
    """RP_BATCH_ACC_GROUP (complex<float>): FP32 sums per pair of tile groups,
    FP64 across, against the oracle's FP64 apply; then an evolution."""
    tb = tile_bits(32, nrhs, large)
    rng = np.random.default_rng(700 + nrhs * 10 + tb)
    lens = [14, 16, 15, 14] if tb > 12 else [12, 14, 13, 12]
    terms = [orc.syk_terms(2 * L, 8.0, seed=800 + i, fp=32) for i, L in enumerate(lens)]
    b = gpu.Batch(lens, terms, nrhs=nrhs, kmax=10, fp=32, tile_bits=tb, acc_group=True)
    states = [np.stack([random_state(rng, L, np.complex64) for _ in range(nrhs)], 1) for L in lens]
    b.load(states)
    b.apply()
    got = b.store()
    worst = 0.0
    for p, L in enumerate(lens):
        bits, coeff = terms[p]
        for s in range(nrhs):
            ref = orc.apply(L, bits, coeff, states[p][:, s].astype(np.complex128), fp=64)
            worst = max(worst, np.abs(got[p][:, s] - ref).max() / np.abs(ref).max())
    print("acc group nrhs=%d tb=%d: apply max rel err %.2e" % (nrhs, tb, worst), flush=True)
    # FP32 partial sums over a pair of groups: a few FP32 ulps of the largest
    # amplitude, the bound of the FP64-accumulating batched apply above
    assert worst < 2e-6
    b.load(states)
    st, rs = b.evolve(2.0, m=10, tol=1e-7)
    got = b.store()
    worst = 0.0
    for p, L in enumerate(lens):
        bits, coeff = terms[p]
        for s in range(nrhs):
            ref, _ = orc.krylov_evolve(L, bits, coeff, states[p][:, s].astype(np.complex128), 2.0,
                                       0.0, krydim=20, tol=1e-14)
            worst = max(worst, np.linalg.norm(got[p][:, s] - ref))
    print("acc group nrhs=%d tb=%d: evolve max |err| %.2e" % (nrhs, tb, worst), flush=True)
    assert worst < 1e-5


def test_batch_selective_second_pass(gpu, orc):
    """The batched Arnoldi step's second Gram-Schmidt pass runs only for
    problems with a state whose first-pass norm is below 1/4 (the reference's
    selective criterion, inc/algo.h:79-101, decided on the device): a
    problem scaled to a small norm, whose every step takes the second pass,
    beside one of ordinary norm, whose steps skip it; both against the
    oracle's FP64 evolution."""
    rng = np.random.default_rng(44)
    lens = [12, 13]
    b1, c1 = orc.syk_terms(24, 8.0, seed=31, fp=64)
    b2, c2 = orc.syk_terms(26, 8.0, seed=32, fp=64)
    terms = [(b1, c1 * 1e-3), (b2, c2)]
    t = 3.0
    states = [np.stack([random_state(rng, L) for _ in range(2)], 1) for L in lens]
    for graph in (True, False):
        b = gpu.Batch(lens, terms, nrhs=2, kmax=10, fp=64)
        b.load(states)
        st, rs = b.evolve(t, m=10, tol=1e-12, graph=graph)
        got = b.store()
        b.close()
        worst = 0.0
        for p, L in enumerate(lens):
            for s in range(2):
                ref, _ = orc.krylov_evolve(L, *terms[p], states[p][:, s], t, 0.0, krydim=20,
                                           tol=1e-14)
                worst = max(worst, np.linalg.norm(got[p][:, s] - ref))
        print("selective second pass (graph=%d): max |err| %.2e, restarts %s" % (
            graph, worst, rs.max(1).tolist()), flush=True)
        assert worst < 1e-11
