"""The parity-block form on the GPU: BlockState (inc/state.h:447-584) under
block operators (inc/blkops.h:700-728), its per-sector Krylov evolution and
Lanczos ground state, and the reference's parity-block SYK driver (E0 shift,
imaginary-time pre-evolution, OTOC and Green's function; examples/syk.cc:
120-284) against dense matrices built by the oracle."""
import numpy as np
import pytest

from conftest import random_pauli_sum, random_state

pytestmark = pytest.mark.gpu


def block_hams(gpu, orc, N, seed, fp, sp=0.0):
    L = N // 2 - 1
    out = []
    for which in (0, 1):
        b, c = orc.syk_block_terms(N, which, sp, seed=seed, fp=fp)
        out.append((gpu.Hamiltonian(L, b, c, fp=fp), orc.dense_matrix(L, b, c)))
    return out


def majorana_blocks(gpu, orc, N, n, fp):
    L = N // 2 - 1
    out = []
    for which in (0, 1):
        b, c = orc.majorana_block(N, n, which, fp=fp)
        bb, cc = np.array([b], np.uint64), np.array([c])
        out.append((gpu.Hamiltonian(L, bb, cc, fp=fp), orc.dense_matrix(L, bb, cc)))
    return out


def ed(H, v, t, beta=0.0):
    w, Q = np.linalg.eigh(H)
    return Q @ (np.exp((-beta - 1j * t) * w) * (Q.conj().T @ v))


@pytest.mark.parametrize("fp", [32, 64])
def test_block_state_layout_and_norm(gpu, fp):
    L = 9
    rng = np.random.default_rng(fp)
    dt = np.complex64 if fp == 32 else np.complex128
    a, b = random_state(rng, L, dt), random_state(rng, L, dt)
    s = gpu.BlockState(L, fp=fp)
    s.load(a, b)
    # full index = sector * 2^L + local, L sector first (inc/state.h:513-534)
    for i in (0, 5, (1 << L) - 1):
        assert s[i] == complex(a[i])
        assert s[(1 << L) + i] == complex(b[i])
    assert s.norm() == pytest.approx(np.sqrt(2.0), rel=1e-6 if fp == 32 else 1e-13)
    s.load(None, b)
    assert s[3] == 0 and s[(1 << L) + 3] == complex(b[3])
    assert s.store()[0] is None
    s.random(7)
    assert s.norm() == pytest.approx(1.0, rel=1e-6 if fp == 32 else 1e-13)


@pytest.mark.parametrize("fp", [32, 64])
def test_block_state_apply_vs_dense(gpu, orc, fp):
    N, L = 14, 6
    tol = 2e-5 if fp == 32 else 1e-12
    dt = np.complex64 if fp == 32 else np.complex128
    rng = np.random.default_rng(11)
    a, b = random_state(rng, L, dt), random_state(rng, L, dt)
    (hLL, mLL), (hRR, mRR) = block_hams(gpu, orc, N, 3, fp)
    s = gpu.BlockState(L, fp=fp)

    # BlockDiag: each sector under its own block
    s.load(a, b)
    s.apply(LL=hLL, RR=hRR)
    L2, R2 = s.store()
    assert np.abs(L2 - mLL @ a).max() < tol and np.abs(R2 - mRR @ b).max() < tol

    # BlockAntiDiag (a Majorana): L' = LR R, R' = RL L
    for n in (3, N - 2, N - 1):
        (hLR, mLR), (hRL, mRL) = majorana_blocks(gpu, orc, N, n, fp)
        s.load(a, b)
        s.apply(LR=hLR, RL=hRL)
        L2, R2 = s.store()
        assert np.abs(L2 - mLR @ b).max() < tol and np.abs(R2 - mRL @ a).max() < tol
        # a null sector stays out of the product: L only -> R only
        s.load(a, None)
        s.apply(LR=hLR, RL=hRL)
        L2, R2 = s.store()
        assert L2 is None and np.abs(R2 - mRL @ a).max() < tol

    # BlockOp with four random blocks
    blocks = []
    for k in range(4):
        bb, cc = random_pauli_sum(rng, L, 9)
        if fp == 32:
            cc = cc.astype(np.complex64).astype(np.complex128)
        blocks.append((gpu.Hamiltonian(L, bb, cc, fp=fp), orc.dense_matrix(L, bb, cc)))
    s.load(a, b)
    s.apply(*[h for h, _ in blocks])
    L2, R2 = s.store()
    full = np.block([[blocks[0][1], blocks[1][1]], [blocks[2][1], blocks[3][1]]])
    want = full @ np.concatenate([a, b])
    scale = np.abs(want).max()
    assert np.abs(np.concatenate([L2, R2]) - want).max() < tol * max(1.0, scale) * 4


@pytest.mark.parametrize("fp", [32, 64])
def test_block_state_evolve_vs_ed(gpu, orc, fp):
    N, L = 16, 7
    dt = np.complex64 if fp == 32 else np.complex128
    rng = np.random.default_rng(5)
    a, b = random_state(rng, L, dt), random_state(rng, L, dt)
    (hLL, mLL), (hRR, mRR) = block_hams(gpu, orc, N, 9, fp)
    s = gpu.BlockState(L, fp=fp)
    s.load(a, b)
    t, beta = 2.0, 0.3
    st = s.evolve(hLL, hRR, t, beta, krydim=12, tol=1e-7 if fp == 32 else 1e-13)
    L2, R2 = s.store()
    eL = np.abs(L2 - ed(mLL, a.astype(np.complex128), t, beta)).max()
    eR = np.abs(R2 - ed(mRR, b.astype(np.complex128), t, beta)).max()
    print("block evolve fp%d: err L %.2e R %.2e, matvecs %d" % (fp, eL, eR, st.matvecs),
          flush=True)
    assert max(eL, eR) < (2e-5 if fp == 32 else 1e-11)


def test_block_ground_state(gpu, orc):
    N, L = 16, 7
    (hLL, mLL), (hRR, mRR) = block_hams(gpu, orc, N, 21, 64)
    wL, wR = np.linalg.eigvalsh(mLL)[0], np.linalg.eigvalsh(mRR)[0]
    s = gpu.BlockState(L, fp=64)
    e, sector = s.ground_state(hLL, hRR, krydim=20, eps=1e-9)
    print("ground state: %.12f (L %.12f, R %.12f), sector %d" % (e, wL, wR, sector), flush=True)
    assert e == pytest.approx(min(wL, wR), abs=1e-8)
    assert sector == (0 if wL <= wR else 1)
    Lv, Rv = s.store()
    assert (Lv is None) == (sector == 1) and (Rv is None) == (sector == 0)
    v = Lv if sector == 0 else Rv
    m = mLL if sector == 0 else mRR
    assert np.linalg.norm(m @ v - e * v) < 1e-4


def test_block_and_non_block_ground_energies_agree(gpu, orc):
    """min over the sectors equals the non-block Hamiltonian's ground state."""
    N, seed = 20, 4
    L = N // 2
    b, c = orc.syk_terms(N, seed=seed, fp=64)
    full = gpu.DeviceState(L, fp=64)
    full.random(3)
    e_full = full.ground_state(gpu.Hamiltonian(L, b, c, fp=64), krydim=24, eps=1e-9)
    (hLL, _), (hRR, _) = block_hams(gpu, orc, N, seed, 64)
    s = gpu.BlockState(L - 1, fp=64)
    e_blk, _ = s.ground_state(hLL, hRR, krydim=24, eps=1e-9)
    print("N=%d ground energy: non-block %.12f, block %.12f" % (N, e_full, e_blk), flush=True)
    assert e_blk == pytest.approx(e_full, abs=1e-8)


def ed_npoint(kind, mLL, mRR, mop, s0, tmax, nsteps, beta):
    """examples/syk.cc:198-284 with dense matrices and no E0 shift (the
    shift multiplies every v_k by the same factor, which v_k / v_0 removes)."""
    init = ed(mLL, s0, 0.0, beta / 8 if kind == "otoc" else beta / 2)
    out = []
    for k in range(nsteps + 1):
        t = k * tmax / nsteps
        s = init
        if kind == "otoc":
            s = mop @ s
            s = ed(mRR, s, t, beta / 4)
            s = ed(mLL, s, -t, beta / 4)
            s = mop @ s
            s = ed(mRR, s, t, beta / 4)
            s = ed(mLL, s, -t, 0.0)
        else:
            s = ed(mRR, s, t)
            s = ed(mLL, s, -t)
        out.append(np.vdot(init, s))
    out = np.array(out)
    return out / out[0]


@pytest.mark.parametrize("kind,beta,fp", [("otoc", 0.0, 64), ("otoc", 4.0, 64), ("green", 2.0, 64),
                                          ("otoc", 4.0, 32), ("green", 0.0, 32)])
def test_syk_block_npoint_vs_ed(gpu, orc, kind, beta, fp):
    N, seed, L = 14, 12, 6
    s0 = random_state(np.random.default_rng(2), L)
    tmax, nsteps = 5.0, 10
    t, v, e0, st = gpu.syk_block_npoint(N, kind, seed=seed, init=s0, beta=beta, tmax=tmax,
                                        nsteps=nsteps, krylov_dim=12,
                                        tol=1e-7 if fp == 32 else 1e-13, fp=fp)
    bL, cL = orc.syk_block_terms(N, 0, seed=seed, fp=fp)
    bR, cR = orc.syk_block_terms(N, 1, seed=seed, fp=fp)
    mLL, mRR = orc.dense_matrix(L, bL, cL), orc.dense_matrix(L, bR, cR)
    ob, oc = orc.majorana_block(N, N - 2, 0, fp=fp)
    mop = orc.dense_matrix(L, np.array([ob], np.uint64), np.array([oc]))
    ref = ed_npoint(kind, mLL, mRR, mop, s0, tmax, nsteps, beta)
    err = np.abs(v - ref).max()
    wmin = np.linalg.eigvalsh(mLL)[0]
    print("%s beta=%g fp%d: max |v - v_ed| = %.2e, E0 %.6f (exact %.6f), matvecs %d" % (
        kind, beta, fp, err, e0, wmin, st.matvecs), flush=True)
    assert np.array_equal(t, np.arange(nsteps + 1) * (tmax / nsteps))
    assert v[0] == 1
    assert err < (2e-5 if fp == 32 else 1e-10)
    # Lanczos with eps 0.1 (examples/syk.cc:137-140): a variational upper bound
    assert wmin - 1e-6 <= e0 <= wmin + 0.1


def test_syk_block_npoint_large(gpu):
    """N = 40 (two sectors of 19 sites): |v_k| <= 1 at beta = 0 for both
    functions (unitary evolutions and op^2 = 1/2), and two random initial
    states give the same curve to within the typical fluctuation of a state
    of 2^19 amplitudes (a few 2^-9.5).  FP32 and FP64 draw different
    couplings from the same seed (the generator's normal draws depend on the
    precision), so they are separate disorder realizations."""
    N = 40
    for fp in (32, 64):
        for kind in ("otoc", "green"):
            vs = []
            for state_seed in ((9, 10) if fp == 32 else (9,)):
                _, v, e0, st = gpu.syk_block_npoint(N, kind, sp=8.0, seed=5,
                                                    state_seed=state_seed, beta=0.0, tmax=4.0,
                                                    nsteps=4, krylov_dim=20,
                                                    tol=1e-7 if fp == 32 else 1e-12, fp=fp)
                print("N=40 %s fp%d state %d: |v| max %.6f, v(tmax) %s, E0 %.6f, matvecs %d" % (
                    kind, fp, state_seed, np.abs(v).max(), v[-1], e0, st.matvecs), flush=True)
                assert v[0] == 1
                assert np.all(np.abs(v) <= 1 + 1e-5)
                vs.append(v)
            if len(vs) == 2:
                assert np.abs(vs[0] - vs[1]).max() < 0.02


def test_syk_block_npoint_with_proposer_vs_ed(gpu, orc):
    from reapers_amd import proposer
    N, seed, L, beta = 14, 12, 6, 2.0
    s0 = random_state(np.random.default_rng(2), L)
    sched = gpu.Schedule.from_mlp(proposer.load_weights())
    t, v, e0, st = gpu.syk_block_npoint(N, "otoc", seed=seed, init=s0, beta=beta, tmax=5.0,
                                        nsteps=5, krylov_dim=12, tol=1e-13, fp=64,
                                        schedule=sched)
    mLL = orc.dense_matrix(L, *orc.syk_block_terms(N, 0, seed=seed, fp=64))
    mRR = orc.dense_matrix(L, *orc.syk_block_terms(N, 1, seed=seed, fp=64))
    ob, oc = orc.majorana_block(N, N - 2, 0, fp=64)
    mop = orc.dense_matrix(L, np.array([ob], np.uint64), np.array([oc]))
    ref = ed_npoint("otoc", mLL, mRR, mop, s0, 5.0, 5, beta)
    print("block OTOC with the proposer: %.2e, matvecs %d" % (np.abs(v - ref).max(), st.matvecs),
          flush=True)
    assert np.abs(v - ref).max() < 1e-10


def test_syk_block_npoint_promoted_fp64_cross_check(gpu):
    """promote=True runs FP64 on the FP32 run's couplings and initial state
    (the reference's cross-check, examples/syk.cc:440-474): the FP32 OTOC
    agrees with it to FP32 accuracy, while a native FP64 run (its own
    couplings) does not."""
    N, seed, beta = 24, 31, 1.0
    kw = dict(seed=seed, state_seed=77, beta=beta, tmax=4.0, nsteps=4, krylov_dim=16)
    _, v32, _, _ = gpu.syk_block_npoint(N, "otoc", tol=1e-7, fp=32, **kw)
    _, v64p, _, _ = gpu.syk_block_npoint(N, "otoc", tol=1e-12, fp=64, promote=True, **kw)
    _, v64n, _, _ = gpu.syk_block_npoint(N, "otoc", tol=1e-12, fp=64, **kw)
    err = np.abs(v32 - v64p).max()
    print("FP32 vs promoted FP64: %.2e; vs native FP64: %.2e" % (err, np.abs(v32 - v64n).max()),
          flush=True)
    assert err < 1e-5
    assert np.abs(v32 - v64n).max() > 1e-3
