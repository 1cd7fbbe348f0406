"""OTOC F(t) = <i|A(t) B A(t) B|i> by the C++ API's otoc() (four Krylov
evolutions per time point, the reference's otoc_krylov sequence) against
exact diagonalization, the closed form at t = 0 and the FP64 run."""
import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu


def ed_otoc(H, A, B, psi, times, beta):
    w, Q = np.linalg.eigh(H)

    def ev(v, t, b):
        return Q @ (np.exp((-b - 1j * t) * w) * (Q.conj().T @ v))

    init = ev(psi, 0.0, beta / 8)
    out = []
    for t in times:
        v = B @ init
        v = ev(v, t, beta / 4)
        v = A @ v
        v = ev(v, -t, beta / 4)
        v = B @ v
        v = ev(v, t, beta / 4)
        v = A @ v
        v = ev(v, -t, 0.0)
        out.append(np.vdot(init, v))
    return np.array(out)


def majorana_ham(gpu, N, n, L, fp):
    b, c = gpu.majorana(N, n, fp=fp)
    return gpu.Hamiltonian(L, np.array([b], np.uint64), np.array([c]), fp=fp), (b, c)


@pytest.mark.parametrize("N,sp,beta", [(16, 0.0, 0.0), (20, 8.0, 0.0), (16, 0.0, 2.0)])
def test_otoc_fp64_vs_ed(gpu, orc, N, sp, beta):
    L = N // 2
    bits, coeff = orc.syk_terms(N, sp, seed=41 + N, fp=64)
    psi = random_state(np.random.default_rng(N), L)
    times = np.array([0.0, 0.1, 0.5, 1.0, 3.0, 8.0])
    ham = gpu.Hamiltonian(L, bits, coeff, fp=64)
    A, (ab, ac) = majorana_ham(gpu, N, 1, L, 64)
    B, (bb, bc) = majorana_ham(gpu, N, 2, L, 64)
    s = gpu.DeviceState(L, fp=64)
    s.load(psi)
    F, st = s.otoc(ham, A, B, times, beta=beta, krydim=12, tol=1e-14)
    H = orc.dense_matrix(L, bits, coeff)
    Am = orc.dense_matrix(L, np.array([ab], np.uint64), np.array([ac]))
    Bm = orc.dense_matrix(L, np.array([bb], np.uint64), np.array([bc]))
    ref = ed_otoc(H, Am, Bm, psi, times, beta)
    print("N=%d beta=%g: max |F - F_ed| = %.2e, matvecs %d" % (
        N, beta, np.abs(F - ref).max(), st.matvecs), flush=True)
    assert np.abs(F - ref).max() < 1e-11
    if beta == 0:
        # chi_1 chi_2 chi_1 chi_2 = -chi_1^2 chi_2^2 = -1/4 (std. normalization)
        assert F[0] == pytest.approx(-0.25, abs=1e-14)
        assert np.all(np.abs(F) <= 0.25 + 1e-12)


def test_otoc_fp32_vs_fp64(gpu, orc):
    N, L = 24, 12
    bits, coeff = orc.syk_terms(N, 8.0, seed=77, fp=32)
    psi = random_state(np.random.default_rng(3), L)
    times = np.geomspace(0.1, 20, 8)
    vals = {}
    for fp in (32, 64):
        ham = gpu.Hamiltonian(L, bits, coeff, fp=fp)
        A, _ = majorana_ham(gpu, N, 1, L, fp)
        B, _ = majorana_ham(gpu, N, 2, L, fp)
        s = gpu.DeviceState(L, fp=fp)
        s.load(psi.astype(np.complex64 if fp == 32 else np.complex128))
        vals[fp], _ = s.otoc(ham, A, B, times, krydim=12, tol=1e-7 if fp == 32 else 1e-14)
    err = np.abs(vals[32] - vals[64]).max()
    print("fp32 vs fp64 OTOC: max |F32 - F64| / |F(0)| = %.2e" % (err / 0.25), flush=True)
    # relative to |F(0)| = 1/4, the reference's normalization (examples/syk.cc:230-235)
    assert err / 0.25 < 1e-5


def test_otoc_with_proposer_schedule_vs_ed(gpu, orc):
    """The OTOC with the trained proposer choosing each restart's Krylov
    dimension agrees with ED like the fixed dimension does."""
    from reapers_amd import proposer
    N, L, beta = 16, 8, 1.0
    bits, coeff = orc.syk_terms(N, 0.0, seed=57, fp=64)
    psi = random_state(np.random.default_rng(8), L)
    times = np.array([0.0, 0.3, 2.0, 6.0])
    ham = gpu.Hamiltonian(L, bits, coeff, fp=64)
    A, (ab, ac) = majorana_ham(gpu, N, 1, L, 64)
    B, (bb, bc) = majorana_ham(gpu, N, 2, L, 64)
    sched = gpu.Schedule.from_mlp(proposer.load_weights())
    s = gpu.DeviceState(L, fp=64)
    s.load(psi)
    F, st = s.otoc(ham, A, B, times, beta=beta, krydim=12, tol=1e-13, schedule=sched)
    H = orc.dense_matrix(L, bits, coeff)
    Am = orc.dense_matrix(L, np.array([ab], np.uint64), np.array([ac]))
    Bm = orc.dense_matrix(L, np.array([bb], np.uint64), np.array([bc]))
    ref = ed_otoc(H, Am, Bm, psi, times, beta)
    print("OTOC with the proposer: max |F - F_ed| = %.2e, matvecs %d, rejected %d" % (
        np.abs(F - ref).max(), st.matvecs, st.rejected), flush=True)
    assert np.abs(F - ref).max() < 1e-11
