"""GPU parity of the restarted Krylov evolution (State::evolve on GPUImpl,
inc/algo.h:266-365) against the oracle (the reference algorithm in FP64) and
against exact diagonalization (numpy eigh).

Tolerances: FP64 states within 1e-11 (max-abs, unit-norm states); FP32 bulk
states within 1e-5 of the FP64 oracle."""
import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu


def ed_evolve(H, psi, t, beta=0.0):
    w, Q = np.linalg.eigh(H)
    return Q @ (np.exp((-beta - 1j * t) * w) * (Q.conj().T @ psi))


@pytest.mark.parametrize("t,beta", [(0.5, 0.0), (2.0, 0.0), (5.0, 0.0), (1.0, 0.3)])
def test_fp64_vs_oracle_and_ed(gpu, orc, t, beta):
    N = 16
    L = N // 2
    bits, coeff = orc.syk_terms(N, 0.0, seed=21, fp=64)
    psi = random_state(np.random.default_rng(2), L)
    ref, st_ref = orc.krylov_evolve(L, bits, coeff, psi, t, beta, krydim=10, tol=1e-12)
    H = orc.dense_matrix(L, bits, coeff)
    ex = ed_evolve(H, psi, t, beta)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=64)
    s = gpu.DeviceState(L, fp=64)
    s.load(psi)
    st = s.evolve(ham, t, beta, krydim=10, tol=1e-12)
    got = s.store()
    print("t=%g beta=%g gpu-oracle %.2e gpu-ed %.2e oracle-ed %.2e matvecs %d/%d" % (
        t, beta, np.abs(got - ref).max(), np.abs(got - ex).max(), np.abs(ref - ex).max(),
        st.matvecs, st_ref.matvecs), flush=True)
    assert np.abs(got - ref).max() < 1e-11
    assert np.abs(got - ex).max() < 1e-11


def test_fp32_bulk_vs_fp64(gpu, orc):
    N = 28
    L = N // 2
    bits, coeff = orc.syk_terms(N, 8.0, seed=5, fp=32)
    psi = random_state(np.random.default_rng(4), L, np.complex64)
    ref, _ = orc.krylov_evolve(L, bits, coeff, psi.astype(np.complex128), 3.0, 0.0,
                               krydim=20, tol=1e-12)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=32)
    s = gpu.DeviceState(L, fp=32)
    s.load(psi)
    st = s.evolve(ham, 3.0, 0.0, krydim=20, tol=1e-7)
    got = s.store().astype(np.complex128)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("fp32 rel err vs fp64 oracle %.3e matvecs %d" % (err, st.matvecs), flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("dims,nsub", [([4], 1), ([3, 5, 8], 2), ([8, 4], 3), ([2], 1)])
def test_schedules_fp64(gpu, orc, dims, nsub):
    """Per-restart Krylov dimensions and sub-steps converge to the same
    exp(-iHt)psi (against ED)."""
    N, L, t = 16, 8, 4.0
    bits, coeff = orc.syk_terms(N, 0.0, seed=31, fp=64)
    psi = random_state(np.random.default_rng(8), L)
    ex = ed_evolve(orc.dense_matrix(L, bits, coeff), psi, t)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=64)
    s = gpu.DeviceState(L, fp=64)
    s.load(psi)
    st = s.evolve_schedule(ham, t, dims, nsub=nsub, tol=1e-13)
    got = s.store()
    print("dims=%s nsub=%d matvecs=%d restarts=%d err=%.2e" % (
        dims, nsub, st.matvecs, st.restarts, np.abs(got - ex).max()), flush=True)
    assert np.abs(got - ex).max() < 1e-11
    assert st.substeps == nsub


def test_schedule_fp32_syk(gpu, orc):
    N, L, t = 32, 16, 10.0
    bits, coeff = orc.syk_terms(N, 8.0, seed=2000, fp=32)
    psi = random_state(np.random.default_rng(9), L, np.complex64)
    ref, _ = orc.krylov_evolve(L, bits, coeff, psi.astype(np.complex128), t, 0.0, krydim=30,
                               tol=1e-14)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=32)
    s = gpu.DeviceState(L, fp=32)
    s.load(psi)
    st = s.evolve_schedule(ham, t, [6], tol=1e-7)
    err = np.linalg.norm(s.store().astype(np.complex128) - ref)
    print("fp32 m=6: matvecs %d err %.2e" % (st.matvecs, err), flush=True)
    assert err < 1e-6
    assert st.matvecs <= 48


def test_xxz_evolution(gpu, orc):
    L, t = 10, 3.0
    bits, coeff = gpu.xxz_terms(L, 1.0, 0.7, 3.0, seed=5, fp=64)
    psi = random_state(np.random.default_rng(10), L)
    ex = ed_evolve(orc.dense_matrix(L, bits, coeff), psi, t, 0.2)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=64)
    s = gpu.DeviceState(L, fp=64)
    s.load(psi)
    s.evolve(ham, t, 0.2, krydim=12, tol=1e-13)
    assert np.abs(s.store() - ex).max() < 1e-11


def test_lanczos_ground_state_beyond_64_vectors(gpu, orc):
    """Krylov::ground_state at krydim 72: each restart's answer combines 72
    basis vectors in one rp_lincomb call; the energy against numpy's lowest
    eigenvalue of the oracle's dense matrix."""
    L = 10
    bits, coeff = orc.syk_terms(20, 0.0, seed=11, fp=64)
    H = gpu.Hamiltonian(L, bits, coeff, fp=64)
    s = gpu.DeviceState(L, fp=64)
    s.random(5)
    e = s.ground_state(H, krydim=72, eps=1e-9)
    dense = orc.dense_matrix(L, bits, coeff)
    e0 = np.linalg.eigvalsh(dense)[0]
    print("krydim 72: E0 %.12f, eigvalsh %.12f" % (e, e0), flush=True)
    s.close()
    H.close()
    assert abs(e - e0) <= 1e-9 * max(1.0, abs(e0))
