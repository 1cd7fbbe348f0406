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
