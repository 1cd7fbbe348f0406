"""Pinning of the oracle's small-matrix exponential and restarted Krylov
evolution (inc/algo.h:76-173, 266-365).

Known answers: exp(c sigma) for c = 1 and c = i from the reference's unit
tests (tests/ops.cc:198-229).  Library routines: scipy.linalg.expm, numpy
eigh (exact diagonalization, the reference's own oracle at small N:
examples/simple-syk-exdiag.cc) and scipy.sparse.linalg.expm_multiply.
Invariants: unit norm under real-time evolution, the group property, and
Arnoldi breakdown on an invariant subspace."""
import numpy as np
import pytest
from scipy.linalg import expm
from scipy.sparse.linalg import expm_multiply

PX = np.array([[0, 1], [1, 0]], complex)
PY = np.array([[0, -1j], [1j, 0]])
PZ = np.array([[1, 0], [0, -1]], complex)

TOL = 100 * np.finfo(float).eps      # tests/ops.cc:23


def test_matexp_known_answers(orc):
    """tests/ops.cc:198-229."""
    e, ch, sh, ie = 2.71828182845905, 1.54308063481524, 1.17520119364380, 0.367879441171442
    co, si = 0.540302305868140, 0.841470984807897
    cases = [
        (np.eye(2), 1, [[e, 0], [0, e]]),
        (PX, 1, [[ch, sh], [sh, ch]]),
        (PY, 1, [[ch, -1j * sh], [1j * sh, ch]]),
        (PZ, 1, [[e, 0], [0, ie]]),
        (PX, 1j, [[co, 1j * si], [1j * si, co]]),
        (PY, 1j, [[co, si], [-si, co]]),
        (PZ, 1j, [[co + 1j * si, 0], [0, co - 1j * si]]),
    ]
    for m, c, want in cases:
        got = orc.expm(c * m)
        # the reference's literals carry 15 significant digits
        assert np.abs(got - np.array(want)).max() < 1e-13


@pytest.mark.parametrize("scale", [1e-3, 0.1, 0.5, 1.5, 5.0, 40.0])
def test_expm_vs_scipy(orc, scale):
    """Norms chosen to exercise every Pade degree (3,5,7,9,13) and squaring."""
    rng = np.random.default_rng(int(scale * 1000))
    n = 12
    a = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    a *= scale / np.abs(a).sum(axis=0).max()
    ref = expm(a)
    got = orc.expm(a)
    assert np.abs(got - ref).max() <= 1e-13 * max(1, np.abs(ref).max())
    got32 = orc.expm(a.astype(np.complex64), fp=32)
    assert np.abs(got32 - ref).max() <= 5e-5 * max(1, np.abs(ref).max())


def _ed(H, psi, t, beta=0.0):
    w, Q = np.linalg.eigh(H)
    return Q @ (np.exp((-beta - 1j * t) * w) * (Q.conj().T @ psi))


def _state(L, seed):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal(1 << L) + 1j * rng.standard_normal(1 << L)
    return v / np.linalg.norm(v)


@pytest.mark.parametrize("N,sp", [(12, 0.0), (16, 0.0), (20, 4.0)])
@pytest.mark.parametrize("t,beta", [(0.5, 0.0), (3.0, 0.0), (1.0, 0.5), (0.0, 2.0)])
def test_krylov_vs_exact_diagonalization(orc, N, sp, t, beta):
    L = N // 2
    bits, coeff = orc.syk_terms(N, sp, seed=N)
    H = orc.dense_matrix(L, bits, coeff)
    psi = _state(L, 1)
    got, st = orc.krylov_evolve(L, bits, coeff, psi, t, beta, krydim=10, tol=1e-13)
    assert np.abs(got - _ed(H, psi, t, beta)).max() < 1e-11
    if beta == 0:
        assert abs(np.linalg.norm(got) - 1) < 1e-12


def test_config1_dense_syk24(orc):
    """Dense SYK N=24 (12 qubits, C(24,4) = 10626 terms), FP64, Krylov dim
    30, tol 1e-12, t in {0.5, 1, 2, 5}: against scipy's expm_multiply."""
    import scipy.sparse as sps
    N, L = 24, 12
    bits, coeff = orc.syk_terms(N, 0.0, seed=1000)
    assert len(bits) == 10626
    H = sps.csr_matrix(orc.dense_matrix(L, bits, coeff))
    psi = _state(L, 1001)
    for t in (0.5, 1.0, 2.0, 5.0):
        got, st = orc.krylov_evolve(L, bits, coeff, psi, t, 0.0, krydim=30, tol=1e-12)
        ref = expm_multiply(-1j * t * H, psi)
        assert np.abs(got - ref).max() < 1e-11, t


def test_group_property(orc):
    L = 8
    bits, coeff = orc.syk_terms(16, 0.0, seed=3)
    psi = _state(L, 2)
    a, _ = orc.krylov_evolve(L, bits, coeff, psi, 1.3, 0.0, krydim=12, tol=1e-14)
    b, _ = orc.krylov_evolve(L, bits, coeff, a, 0.9, 0.0, krydim=12, tol=1e-14)
    c, _ = orc.krylov_evolve(L, bits, coeff, psi, 2.2, 0.0, krydim=12, tol=1e-14)
    assert np.abs(b - c).max() < 1e-12


def test_breakdown_on_invariant_subspace(orc):
    """H = 0.5*Id on a basis vector: H v is exactly in the span, the
    Gram-Schmidt residual is exactly zero and the restart loop stops
    (inc/algo.h:104-106, 162-166), giving exp(-0.5 i t) v.  (With a generic
    vector the residual is a rounding error, not zero, and the reference's
    criterion does not report a breakdown; the result is the same.)"""
    L = 6
    psi = np.zeros(1 << L, complex)
    psi[3] = 1
    got, st = orc.krylov_evolve(L, np.array([0], np.uint64), np.array([0.5 + 0j]), psi, 2.0,
                                0.0, krydim=5, tol=1e-14)
    assert st.breakdown == 1 and st.matvecs == 1
    assert np.abs(got - np.exp(-0.5j * 2.0) * psi).max() < 1e-15
    psi = _state(L, 3)
    got, st = orc.krylov_evolve(L, np.array([0], np.uint64), np.array([0.7 + 0j]), psi, 2.0,
                                0.0, krydim=5, tol=1e-14)
    assert np.abs(got - np.exp(-0.7j * 2.0) * psi).max() < 1e-14
    # start from an eigenvector of a diagonal H: breakdown again
    bits = np.array([2 << 0, 2 << 4], np.uint64)          # sz(0) + 0.5 sz(2)
    coeff = np.array([1.0, 0.5], complex)
    e = np.zeros(1 << L, complex); e[5] = 1
    got, st = orc.krylov_evolve(L, bits, coeff, e, 1.0, 0.0, krydim=5, tol=1e-14)
    assert st.breakdown == 1
    energy = -1.0 + 0.5 * (-1)        # index 5 = 0b101: bit0 = 1, bit2 = 1
    assert abs(got[5] - np.exp(-1j * energy)) < 1e-14


def test_max_iters_caps_restarts(orc):
    L = 8
    bits, coeff = orc.syk_terms(16, 0.0, seed=9)
    psi = _state(L, 4)
    _, st = orc.krylov_evolve(L, bits, coeff, psi, 20.0, 0.0, krydim=4, tol=1e-14, max_iters=3)
    assert st.restarts == 3 and st.matvecs == 12


def test_lanczos_ground_state(orc):
    L = 8
    bits, coeff = orc.syk_terms(16, 0.0, seed=6)
    H = orc.dense_matrix(L, bits, coeff)
    e0 = np.linalg.eigvalsh(H)[0]
    e, v, it = orc.lanczos_ground_state(L, bits, coeff, _state(L, 5), krydim=20, eps=1e-12,
                                        maxiter=200)
    assert abs(e - e0) < 1e-10
    assert np.linalg.norm(H @ v - e0 * v) < 1e-6
