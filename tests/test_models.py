"""The random-field XXZ chain generator of the C++ API (a second model
family for the Krylov schedule's training traces) against a numpy Kronecker
construction of J sum (XX + YY + Delta ZZ) + sum h_i Z."""
import numpy as np

PX = np.array([[0, 1], [1, 0]], complex)
PY = np.array([[0, -1j], [1j, 0]])
PZ = np.array([[1, 0], [0, -1]], complex)


def site_op(op, i, L):
    m = np.ones((1, 1), complex)
    for s in reversed(range(L)):
        m = np.kron(m, op if s == i else np.eye(2))
    return m


def test_xxz_matches_kron(ra, orc):
    L, J, D, W = 6, 0.9, 0.7, 2.5
    for fp in (32, 64):
        bits, coeff = ra.xxz_terms(L, J, D, W, seed=12, fp=fp)
        assert len(bits) == 4 * L - 3
        H = orc.dense_matrix(L, bits, coeff)
        fields = coeff[-L:].real
        assert np.all(np.abs(fields) <= W)
        ref = sum(J * (site_op(PX, i, L) @ site_op(PX, i + 1, L)
                       + site_op(PY, i, L) @ site_op(PY, i + 1, L)
                       + D * site_op(PZ, i, L) @ site_op(PZ, i + 1, L)) for i in range(L - 1))
        ref = ref + sum(fields[i] * site_op(PZ, i, L) for i in range(L))
        tol = 1e-6 if fp == 32 else 1e-14
        assert np.abs(H - ref).max() < tol
        assert np.abs(H - H.conj().T).max() == 0


def test_xxz_fields_are_seeded(ra):
    a = ra.xxz_terms(8, 1, 1, 1, seed=3, fp=64)[1]
    b = ra.xxz_terms(8, 1, 1, 1, seed=3, fp=64)[1]
    c = ra.xxz_terms(8, 1, 1, 1, seed=4, fp=64)[1]
    assert np.array_equal(a, b) and not np.array_equal(a, c)
