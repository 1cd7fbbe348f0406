"""The pinning checks catch plausible mistakes.  Each planted variant of the
oracle (built with -DOR_PLANT=n, see oracle/oracle.c) differs from the
restatement in one place; the same known-answer or library checks that pass
for the restatement must fail for it.  Mistakes are planted on the CPU, in
the oracle only."""
import numpy as np
import pytest

import oracle


def _product_table_ok():
    """tests/ops.cc:26-44 through the oracle's product."""
    SX, SZ, SY = 1, 2, 3
    ops = {"sx": (SX, 1.0), "sy": (SY, 1j), "sz": (SZ, 1.0)}
    table = {("sx", "sy"): ("sz", 1j), ("sy", "sx"): ("sz", -1j), ("sz", "sx"): ("sy", 1j),
             ("sx", "sz"): ("sy", -1j), ("sy", "sz"): ("sx", 1j), ("sz", "sy"): ("sx", -1j)}
    for (a, b), (r, f) in table.items():
        bb, minus = oracle.pauli_mul(ops[a][0], ops[b][0])
        c = (-ops[a][1] if minus else ops[a][1]) * ops[b][1]
        if bb != ops[r][0] or c != f * ops[r][1]:
            return False
    return True


def _matrix_kat_ok():
    """tests/ops.cc:132-136 (sy1*sx) through the oracle's apply."""
    want = np.array([[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]])
    bits = np.array([(3 << 2) | 1], np.uint64)
    m = np.stack([oracle.apply(2, bits, np.array([1j]), np.eye(4)[k]) for k in range(4)], axis=1)
    return np.array_equal(m, want)


def _krylov_ed_ok():
    """restarted Krylov (several restarts) against exact diagonalization"""
    L = 6
    bits, coeff = oracle.syk_terms(12, 0.0, seed=1)
    H = oracle.dense_matrix(L, bits, coeff)
    rng = np.random.default_rng(0)
    psi = rng.standard_normal(64) + 1j * rng.standard_normal(64)
    psi /= np.linalg.norm(psi)
    w, Q = np.linalg.eigh(H)
    ex = Q @ (np.exp(-4j * w) * (Q.conj().T @ psi))
    got, st = oracle.krylov_evolve(L, bits, coeff, psi, 4.0, 0.0, krydim=4, tol=1e-13)
    return st.restarts > 1 and np.abs(got - ex).max() < 1e-10


CHECKS = [_product_table_ok, _matrix_kat_ok, _krylov_ed_ok]


def test_restatement_passes_every_check():
    assert all(c() for c in CHECKS)


@pytest.mark.parametrize("plant,check", [(1, _product_table_ok), (2, _matrix_kat_ok),
                                         (3, _krylov_ed_ok)])
def test_planted_mistake_is_caught(plant, check):
    with oracle.use(oracle.planted(plant)):
        assert not check()
