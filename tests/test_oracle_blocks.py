"""Pinning of the parity-block form: FermionOpBlockForm (inc/fermions.h:
308-368), the block arithmetic (inc/blkops.h) and SYKBlockForm's Hamiltonian
blocks, in the oracle and the shipped C++ API.

Known answers: the reference's description of the blocks of b_n
(inc/fermions.h:311-327: for n <= N-3 both blocks are the (N-2)-fermion
b_n, for n = N-2 both are sigma_z(N/2-2), for n = N-1 LR = -i 1 and
RL = +i 1, all over sqrt(2) in the standard normalization) and the
anticommutation relations of the reference's unit test
(tests/ops.cc:321-346).  What the mathematics fixes: the block matrices
assembled into full 2^(N/2) x 2^(N/2) operators must satisfy the Clifford
algebra, the Hamiltonian blocks must be the diagonal blocks of
sum J_ijkl b_i b_j b_k b_l formed from those full operators, and the
spectrum of LL (+) RR must equal the non-block Hamiltonian's (both are
representations of the same Clifford algebra with the same couplings)."""
import math

import numpy as np
import pytest

SQ = 1.0 / math.sqrt(2.0)


def block_majorana_full(orc, N, n):
    """The full block anti-diagonal matrix [[0, LR], [RL, 0]] of b_n."""
    L = N // 2 - 1
    blocks = []
    for which in (0, 1):
        b, c = orc.majorana_block(N, n, which)
        blocks.append(orc.dense_matrix(L, np.array([b], np.uint64), np.array([c])))
    z = np.zeros_like(blocks[0])
    return np.block([[z, blocks[0]], [blocks[1], z]])


@pytest.mark.parametrize("N", [4, 6, 8, 12, 20])
def test_block_majorana_known_answers(ra, orc, N):
    for n in range(N):
        for fp in (32, 64):
            inv = float(np.float32(SQ)) if fp == 32 else SQ
            lr = orc.majorana_block(N, n, 0, fp=fp)
            rl = orc.majorana_block(N, n, 1, fp=fp)
            if n == N - 1:
                assert lr == (0, -1j * inv) and rl == (0, 1j * inv)
            elif n == N - 2:
                sz = 2 << (2 * (N // 2 - 2))
                assert lr == rl == (sz, inv)
            else:
                b, c = orc.majorana(N - 2, n, fp=fp, std_gamma=False)
                assert lr == rl == (b, c * inv)
            # the C++ API builds the same blocks
            assert ra.majorana_block(N, n, 0, fp=fp) == lr
            assert ra.majorana_block(N, n, 1, fp=fp) == rl
        # non-standard normalization: no 1/sqrt(2)
        lr = orc.majorana_block(N, n, 0, std_gamma=False)
        assert abs(abs(lr[1]) - 1.0) == 0


@pytest.mark.parametrize("N", [4, 6, 8, 10])
def test_block_majoranas_form_a_clifford_algebra(orc, N):
    """{b_i, b_j} = delta_ij for the assembled full operators."""
    g = [block_majorana_full(orc, N, n) for n in range(N)]
    eye = np.eye(g[0].shape[0])
    for i in range(N):
        assert np.allclose(g[i], g[i].conj().T, atol=0)
        for j in range(N):
            ac = g[i] @ g[j] + g[j] @ g[i]
            assert np.abs(ac - (eye if i == j else 0)).max() < 1e-15


@pytest.mark.parametrize("N", range(4, 21, 2))
def test_block_anticommutation_cpp(ra, N):
    """tests/ops.cc:321-346: {g_i, g_j} has LL == RR == 0 off the diagonal
    and 2 (1/sqrt2)^2 (in FP64) on it; 2 delta_ij without the 1/sqrt2."""
    one = SQ * SQ * 2
    for i in range(N):
        for j in range(0, N, max(1, N // 6)):
            for which in (0, 1):
                b, c, null = ra.block_anticomm(N, i, j, which)
                # a sum's blocks are null when empty (inc/blkops.h:465-471)
                assert null == (i != j)
                if i != j:
                    assert len(b) == 0
                else:
                    assert list(b) == [0] and c[0] == one
                b2, c2, _ = ra.block_anticomm(N, i, j, which, std_gamma=False)
                if i != j:
                    assert len(b2) == 0
                else:
                    assert list(b2) == [0] and c2[0] == 2


def dense_couplings(orc, N, seed, fp=64):
    """scale * J_ijkl in the generator's loop order i > j > k > l."""
    J = orc.normal_stream(seed, math.comb(N, 4), fp=fp).astype(np.float64)
    scale = math.sqrt(6.0 / N ** 3)
    quads = [(i, j, k, l) for i in range(N) for j in range(i) for k in range(j) for l in range(k)]
    return quads, scale * J


@pytest.mark.parametrize("N", [6, 8, 10])
def test_block_hamiltonian_is_diagonal_of_full_sum(orc, N):
    """ham.LL and ham.RR are the diagonal blocks of sum c_q b_i b_j b_k b_l
    built from the full block operators, and the off-diagonal blocks of that
    sum vanish (parity conservation)."""
    seed = 17 + N
    quads, cq = dense_couplings(orc, N, seed)
    g = [block_majorana_full(orc, N, n) for n in range(N)]
    H = sum(c * (g[i] @ g[j] @ g[k] @ g[l]) for (i, j, k, l), c in zip(quads, cq))
    half = H.shape[0] // 2
    L = N // 2 - 1
    LL = orc.dense_matrix(L, *orc.syk_block_terms(N, 0, seed=seed))
    RR = orc.dense_matrix(L, *orc.syk_block_terms(N, 1, seed=seed))
    scale = np.abs(H).max()
    assert np.abs(H[:half, half:]).max() < 1e-15 * scale * len(quads)
    assert np.abs(H[half:, :half]).max() < 1e-15 * scale * len(quads)
    assert np.abs(H[:half, :half] - LL).max() < 1e-14 * scale * len(quads)
    assert np.abs(H[half:, half:] - RR).max() < 1e-14 * scale * len(quads)


@pytest.mark.parametrize("N,sp,reg", [(8, 0.0, 0), (10, 0.0, 0), (12, 0.0, 0), (12, 2.0, 0),
                                      (12, 2.0, 1)])
def test_block_spectrum_equals_non_block(orc, N, sp, reg):
    seed = 5 + N
    L = N // 2
    H = orc.dense_matrix(L, *orc.syk_terms(N, sp, seed=seed, regularize=bool(reg)))
    LL = orc.dense_matrix(L - 1, *orc.syk_block_terms(N, 0, sp, seed=seed, regularize=bool(reg)))
    RR = orc.dense_matrix(L - 1, *orc.syk_block_terms(N, 1, sp, seed=seed, regularize=bool(reg)))
    w = np.linalg.eigvalsh(H)
    wb = np.sort(np.concatenate([np.linalg.eigvalsh(LL), np.linalg.eigvalsh(RR)]))
    assert np.abs(w - wb).max() < 1e-12 * max(1.0, np.abs(w).max())
    # the two sectors differ (the terms with b_{N-1} change sign)
    assert np.abs(LL - RR).max() > 1e-3 * np.abs(LL).max()


CASES = [(N, sp, fp, reg) for (N, sp) in [(8, 0.0), (12, 0.0), (24, 0.0), (32, 8.0), (48, 8.0),
                                        (60, 8.0)]
         for fp in (32, 64) for reg in (0, 1) if not (reg and sp == 0.0)]


@pytest.mark.parametrize("N,sp,fp,reg", CASES)
def test_syk_block_cpp_matches_oracle_bit_exact(ra, orc, N, sp, fp, reg):
    for seed in (1, 2000):
        for which in (0, 1):
            b1, c1 = ra.syk_block_terms(N, which, sp, seed=seed, fp=fp, regularize=bool(reg))
            b2, c2 = orc.syk_block_terms(N, which, sp, seed=seed, fp=fp, regularize=bool(reg))
            assert np.array_equal(b1, b2)
            assert np.array_equal(c1.view(np.float64), c2.view(np.float64))


def test_syk_block_term_counts(orc):
    """No two quadruples share a string in a sector for N > 8; for N = 8 a
    quadruple and its complement coincide (their product is the sector's
    parity), so each block has C(8,4)/2 strings."""
    assert len(orc.syk_block_terms(8, 0)[0]) == 35
    for N in (10, 16, 24):
        assert len(orc.syk_block_terms(N, 0)[0]) == math.comb(N, 4)
        assert len(orc.syk_block_terms(N, 1)[0]) == math.comb(N, 4)


def test_result_file_format(ra, tmp_path):
    """One 't Re Im' line per time point in an ostream's default format
    (examples/syk.cc:230-235); 17 digits round-trip exactly."""
    t = np.array([0.0, 0.5, 1.0, 12.5])
    v = np.array([1.0 + 0j, 0.25 - 0.5j, -1e-3 + 2e-9j, 0.123456789 - 1e-300j])
    assert ra.format_result_lines(t, v) == [
        "0 1 0", "0.5 0.25 -0.5", "1 -0.001 2e-09", "12.5 0.123457 -1e-300"]
    p = tmp_path / "otoc.txt"
    ra.write_result_file(str(p), t, v, digits=17)
    t2, v2 = ra.read_result_file(str(p))
    assert np.array_equal(t2, t) and np.array_equal(v2, v)
