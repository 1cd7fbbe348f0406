"""Pinning of the Jordan-Wigner Majorana strings (inc/fermions.h:254-306)
and the SYK generators (inc/syk.h:88-262), in the oracle and the shipped C++
API.

Known answers: the reference's worked examples of b_n for N = 2, 4, 6, 8
(inc/fermions.h:236-245) and the anticommutation relations the reference's
unit tests check for N = 4..20 (tests/ops.cc:299-346).  The SYK term lists
from the C++ API (which draws from libstdc++ <random>) and from the oracle
(its C restatement) must agree bit for bit; their statistics must match the
model's closed forms."""
import math

import numpy as np
import pytest

CODE = {"I": 0, "sx": 1, "sz": 2, "sy": 3}


def string_bits(ops_high_to_low):
    """Bits of a string written as in the reference's comments, highest site
    first (sx.sx.sy means sx(2).sx(1).sy(0))."""
    b = 0
    n = len(ops_high_to_low)
    for k, op in enumerate(ops_high_to_low):
        b |= CODE[op] << (2 * (n - 1 - k))
    return b


WORKED = {  # inc/fermions.h:236-245
    2: [["sx"], ["sy"]],
    4: [["sx", "sx"], ["sx", "sy"], ["sx", "sz"], ["sy", "I"]],
    6: [["sx", "sx", "sx"], ["sx", "sx", "sy"], ["sx", "sx", "sz"], ["sx", "sy", "I"],
        ["sx", "sz", "I"], ["sy", "I", "I"]],
    8: [["sx", "sx", "sx", "sx"], ["sx", "sx", "sx", "sy"], ["sx", "sx", "sx", "sz"],
        ["sx", "sx", "sy", "I"], ["sx", "sx", "sz", "I"], ["sx", "sy", "I", "I"],
        ["sx", "sz", "I", "I"], ["sy", "I", "I", "I"]],
}


@pytest.mark.parametrize("N", sorted(WORKED))
def test_majorana_worked_examples(ra, orc, N):
    for n, ops in enumerate(WORKED[N]):
        want = string_bits(ops)
        for fp in (32, 64):
            b, c = orc.majorana(N, n, fp=fp)
            assert b == want, (N, n)
            # coefficient: i for odd n (sigma_y), over sqrt(2)
            expect = (1j if n & 1 else 1.0) * float(np.float32(1) / np.float32(np.sqrt(2.0))
                                                 if fp == 32 else 1 / np.sqrt(2.0))
            assert c == expect
            assert ra.majorana(N, n, fp=fp) == (b, c)


def test_majorana_closed_form_masks(orc):
    """flip mask = sites [n/2, N/2), phase mask = site (n-1)/2 for n >= 1."""
    for N in range(2, 65, 2):
        L = N // 2
        for n in range(N):
            b, _ = orc.majorana(N, n)
            x = sum(((b >> (2 * s)) & 1) << s for s in range(L))
            z = sum(((b >> (2 * s + 1)) & 1) << s for s in range(L))
            assert x == ((1 << L) - 1) & ~((1 << (n // 2)) - 1)
            assert z == (0 if n == 0 else 1 << ((n - 1) // 2))


@pytest.mark.parametrize("N", range(4, 21, 2))
def test_anticommutation(ra, orc, N):
    """{g_i, g_j} = delta_ij (standard), with 2*(1/sqrt2)^2 in FP64 on the
    diagonal (tests/ops.cc:299-319); 2*delta_ij for the non-standard form."""
    one = (1.0 / math.sqrt(2)) * (1.0 / math.sqrt(2)) * 2
    for i in range(N):
        for j in range(N):
            bits, c = orc.majorana_anticomm(N, i, j)
            if i != j:
                assert len(bits) == 0
            else:
                assert list(bits) == [0] and c[0] == one
            bits2, c2 = orc.majorana_anticomm(N, i, j, std_gamma=False)
            if i != j:
                assert len(bits2) == 0
            else:
                assert list(bits2) == [0] and c2[0] == 2
    # the same products through the C++ API's HostSumOps
    for i in range(0, N, 3):
        for j in range(0, N, 2):
            gi = ra.majorana(N, i, fp=64)
            gj = ra.majorana(N, j, fp=64)
            A = (np.array([gi[0]], np.uint64), np.array([gi[1]]))
            B = (np.array([gj[0]], np.uint64), np.array([gj[1]]))
            ab = ra.ops_mul(A, B)
            ba = ra.ops_mul(B, A)
            total = {}
            for bb, cc in list(zip(*ab)) + list(zip(*ba)):
                total[int(bb)] = total.get(int(bb), 0) + cc
            total = {k: v for k, v in total.items() if v != 0}
            assert total == ({0: one} if i == j else {})


CASES = [(N, sp, fp, reg) for (N, sp) in [(8, 0.0), (12, 0.0), (24, 0.0), (16, 2.0), (32, 8.0),
                                        (48, 8.0), (60, 8.0), (40, 4.0)]
         for fp in (32, 64) for reg in (0, 1) if not (reg and sp == 0.0)]


@pytest.mark.parametrize("N,sp,fp,reg", CASES)
def test_syk_cpp_matches_oracle_bit_exact(ra, orc, N, sp, fp, reg):
    for seed in (1, 2000):
        b1, c1 = ra.syk_terms(N, sp, seed=seed, fp=fp, regularize=bool(reg))
        b2, c2 = orc.syk_terms(N, sp, seed=seed, fp=fp, regularize=bool(reg))
        assert np.array_equal(b1, b2)
        assert np.array_equal(c1.view(np.float64), c2.view(np.float64))


def test_syk_dense_structure(orc):
    """Dense SYK: C(N,4) distinct Hermitian terms, one per quadruple, coupling
    |c| = scale*|J|/4 with J ~ N(0,1) and scale = sqrt(6/N^3)
    (inc/syk.h:106-115)."""
    N = 24
    bits, coeff = orc.syk_terms(N, 0.0, seed=5)
    assert len(bits) == math.comb(N, 4) == len(set(bits.tolist()))
    L = N // 2
    x = np.array([sum(((int(b) >> (2 * s)) & 1) << s for s in range(L)) for b in bits])
    z = np.array([sum(((int(b) >> (2 * s + 1)) & 1) << s for s in range(L)) for b in bits])
    nsy = np.array([bin(int(a) & int(c)).count("1") for a, c in zip(x, z)])
    std = coeff * (-1j) ** nsy         # coefficient() (inc/ops.h:366-377)
    assert np.all(np.abs(std.imag) <= 1e-15 * np.abs(std.real))
    J = std.real * 4 / math.sqrt(6.0 / N ** 3)
    assert abs(J.mean()) < 0.05 and abs(J.var() - 1) < 0.05
    # the top qubit's X flips cancel in every product of four Majoranas
    assert np.all(((x >> (L - 1)) & 1) == 0)


def test_syk_sparse_statistics(orc):
    """Sparse SYK keeps each quadruple with p = kN/C(N,4) and rescales by
    1/sqrt(p) (inc/syk.h:111-115): T ~ Binomial(C(N,4), p)."""
    N, k = 48, 8.0
    total = math.comb(N, 4)
    p = k * N / total
    counts = [len(orc.syk_terms(N, k, seed=s)[0]) for s in range(30)]
    mean, sd = np.mean(counts), np.sqrt(total * p * (1 - p))
    assert abs(mean - total * p) < 4 * sd / np.sqrt(30)
    bits, coeff = orc.syk_terms(N, k, seed=3)
    scale = math.sqrt(6.0 / N ** 3) / math.sqrt(p)
    J = np.abs(coeff) * 4 / scale
    assert abs(np.mean(J ** 2) - 1) < 0.25


def test_syk_regularized_degrees(orc):
    """The regularized generator builds an almost 4k-regular hypergraph with
    about kN edges (inc/syk.h:152-234)."""
    N, k = 32, 8.0
    bits, _ = orc.syk_terms(N, k, seed=4, regularize=True)
    assert abs(len(bits) - k * N) <= 0.05 * k * N
