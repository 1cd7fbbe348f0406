"""Pinning of the Pauli-string algebra and apply semantics (inc/ops.h) in
the oracle and in the shipped C++ API.

Known answers from the reference's own unit tests: the single-site product
table (tests/ops.cc:26-44) and the explicit 1-, 2- and 4-dimensional matrices
of id, sx, sy, sz and their two-site products (tests/ops.cc:46-161), which fix
the site <-> index-bit order.  Independently: dense Kronecker products built
with numpy from the textbook Pauli matrices, the COO construction
(inc/ops.h:553-586) against the bit-loop apply (inc/ops.h:421-440), and the
oracle's apply against both."""
import itertools

import numpy as np
import pytest

ID, SX, SZ, SY = 0, 1, 2, 3          # 2-bit codes (inc/ops.h:133-143)
I2 = np.eye(2)
PX = np.array([[0, 1], [1, 0]], complex)
PY = np.array([[0, -1j], [1j, 0]])
PZ = np.array([[1, 0], [0, -1]], complex)


def sigma_bits(code, site):
    return code << (2 * site)


def coeff_of(code):
    """SpinOp coefficient of sigma_{x,y,z}: sigma_y carries its i
    (inc/ops.h:285-288)."""
    return 1j if code == SY else 1.0


def kron_matrix(bits, coeff, length):
    """Textbook matrix of the string: site s acts on index bit s; the SpinOp
    coefficient multiplies sy = sx.sz = -i sigma_y."""
    m = np.ones((1, 1), complex)
    for s in reversed(range(length)):
        code = (bits >> (2 * s)) & 3
        op = {ID: I2, SX: PX, SZ: PZ, SY: -1j * PY}[code]
        m = np.kron(m, op)
    return coeff * m


def product(orc, b0, c0, b1, c1):
    b, minus = orc.pauli_mul(b0, b1)
    return b, (-c0 if minus else c0) * c1


def test_single_site_product_table(orc):
    """tests/ops.cc:26-44 (i = the imaginary unit, sigma_y = i*sy)."""
    i = 1j
    ops = {"id": (0, 1.0), "sx": (SX, 1.0), "sy": (SY, 1j), "sz": (SZ, 1.0)}
    table = {
        ("id", "id"): ("id", 1), ("id", "sx"): ("sx", 1), ("id", "sy"): ("sy", 1),
        ("id", "sz"): ("sz", 1), ("sx", "id"): ("sx", 1), ("sx", "sx"): ("id", 1),
        ("sx", "sy"): ("sz", i), ("sx", "sz"): ("sy", -i), ("sy", "id"): ("sy", 1),
        ("sy", "sx"): ("sz", -i), ("sy", "sy"): ("id", 1), ("sy", "sz"): ("sx", i),
        ("sz", "id"): ("sz", 1), ("sz", "sx"): ("sy", i), ("sz", "sy"): ("sx", -i),
        ("sz", "sz"): ("id", 1),
    }
    for (a, b), (r, f) in table.items():
        bb, cc = product(orc, ops[a][0], ops[a][1], ops[b][0], ops[b][1])
        assert bb == ops[r][0], (a, b)
        assert cc == f * ops[r][1], (a, b, cc)


def test_product_sign_rule_closed_form(orc):
    """minus iff an odd number of sites have a Z bit on the left factor and an
    X bit on the right (inc/ops.h:151-161)."""
    rng = np.random.default_rng(0)
    for _ in range(2000):
        b0 = int(rng.integers(0, 1 << 40))
        b1 = int(rng.integers(0, 1 << 40))
        b, minus = orc.pauli_mul(b0, b1)
        z0 = (b0 >> 1) & 0x5555555555555555
        x1 = b1 & 0x5555555555555555
        assert b == b0 ^ b1
        assert minus == bool(bin(z0 & x1).count("1") & 1)


MATRICES = {
    # tests/ops.cc:46-161: the string, the chain length, the expected matrix
    "id1": (0, 1.0, 1, [[1, 0], [0, 1]]),
    "sx": (SX, 1.0, 1, [[0, 1], [1, 0]]),
    "sy": (SY, 1j, 1, [[0, -1j], [1j, 0]]),
    "sz": (SZ, 1.0, 1, [[1, 0], [0, -1]]),
    "id*sx": (SX, 1.0, 2, [[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]]),
    "id*sy": (SY, 1j, 2, [[0, -1j, 0, 0], [1j, 0, 0, 0], [0, 0, 0, -1j], [0, 0, 1j, 0]]),
    "id*sz": (SZ, 1.0, 2, [[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, 1, 0], [0, 0, 0, -1]]),
    "sx1*id": (SX << 2, 1.0, 2, [[0, 0, 1, 0], [0, 0, 0, 1], [1, 0, 0, 0], [0, 1, 0, 0]]),
    "sx1*sx": ((SX << 2) | SX, 1.0, 2, [[0, 0, 0, 1], [0, 0, 1, 0], [0, 1, 0, 0], [1, 0, 0, 0]]),
    "sx1*sy": ((SX << 2) | SY, 1j, 2, [[0, 0, 0, -1j], [0, 0, 1j, 0], [0, -1j, 0, 0], [1j, 0, 0, 0]]),
    "sx1*sz": ((SX << 2) | SZ, 1.0, 2, [[0, 0, 1, 0], [0, 0, 0, -1], [1, 0, 0, 0], [0, -1, 0, 0]]),
    "sy1*id": (SY << 2, 1j, 2, [[0, 0, -1j, 0], [0, 0, 0, -1j], [1j, 0, 0, 0], [0, 1j, 0, 0]]),
    "sy1*sx": ((SY << 2) | SX, 1j, 2, [[0, 0, 0, -1j], [0, 0, -1j, 0], [0, 1j, 0, 0], [1j, 0, 0, 0]]),
    "sy1*sy": ((SY << 2) | SY, -1.0, 2, [[0, 0, 0, -1], [0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0]]),
    "sy1*sz": ((SY << 2) | SZ, 1j, 2, [[0, 0, -1j, 0], [0, 0, 0, 1j], [1j, 0, 0, 0], [0, -1j, 0, 0]]),
    "sz1*id": (SZ << 2, 1.0, 2, [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, -1, 0], [0, 0, 0, -1]]),
    "sz1*sx": ((SZ << 2) | SX, 1.0, 2, [[0, 1, 0, 0], [1, 0, 0, 0], [0, 0, 0, -1], [0, 0, -1, 0]]),
    "sz1*sy": ((SZ << 2) | SY, 1j, 2, [[0, -1j, 0, 0], [1j, 0, 0, 0], [0, 0, 0, 1j], [0, 0, -1j, 0]]),
    "sz1*sz": ((SZ << 2) | SZ, 1.0, 2, [[1, 0, 0, 0], [0, -1, 0, 0], [0, 0, -1, 0], [0, 0, 0, 1]]),
}


@pytest.mark.parametrize("name", sorted(MATRICES))
def test_matrix_representation(orc, name):
    bits, coeff, length, expect = MATRICES[name]
    m = orc.dense_matrix(length, np.array([bits], np.uint64), np.array([coeff]))
    assert np.array_equal(m, np.array(expect, complex)), name
    # the same matrix from textbook Kronecker products
    assert np.allclose(kron_matrix(bits, coeff, length), m, atol=0)


def test_two_site_products_via_algebra(orc):
    """The two-site strings of the table above, formed as the reference forms
    them (sigma(1) * sigma(0) products), reproduce the same matrices."""
    for c1, c0 in itertools.product((ID, SX, SY, SZ), repeat=2):
        b1, k1 = sigma_bits(c1, 1), coeff_of(c1)
        b0, k0 = sigma_bits(c0, 0), coeff_of(c0)
        b, c = product(orc, b1, k1, b0, k0)
        m = orc.dense_matrix(2, np.array([b], np.uint64), np.array([c]))
        ref = np.kron({ID: I2, SX: PX, SY: PY, SZ: PZ}[c1], {ID: I2, SX: PX, SY: PY, SZ: PZ}[c0])
        assert np.allclose(m, ref, atol=0)


def test_apply_vs_sparse_matrix_vs_kron(orc):
    """Three constructions of one operator: the bit-loop apply, the COO
    doubling construction and numpy Kronecker products (length 7)."""
    rng = np.random.default_rng(2)
    length = 7
    for _ in range(40):
        bits = int(rng.integers(0, 1 << (2 * length)))
        cols, neg = orc.sparse_matrix(bits, length)
        K = kron_matrix(bits, 1.0, length)
        for n in range(1 << length):
            res, minus = orc.pauli_apply(bits, n)
            # row n of the operator has its one entry at column res
            assert res == cols[n]
            assert minus == neg[n]
            assert K[n, res] == (-1 if minus else 1)
            assert np.count_nonzero(K[n]) == 1


def test_oracle_apply_matches_dense(orc):
    rng = np.random.default_rng(3)
    length = 8
    bits = rng.integers(0, 1 << (2 * length), size=30).astype(np.uint64)
    coeff = rng.standard_normal(30) + 1j * rng.standard_normal(30)
    psi = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    H = sum(kron_matrix(int(b), c, length) for b, c in zip(bits, coeff))
    assert np.allclose(orc.apply(length, bits, coeff, psi), H @ psi, atol=1e-12)
    assert np.allclose(orc.dense_matrix(length, bits, coeff), H, atol=1e-12)


def test_product_merge_semantics_cpp_vs_oracle(ra, orc):
    """HostSumOps::operator* (all products, merged in order; inc/cpuimpl.h:
    233-253, 325-333) in the C++ API against the oracle's product rule."""
    rng = np.random.default_rng(4)
    for _ in range(20):
        a = (rng.integers(0, 1 << 16, size=5).astype(np.uint64),
             rng.standard_normal(5) + 1j * rng.standard_normal(5))
        b = (rng.integers(0, 1 << 16, size=4).astype(np.uint64),
             rng.standard_normal(4) + 1j * rng.standard_normal(4))
        bo, co = ra.ops_mul(a, b)
        # reference semantics restated in Python: merge by bits, drop zeros
        merged = {}
        order = []
        for b0, c0 in zip(*a):
            for b1, c1 in zip(*b):
                bb, cc = product(orc, int(b0), c0, int(b1), c1)
                if bb in merged:
                    merged[bb] += cc
                    if merged[bb] == 0:
                        del merged[bb]
                        order.remove(bb)
                else:
                    merged[bb] = cc
                    order.append(bb)
        assert [int(x) for x in bo] == order
        assert np.allclose(co, [merged[x] for x in order], rtol=0, atol=1e-14)


def test_oracle_row_windows_match_dense(orc):
    """apply_rows (the row-window form the full-size GPU tests sample)
    returns exactly the rows of the Kronecker-product matrix times psi."""
    rng = np.random.default_rng(9)
    length = 8
    bits = rng.integers(0, 1 << (2 * length), size=25).astype(np.uint64)
    coeff = rng.standard_normal(25) + 1j * rng.standard_normal(25)
    psi = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    H = sum(kron_matrix(int(b), c, length) for b, c in zip(bits, coeff))
    ref = H @ psi
    for lo, hi in [(0, 256), (0, 1), (37, 101), (255, 256)]:
        got = orc.apply_rows(length, bits, coeff, psi, lo, hi)
        assert got.shape == (hi - lo,)
        assert np.abs(got - ref[lo:hi]).max() < 1e-12 * np.abs(ref).max()
    # FP32 arithmetic on a complex64 state read in place
    c32 = coeff.astype(np.complex64).astype(np.complex128)
    p32 = psi.astype(np.complex64)
    ref32 = sum(kron_matrix(int(b), c, length) for b, c in zip(bits, c32)) @ p32.astype(np.complex128)
    got = orc.apply_rows(length, bits, c32, p32, 64, 200, fp=32)
    assert np.abs(got - ref32[64:200]).max() < 1e-5 * np.abs(ref32).max()
