"""Pinning of the oracle's restatement of libstdc++ <random>, which the SYK
generators (inc/syk.h:58, inc/syk.h:223-245) draw from.

Checks: the C++ standard's mt19937 known answer; numpy's independent MT19937
(legacy RandomState seeding is the standard init_genrand); and stream-for-
stream equality with libstdc++ itself (normal_distribution with its cached
second polar sample, bernoulli_distribution, uniform_int_distribution,
generate_canonical) called through the shipped C++ library."""
import numpy as np
import pytest


def test_mt19937_known_answer(orc):
    # C++ [rand.predef]: the 10000th invocation of a default-constructed
    # mt19937 produces 4123659995
    assert orc.lib().or_mt19937_kat() == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 5489, 2000, 123456789])
def test_mt19937_matches_numpy(orc, seed):
    got = orc.mt_stream(seed, 5000)
    ref = np.random.RandomState(seed).randint(0, 2 ** 32, size=5000, dtype=np.uint64).astype(np.uint32)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("fp", [32, 64])
@pytest.mark.parametrize("seed", [0, 7, 2000])
def test_normal_distribution_matches_libstdcxx(ra, orc, fp, seed):
    n = 20001                       # odd: ends on a freshly generated pair
    got = orc.normal_stream(seed, n, fp=fp).astype(np.float64)
    ref = ra.random_stream("normal", n, seed, fp=fp)
    assert np.array_equal(got, ref)
    assert abs(got.mean()) < 0.03 and abs(got.var() - 1) < 0.05


@pytest.mark.parametrize("p", [0.001973, 0.5, 0.9999])
def test_bernoulli_matches_libstdcxx(ra, orc, p):
    got = orc.bernoulli_stream(3, p, 20000)
    ref = ra.random_stream("bernoulli", 20000, 3, a=p)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("a,b", [(0, 47), (0, 59), (0, 3), (5, 5)])
def test_uniform_int_matches_libstdcxx(ra, orc, a, b):
    got = orc.uniform_int_stream(11, a, b, 20000)
    ref = ra.random_stream("uniform_int", 20000, 11, a=a, b=b)
    assert np.array_equal(got, ref)
    assert got.min() >= a and got.max() <= b


@pytest.mark.parametrize("fp", [32, 64])
def test_generate_canonical_matches_libstdcxx(ra, orc, fp):
    got = orc.canonical_stream(99, 20000, fp=fp)
    ref = ra.random_stream("canonical", 20000, 99, fp=fp)
    assert np.array_equal(got, ref)
    assert got.min() >= 0 and got.max() < 1
