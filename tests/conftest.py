"""Shared fixtures.  Tests marked `gpu` need an MI355X and the built HIP
library; everything else runs on the CPU (oracle, host C++ API, ABI exports)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a GPU (MI355X) and libreapers_hip.so")


@pytest.fixture(scope="session")
def ra():
    import reapers_amd
    return reapers_amd


@pytest.fixture(scope="session")
def gpu(ra):
    """Initialized HIP backend; the test fails (does not skip) if it cannot
    initialize, so a GPU run never passes on a silent fallback."""
    ra.init(0)
    return ra


@pytest.fixture(scope="session")
def orc():
    import oracle
    return oracle


def random_pauli_sum(rng, length, nterms, hermitian=False, max_sites=None):
    """Random term list: SpinOp bits over `length` sites and complex coeffs."""
    bits = []
    for _ in range(nterms):
        b = 0
        sites = range(length) if max_sites is None else rng.choice(length, size=min(max_sites, length), replace=False)
        for s in sites:
            b |= int(rng.integers(0, 4)) << (2 * int(s))
        bits.append(b)
    bits = np.array(bits, dtype=np.uint64)
    coeff = rng.standard_normal(nterms) + 1j * rng.standard_normal(nterms)
    return bits, coeff


def random_state(rng, length, dtype=np.complex128):
    dim = 1 << length
    v = rng.standard_normal(dim) + 1j * rng.standard_normal(dim)
    v /= np.linalg.norm(v)
    return v.astype(dtype)
