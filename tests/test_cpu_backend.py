"""The CPU backend (CPUImpl, the C++ API built with REAPERS_NOGPU; the
reference's CPUImpl, inc/cpuimpl.h:513-755, and its backend switch,
inc/reapers.h:96-100) against the oracle and numpy, and the reference's own
example programs compiled against this build's headers.

Tolerances: FP64 applies within 1e-13 of the oracle (both sum in FP64, in
different orders); FP32 applies within 1e-6 of the largest amplitude (FP32
inputs and outputs); FP64 evolutions within 1e-11 of the oracle and of exact
diagonalization; eigenvalues within 1e-12 and A V = V w within 1e-12."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import random_pauli_sum, random_state

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = [os.path.join(ROOT, "reapers_amd", "include", "reapers"), os.path.join(ROOT, "include")]
REF_EXAMPLES = "/root/reference/examples"


@pytest.fixture(scope="module")
def cpu():
    from reapers_amd import cpu as c
    return c


def ed_evolve(H, psi, t, beta=0.0):
    w, Q = np.linalg.eigh(H)
    return Q @ (np.exp((-beta - 1j * t) * w) * (Q.conj().T @ psi))


@pytest.mark.parametrize("length", [1, 3, 8, 13, 14])
def test_apply_random_sums_vs_oracle(cpu, orc, length):
    rng = np.random.default_rng(length)
    bits, coeff = random_pauli_sum(rng, length, 17)
    psi = random_state(rng, length)
    acc = random_state(rng, length)
    ref = orc.apply(length, bits, coeff, psi)
    assert np.abs(cpu.apply(length, bits, coeff, psi) - ref).max() < 1e-13
    ref2 = orc.apply(length, bits, coeff, psi, out=acc)
    assert np.abs(cpu.apply(length, bits, coeff, psi, out=acc) - ref2).max() < 1e-13


def test_apply_fp32_sparse_syk_vs_oracle(cpu, orc):
    L = 16
    bits, coeff = orc.syk_terms(32, 8.0, seed=7, fp=32)
    psi = random_state(np.random.default_rng(3), L, np.complex64)
    ref = orc.apply(L, bits, coeff, psi.astype(np.complex128))
    got = cpu.apply(L, bits, coeff, psi, fp=32).astype(np.complex128)
    assert np.abs(got - ref).max() / np.abs(ref).max() < 1e-6


def test_apply_dense_switch_vs_oracle(cpu, orc):
    """At least 2^len terms: the reference applies the dense matrix
    (inc/cpuimpl.h:737-741); dense SYK N=12 has 495 >= 2^6 strings."""
    L = 6
    bits, coeff = orc.syk_terms(12, 0.0, seed=11, fp=64)
    assert len(bits) >= 1 << L
    psi = random_state(np.random.default_rng(5), L)
    ref = orc.apply(L, bits, coeff, psi)
    assert np.abs(cpu.apply(L, bits, coeff, psi) - ref).max() < 1e-13


@pytest.mark.parametrize("t,beta", [(0.5, 0.0), (3.0, 0.0), (1.0, 0.4)])
def test_evolve_vs_oracle_and_ed(cpu, orc, t, beta):
    N, L = 16, 8
    bits, coeff = orc.syk_terms(N, 0.0, seed=21, fp=64)
    psi = random_state(np.random.default_rng(2), L)
    ref, _ = orc.krylov_evolve(L, bits, coeff, psi, t, beta, krydim=10, tol=1e-12)
    ex = ed_evolve(orc.dense_matrix(L, bits, coeff), psi, t, beta)
    got, mv = cpu.evolve(L, bits, coeff, psi, t, beta, krydim=10, tol=1e-12)
    assert mv > 0
    assert np.abs(got - ref).max() < 1e-11
    assert np.abs(got - ex).max() < 1e-11


def test_evolve_fp32_bulk_vs_fp64_oracle(cpu, orc):
    L = 12
    bits, coeff = orc.syk_terms(24, 8.0, seed=5, fp=32)
    psi = random_state(np.random.default_rng(4), L, np.complex64)
    ref, _ = orc.krylov_evolve(L, bits, coeff, psi.astype(np.complex128), 3.0, 0.0,
                               krydim=20, tol=1e-12)
    got, _ = cpu.evolve(L, bits, coeff, psi, 3.0, 0.0, krydim=20, tol=1e-7, fp=32)
    err = np.linalg.norm(got.astype(np.complex128) - ref) / np.linalg.norm(ref)
    assert err < 1e-5


def test_exact_diagonalization_evolve(cpu, orc):
    L = 7
    bits, coeff = orc.syk_terms(14, 0.0, seed=9, fp=64)
    psi = random_state(np.random.default_rng(8), L)
    ex = ed_evolve(orc.dense_matrix(L, bits, coeff), psi, 2.0, 0.3)
    assert np.abs(cpu.ed_evolve(L, bits, coeff, psi, 2.0, 0.3) - ex).max() < 1e-11


@pytest.mark.parametrize("N,sp", [(16, 0.0), (18, 0.0), (20, 4.0)])
def test_eigensystem_vs_numpy(cpu, orc, N, sp):
    """Hermitian eigensystem of the dense form, including parity-conserving
    SYK Hamiltonians whose matrices split into decoupled blocks with
    degenerate spectra (N=16, 20: two-fold degenerate levels)."""
    L = N // 2
    bits, coeff = orc.syk_terms(N, sp, seed=N, fp=64)
    H = orc.dense_matrix(L, bits, coeff)
    w, V = cpu.eigensystem(L, bits, coeff)
    assert np.all(np.diff(w) >= -1e-13)
    assert np.abs(w - np.linalg.eigvalsh(H)).max() < 1e-12
    assert np.abs(H @ V - V * w).max() < 1e-12
    assert np.abs(V.conj().T @ V - np.eye(1 << L)).max() < 1e-12


def test_eigensystem_random_complex_hermitian(cpu, orc):
    """A random (non-Hermitian) Pauli sum: the eigensystem is that of the
    Hermitian matrix formed from its dense form's lower triangle, as
    get_eigensystem documents; every entry complex, no degeneracies."""
    rng = np.random.default_rng(77)
    L = 6
    bits, coeff = random_pauli_sum(rng, L, 40)
    H0 = orc.dense_matrix(L, bits, coeff)
    Hl = np.tril(H0, -1) + np.tril(H0, -1).conj().T + np.diag(np.diag(H0).real)
    w, V = cpu.eigensystem(L, bits, coeff)
    assert np.abs(w - np.linalg.eigvalsh(Hl)).max() < 1e-12
    assert np.abs(Hl @ V - V * w).max() < 1e-12


def test_ground_state_vs_eigvalsh(cpu, orc):
    L = 8
    bits, coeff = orc.syk_terms(16, 0.0, seed=33, fp=64)
    w = np.linalg.eigvalsh(orc.dense_matrix(L, bits, coeff))
    psi = random_state(np.random.default_rng(1), L)
    e, v = cpu.ground_state(L, bits, coeff, psi, krydim=30, eps=1e-12)
    assert abs(e - w[0]) < 1e-9
    assert abs(np.linalg.norm(v) - 1) < 1e-12


def test_random_state_is_seeded_and_normalized(cpu):
    a = cpu.random_state(12, 4242)
    b = cpu.random_state(12, 4242)
    c = cpu.random_state(12, 4243)
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    assert abs(np.linalg.norm(a) - 1) < 1e-14
    # complex Gaussian entries: E|a_i|^2 = 1/dim, Re and Im uncorrelated
    x = a * np.sqrt(a.size)
    assert abs(np.mean(np.abs(x) ** 2) - 1) < 1e-12
    assert abs(np.mean(x.real * x.imag)) < 0.05


def _gxx():
    return shutil.which("g++")


@pytest.mark.skipif(not os.path.isdir(REF_EXAMPLES), reason="reference examples not present")
@pytest.mark.parametrize("prog", ["simple-syk-krylov.cc", "simple-syk-exdiag.cc"])
@pytest.mark.parametrize("nogpu", [False, True])
def test_reference_examples_compile_against_this_api(prog, nogpu):
    """The reference's short example programs, unchanged, compile against
    this build's headers with the GPU backend (the default) and with
    REAPERS_NOGPU (compile-only: they are read where they lie)."""
    cmd = [_gxx(), "-std=c++20", "-fsyntax-only", "-fopenmp"] + ["-I" + d for d in INC]
    if nogpu:
        cmd.append("-DREAPERS_NOGPU")
    cmd.append(os.path.join(REF_EXAMPLES, prog))
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]


@pytest.mark.parametrize("blocks", [False, True])
def test_driver_api_check_on_cpu_backend(tmp_path, blocks):
    """The promotion, backend-swap, dense-view and ED checks of
    examples/driver_api_check.cc, built with REAPERS_NOGPU and run."""
    exe = str(tmp_path / "driver_api_check")
    cmd = [_gxx(), "-std=c++20", "-O2", "-fopenmp", "-DREAPERS_NOGPU"] + ["-I" + d for d in INC]
    if blocks:
        cmd.append("-DREAPERS_USE_PARITY_BLOCKS")
    cmd += [os.path.join(ROOT, "reapers_amd", "examples", "driver_api_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe, "16"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout


@pytest.mark.parametrize("blocks", [False, True])
def test_driver_api_check_compiles_for_gpu_backend(blocks):
    cmd = [_gxx(), "-std=c++20", "-fsyntax-only", "-fopenmp"] + ["-I" + d for d in INC]
    if blocks:
        cmd.append("-DREAPERS_USE_PARITY_BLOCKS")
    cmd.append(os.path.join(ROOT, "reapers_amd", "examples", "driver_api_check.cc"))
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
