"""Dense matrices and exact diagonalization on the GPU (the reference's
DevSumOps::get_matrix / get_eigensystem / matexp and ExactDiagonalization,
inc/gpuimpl.h:1179-1263 and inc/algo.h:28-54) against the oracle's dense
matrices and numpy/scipy, the Krylov path against ED at sizes the oracle
cannot reach quickly, and the trace-definition OTOC and Green's function of
the parity-block SYK driver (examples/syk.cc:171-185, 286-298)."""
import numpy as np
import pytest
import scipy.linalg

from conftest import random_pauli_sum, random_state

pytestmark = pytest.mark.gpu


def rounded(c, fp):
    return c.astype(np.complex64).astype(np.complex128) if fp == 32 else c


@pytest.mark.parametrize("fp", [32, 64])
def test_dense_matrix_matches_oracle(gpu, orc, fp):
    rng = np.random.default_rng(4)
    for L, T in [(1, 3), (4, 40), (7, 200)]:
        bits, coeff = random_pauli_sum(rng, L, T)
        coeff = rounded(coeff, fp)
        h = gpu.Hamiltonian(L, bits, coeff, fp=fp)
        want = orc.dense_matrix(L, bits, coeff)
        got = h.matrix()
        tol = 1e-6 * np.abs(coeff).sum() if fp == 32 else 1e-14 * np.abs(coeff).sum()
        assert np.abs(got - want).max() <= tol, (L, T)


@pytest.mark.parametrize("fp", [32, 64])
def test_eigensystem_and_matexp(gpu, orc, fp):
    N, L = 14, 6
    bits, coeff = orc.syk_block_terms(N, 0, seed=8, fp=fp)
    h = gpu.Hamiltonian(L, bits, coeff, fp=fp)
    H = orc.dense_matrix(L, bits, coeff)
    w, V = h.eigensystem()
    wref = np.linalg.eigvalsh(H)
    tol = 2e-6 if fp == 32 else 1e-13
    assert np.all(np.diff(w) >= 0)
    assert np.abs(w - wref).max() < tol
    assert np.abs(V.conj().T @ V - np.eye(1 << L)).max() < 10 * tol
    assert np.abs(H @ V - V * w).max() < 10 * tol
    for c in (-0.3 - 2.0j, 1.5j, -1.0):
        E = h.matexp(c)
        ref = scipy.linalg.expm(c * H)
        err = np.abs(E - ref).max() / max(1.0, np.abs(ref).max())
        print("fp%d matexp(%s): rel err %.2e" % (fp, c, err), flush=True)
        assert err < (2e-5 if fp == 32 else 1e-12)
    # cached by c: a repeated c adds no entry, a new one does
    n0 = h.matexp_cache_entries()
    h.matexp(1.5j)
    assert h.matexp_cache_entries() == n0
    h.matexp(0.25j)
    assert h.matexp_cache_entries() == n0 + 1


@pytest.mark.parametrize("N,sp", [(20, 0.0), (24, 0.0), (24, 8.0)])
def test_eigensystem_of_parity_conserving_matrix(gpu, orc, N, sp):
    """The non-block SYK Hamiltonian conserves parity, so in the computational
    basis its matrix splits into decoupled blocks (and for N = 20 every level
    is doubly degenerate); the eigenvectors must come back finite and solve
    the eigenproblem, and the exponential must match scipy's."""
    L = N // 2
    bits, coeff = orc.syk_terms(N, sp, seed=2000, fp=64)
    h = gpu.Hamiltonian(L, bits, coeff, fp=64)
    H = orc.dense_matrix(L, bits, coeff)
    w, V = h.eigensystem()
    assert np.all(np.isfinite(V))
    assert np.abs(w - np.linalg.eigvalsh(H)).max() < 1e-12
    assert np.abs(H @ V - V * w).max() < 1e-11
    E = h.matexp(-0.5j)
    assert np.abs(E - scipy.linalg.expm(-0.5j * H)).max() < 1e-11


@pytest.mark.parametrize("fp", [32, 64])
def test_ed_evolve_and_ground_state(gpu, orc, fp):
    N, L = 20, 9
    bits, coeff = orc.syk_block_terms(N, 1, seed=2, fp=fp)
    h = gpu.Hamiltonian(L, bits, coeff, fp=fp)
    H = orc.dense_matrix(L, bits, coeff)
    psi = random_state(np.random.default_rng(6), L)
    dt = np.complex64 if fp == 32 else np.complex128
    s = gpu.DeviceState(L, fp=fp)
    s.load(psi.astype(dt))
    t, beta = 3.0, 0.4
    s.evolve_ed(h, t, beta)
    ref = scipy.linalg.expm((-beta - 1j * t) * H) @ psi
    err = np.abs(s.store() - ref).max()
    print("fp%d ED evolve: max err %.2e" % (fp, err), flush=True)
    assert err < (3e-5 if fp == 32 else 1e-12)
    e = s.ground_state_ed(h)
    w = np.linalg.eigvalsh(H)
    assert e == pytest.approx(w[0], abs=2e-6 if fp == 32 else 1e-12)
    v = s.store().astype(np.complex128)
    assert np.linalg.norm(H @ v - e * v) < (1e-4 if fp == 32 else 1e-10)


def test_krylov_against_on_gpu_ed(gpu, orc):
    """The GPU's own ED as the oracle of the Krylov path at 12 sites (4096
    amplitudes, 1820 terms): FP64 Krylov to 1e-12 and FP32 to 1e-5."""
    N, L = 26, 12
    for fp, tol, bound in ((64, 1e-13, 1e-10), (32, 1e-7, 2e-5)):
        bits, coeff = orc.syk_block_terms(N, 0, sp=4.0, seed=31, fp=fp)
        h = gpu.Hamiltonian(L, bits, coeff, fp=fp)
        dt = np.complex64 if fp == 32 else np.complex128
        psi = random_state(np.random.default_rng(1), L).astype(dt)
        a, b = gpu.DeviceState(L, fp=fp), gpu.DeviceState(L, fp=fp)
        a.load(psi)
        b.load(psi)
        a.evolve_ed(h, 5.0, 0.2)
        st = b.evolve(h, 5.0, 0.2, krydim=20, tol=tol)
        err = np.abs(a.store() - b.store()).max()
        print("fp%d Krylov vs ED at L=12: %.2e (%d matvecs)" % (fp, err, st.matvecs), flush=True)
        assert err < bound


def dense_trace_npoint(kind, mLL, mRR, mop, tmax, nsteps, beta):
    E = scipy.linalg.expm(-beta * mLL)
    out = []
    for k in range(nsteps + 1):
        t = k * tmax / nsteps
        m0 = scipy.linalg.expm(-1j * t * mRR)
        m1 = scipy.linalg.expm(1j * t * mLL)
        if kind == "otoc-trace":
            out.append(np.trace(m1 @ m0 @ mop @ m1 @ m0 @ mop @ E))
        else:
            out.append(np.trace(m1 @ m0 @ E))
    out = np.array(out)
    return out / out[0]


@pytest.mark.parametrize("kind,beta,fp", [("otoc-trace", 0.0, 64), ("otoc-trace", 2.0, 64),
                                          ("green-trace", 1.0, 64), ("otoc-trace", 2.0, 32)])
def test_trace_npoint_vs_dense(gpu, orc, kind, beta, fp):
    N, seed, L = 12, 3, 5
    tmax, nsteps = 4.0, 8
    t, v, e0, _ = gpu.syk_block_npoint(N, kind, seed=seed, beta=beta, tmax=tmax, nsteps=nsteps,
                                       fp=fp)
    mLL = orc.dense_matrix(L, *orc.syk_block_terms(N, 0, seed=seed, fp=fp))
    mRR = orc.dense_matrix(L, *orc.syk_block_terms(N, 1, seed=seed, fp=fp))
    ob, oc = orc.majorana_block(N, N - 2, 0, fp=fp)
    mop = orc.dense_matrix(L, np.array([ob], np.uint64), np.array([oc]))
    ref = dense_trace_npoint(kind, mLL, mRR, mop, tmax, nsteps, beta)
    err = np.abs(v - ref).max()
    print("%s beta=%g fp%d: max err %.2e" % (kind, beta, fp, err), flush=True)
    assert v[0] == 1 and e0 == 0
    assert err < (5e-5 if fp == 32 else 1e-11)


def test_trace_and_state_definitions_agree_statistically(gpu):
    """At beta = 0 the state definition with a random state estimates the
    trace definition (typicality): 12 sites (N = 26), error ~ 2^-6."""
    N = 26
    _, vt, _, _ = gpu.syk_block_npoint(N, "green-trace", seed=7, tmax=3.0, nsteps=3, fp=64)
    _, vs, _, _ = gpu.syk_block_npoint(N, "green", seed=7, state_seed=5, tmax=3.0, nsteps=3,
                                       krylov_dim=20, tol=1e-12, fp=64)
    print("green trace %s\ngreen state %s" % (vt, vs), flush=True)
    assert np.abs(vt - vs).max() < 0.1


def test_matexp_cache_budget_evicts(gpu, orc):
    """With a budget of three 2^8 x 2^8 complex<double> matrices, a fourth
    matexp drops the least-used entry; results stay correct."""
    L = 8
    bits, coeff = orc.syk_block_terms(18, 0, seed=3, fp=64)
    h = gpu.Hamiltonian(L, bits, coeff, fp=64)
    H = orc.dense_matrix(L, bits, coeff)
    old_max, _ = gpu.matexp_cache_budget()
    one = (1 << L) ** 2 * 16
    try:
        _, used0 = gpu.matexp_cache_budget()
        gpu.matexp_cache_budget(used0 + 3 * one)
        for k, t in enumerate([0.1, 0.2, 0.3]):
            h.matexp(-1j * t)
        h.matexp(-0.1j)          # one more use of t = 0.1
        assert h.matexp_cache_entries() == 3
        E = h.matexp(-0.4j)      # evicts the least used (t = 0.2 or 0.3)
        assert h.matexp_cache_entries() == 3
        mx, used = gpu.matexp_cache_budget()
        assert used <= mx
        assert np.abs(E - scipy.linalg.expm(-0.4j * H)).max() < 1e-12
    finally:
        gpu.matexp_cache_budget(old_max)


@pytest.mark.parametrize("fp", [32, 64])
def test_state_element_write(gpu, fp):
    L = 6
    s = gpu.DeviceState(L, fp=fp)
    s.random(3)
    v = s.store()
    s[5] = 0.25 - 0.5j
    s[63] = 2.0
    w = s.store()
    v[5], v[63] = 0.25 - 0.5j, 2.0
    assert np.array_equal(w, v.astype(w.dtype))


@pytest.mark.parametrize("fp", [32, 64])
def test_dense_apply_switch(gpu, orc, fp):
    """With the switch on, a Hamiltonian with at least 2^len terms is applied
    through its dense matrix (dense SYK N=24: 10626 strings on 4096
    amplitudes); the result and a Krylov evolution match the oracle."""
    import time
    L = 12
    bits, coeff = orc.syk_terms(24, 0.0, seed=6, fp=fp)
    psi = random_state(np.random.default_rng(4), L)
    dt = np.complex64 if fp == 32 else np.complex128
    ref = orc.apply(L, bits, coeff, psi, fp=64)
    times = {}
    prev = gpu.set_dense_apply(False)
    try:
        for on in (False, True):
            gpu.set_dense_apply(on)
            h = gpu.Hamiltonian(L, bits, coeff, fp=fp)
            s = gpu.DeviceState(L, fp=fp)
            s.load(psi.astype(dt))
            s.apply(h)
            got = s.store().astype(np.complex128)
            err = np.abs(got - ref).max() / np.abs(ref).max()
            assert err < (2e-6 if fp == 32 else 1e-13), (on, err)
            s.load(psi.astype(dt))
            t0 = time.perf_counter()
            st = s.evolve(h, 1.0, krydim=20, tol=1e-7 if fp == 32 else 1e-12)
            times[on] = time.perf_counter() - t0
            ev, _ = orc.krylov_evolve(L, bits, coeff, psi, 1.0, 0.0, krydim=20, tol=1e-14)
            assert np.linalg.norm(s.store().astype(np.complex128) - ev) < (2e-5 if fp == 32 else 1e-10)
        print("fp%d evolve: term-split %.2f ms, dense %.2f ms" % (fp, 1e3 * times[False],
                                                                1e3 * times[True]), flush=True)
    finally:
        gpu.set_dense_apply(prev)
