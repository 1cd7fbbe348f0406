"""The five benchmark workloads at their own sizes, through the C ABI.

1. Dense SYK N=24 (12 qubits, C(24,4) = 10626 strings), FP64, Krylov
   dimension 30 at tol 1e-12, t in {0.5, 1, 2, 5}: against exact
   diagonalization (numpy eigh of the oracle's dense matrix) and the oracle's
   restatement of Krylov::evolve (inc/algo.h:266-365), within 1e-11.
2. Sparse SYK N=48, k=8 (24 qubits): H psi of the whole 2^24-amplitude state
   against the oracle's apply (inc/cpuimpl.h:737-754) for FP64 and for the
   benchmark's FP32 upload (autotuned tiling, group accumulation), on the
   same FP32 state and FP32-rounded couplings; FP64 within 1e-11 and FP32
   within 4e-7 of the largest amplitude.
3. The batch: 64 realizations of sparse SYK k=8 with N drawn from
   {32, 36, 40, 44}, 4 states each (256 right-hand sides), t=5, tol 1e-6,
   FP32, autotuned batched apply, graph-replayed restarts; a sample of
   right-hand sides against the oracle's single-state FP64 evolution (N=32,
   36) and one per N against the single-state FP64 path, within 1e-5
   relative.
4. OTOC at N=60 (30 qubits, 8.6 GB FP32 state): F(0) = -1/4 at beta = 0
   (chi_1 chi_2 chi_1 chi_2 = -chi_1^2 chi_2^2 = -1/4), and at one point of
   the workload's time grid (t = 0.1 * 200^(5/19) = 0.403) the library's
   OTOC driver with the proposer agrees with the four evolutions done one by
   one at the fixed dimension 12 (1e-5), |F| <= 1/4 and the final state has
   norm 1/4 (each chi_n is a unitary over sqrt(2)).
5. The trained proposer on unseen XXZ L=26 and SYK N=48 at tol 1e-6 (FP32
   states): error against an FP64 evolution at tol 1e-14 within 1e-5, and
   fewer applies, fewer state-vector passes and less wall time than the fixed
   Krylov dimension 20.
"""
import ctypes

import numpy as np
import pytest

from conftest import random_state
from gpu_util import DeviceBuffer, Terms

pytestmark = pytest.mark.gpu


def _banner(msg):
    print(msg, flush=True)


# ---------------------------------------------------------------------------
# configuration 1
# ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def dense_syk24(orc):
    L = 12
    bits, coeff = orc.syk_terms(24, 0.0, seed=1000, fp=64)
    H = orc.dense_matrix(L, bits, coeff)
    w, Q = np.linalg.eigh(H)
    psi = random_state(np.random.default_rng(1001), L)
    return L, bits, coeff, w, Q, psi


@pytest.mark.parametrize("t", [0.5, 1.0, 2.0, 5.0])
def test_config1_dense_syk24_dim30_vs_ed_and_oracle(gpu, orc, dense_syk24, t):
    L, bits, coeff, w, Q, psi = dense_syk24
    assert len(bits) == 10626
    _banner("config 1: dense SYK N=24, t=%g" % t)
    ex = Q @ (np.exp(-1j * t * w) * (Q.conj().T @ psi))
    ref, st_ref = orc.krylov_evolve(L, bits, coeff, psi, t, 0.0, krydim=30, tol=1e-12)
    ham = gpu.Hamiltonian(L, bits, coeff, fp=64)
    s = gpu.DeviceState(L, fp=64)
    s.load(psi)
    st = s.evolve(ham, t, 0.0, krydim=30, tol=1e-12)
    got = s.store()
    s.close()
    ham.close()
    e_ed, e_or = np.abs(got - ex).max(), np.abs(got - ref).max()
    print("t=%g: |gpu-ed| %.2e |gpu-oracle| %.2e |oracle-ed| %.2e, matvecs %d (oracle %d)"
          % (t, e_ed, e_or, np.abs(ref - ex).max(), st.matvecs, st_ref.matvecs), flush=True)
    assert e_ed < 1e-11
    assert e_or < 1e-11


# ---------------------------------------------------------------------------
# configuration 2
# ---------------------------------------------------------------------------
def test_config2_whole_state_apply_n48_vs_oracle(gpu, orc):
    N, L = 48, 24
    dim = 1 << L
    _banner("config 2: whole-state H psi at N=48")
    bits, coeff = orc.syk_terms(N, 8.0, seed=2000, fp=32)
    coeff = coeff.astype(np.complex64).astype(np.complex128)
    hip = gpu.hip()
    vin = DeviceBuffer(8 * dim)
    gpu.check(hip.rp_init_random(gpu.RP_C64, vin.ptr, dim, 2001))
    psi32 = vin.to_array(np.complex64, dim)
    psi64 = psi32.astype(np.complex128)
    ref = orc.apply(L, bits, coeff, psi64)
    scale = np.abs(ref).max()
    # the benchmark's FP32 upload
    t32 = Terms(L, bits, coeff, gpu.RP_C64, flags=gpu.RP_TERMS_AUTOTUNE | gpu.RP_TERMS_ACC_GROUP)
    info = t32.info()
    vout = DeviceBuffer(8 * dim)
    gpu.check(hip.rp_apply_ops(t32.h, vout.ptr, vin.ptr, 0))
    got32 = vout.to_array(np.complex64, dim).astype(np.complex128)
    t32.free()
    vout.free()
    vin.free()
    e32 = np.abs(got32 - ref).max() / scale
    # FP64 on the same (promoted) state
    t64 = Terms(L, bits, coeff, gpu.RP_C128)
    vin = DeviceBuffer.from_array(psi64)
    vout = DeviceBuffer(16 * dim)
    gpu.check(hip.rp_apply_ops(t64.h, vout.ptr, vin.ptr, 0))
    got64 = vout.to_array(np.complex128, dim)
    t64.free()
    vout.free()
    vin.free()
    e64 = np.abs(got64 - ref).max() / scale
    print("N=48 whole state (%d terms, tiling %d/%d/%d): FP32 group %.2e, FP64 %.2e of max |H psi|"
          % (len(bits), info.tile_bits, info.threads, info.depth, e32, e64), flush=True)
    assert e64 < 1e-11
    assert e32 < 4e-7


# ---------------------------------------------------------------------------
# configuration 3
# ---------------------------------------------------------------------------
def test_config3_batch_64x4_vs_oracle(gpu, orc):
    nreal, nrhs, t, tol, m, kmax = 64, 4, 5.0, 1e-6, 12, 24
    rng = np.random.default_rng(3)
    Ns = rng.choice([32, 36, 40, 44], size=nreal)
    lens = [int(n) // 2 for n in Ns]
    _banner("config 3: %d problems x %d states, N counts %s"
            % (nreal, nrhs, {int(k): int((Ns == k).sum()) for k in (32, 36, 40, 44)}))
    terms = [gpu.syk_terms(int(n), 8.0, seed=3000 + r, fp=32) for r, n in enumerate(Ns)]
    states = []
    for L in lens:
        x = rng.normal(size=(1 << L, nrhs)) + 1j * rng.normal(size=(1 << L, nrhs))
        states.append((x / np.linalg.norm(x, axis=0)).astype(np.complex64))
    b = gpu.Batch(lens, terms, nrhs=nrhs, kmax=kmax, fp=32, autotune=True)
    b.load(states)
    st, rs = b.evolve(t, m=m, tol=tol, graph=True)
    got = b.store()
    b.close()
    print("batch: %d matvecs per rhs, restarts %d..%d, %.0f ms"
          % (st.matvecs, rs.min(), rs.max(), st.wall_ms), flush=True)
    # the oracle's FP64 single-state evolution: two states of an N=32
    # problem and one of an N=36 problem
    picks = []
    for n, cnt in ((32, 2), (36, 1)):
        p = int(np.flatnonzero(Ns == n)[0])
        picks += [(p, q) for q in range(cnt)]
    worst = 0.0
    for p, q in picks:
        L = lens[p]
        bits, coeff = terms[p]
        ref, _ = orc.krylov_evolve(L, bits, coeff, states[p][:, q].astype(np.complex128), t, 0.0,
                                   krydim=20, tol=1e-12)
        err = np.linalg.norm(got[p][:, q] - ref) / np.linalg.norm(ref)
        print("  oracle: problem %d (N=%d) state %d: %.2e" % (p, 2 * L, q, err), flush=True)
        worst = max(worst, err)
    # one state per N against the single-state FP64 path at tol 1e-13
    for n in (32, 36, 40, 44):
        p = int(np.flatnonzero(Ns == n)[-1])
        L = lens[p]
        H = gpu.Hamiltonian(L, *terms[p], fp=64)
        s = gpu.DeviceState(L, fp=64)
        s.load(states[p][:, 3].astype(np.complex128))
        s.evolve(H, t, 0.0, krydim=30, tol=1e-13)
        ref = s.store()
        s.close()
        H.close()
        err = np.linalg.norm(got[p][:, 3] - ref) / np.linalg.norm(ref)
        print("  fp64 path: problem %d (N=%d) state 3: %.2e" % (p, n, err), flush=True)
        worst = max(worst, err)
    assert worst < 1e-5


# ---------------------------------------------------------------------------
# configuration 4
# ---------------------------------------------------------------------------
def test_config4_otoc_n60_point(gpu):
    from reapers_amd import proposer
    N, L = 60, 30
    _banner("config 4: N=60 OTOC")
    bits, coeff = gpu.syk_terms(N, 8.0, seed=4000, fp=32)
    gpu.set_upload_flags(gpu.RP_TERMS_ACC_GROUP)
    try:
        H = gpu.Hamiltonian(L, bits, coeff, fp=32)
        ops = []
        for idx in (1, 2):
            mb, mc = gpu.majorana(N, idx, fp=32)
            ops.append(gpu.Hamiltonian(L, np.array([mb], np.uint64), np.array([mc]), fp=32))
    finally:
        gpu.set_upload_flags(0)
    W, V = ops
    init = gpu.DeviceState(L, fp=32)
    init.random(4001)
    F0, _ = init.otoc(H, W, V, [0.0], krydim=12, tol=1e-7)
    print("F(0) = %.9f%+.2ei" % (F0[0].real, F0[0].imag), flush=True)
    assert abs(F0[0] + 0.25) < 1e-6
    t = float(np.geomspace(0.1, 20.0, 20)[5])
    sched = gpu.Schedule.from_mlp(proposer.load_weights(), max_dim=16)
    Fp, stp = init.otoc(H, W, V, [t], krydim=12, tol=1e-7, schedule=sched)
    sched.close()
    print("t=%.4f proposer: F = %.8f%+.2ei, %d matvecs" % (t, Fp[0].real, Fp[0].imag,
                                                        stp.matvecs), flush=True)
    # the same point as four evolutions one by one, fixed dimension 12
    psi = gpu.DeviceState(L, fp=32)
    psi.copy_from(init)
    for op, tk in ((V, t), (W, -t), (V, t), (W, -t)):
        psi.apply(op)
        psi.evolve(H, tk, 0.0, krydim=12, tol=1e-7)
    Fm = init.inner(psi)
    nrm = psi.norm()
    psi.close()
    init.close()
    for h in (H, W, V):
        h.close()
    print("t=%.4f fixed 12: F = %.8f%+.2ei, final norm %.8f" % (t, Fm.real, Fm.imag, nrm),
          flush=True)
    assert abs(Fm - Fp[0]) < 1e-5
    assert abs(Fp[0]) <= 0.25 + 1e-6
    assert abs(nrm - 0.25) < 1e-5


# ---------------------------------------------------------------------------
# configuration 5
# ---------------------------------------------------------------------------
def test_config5_proposer_unseen_xxz26_syk48(gpu):
    from reapers_amd import proposer
    _banner("config 5: proposer on XXZ L=26 and SYK N=48")
    rng = np.random.default_rng(5155)
    insts = [proposer.make_instance(rng, "xxz", 26, fp=32),
             proposer.make_instance(rng, "syk", 24, fp=32)]
    for d, *_ in insts:
        d["tol"] = 1e-6
    rows = proposer.evaluate(proposer.load_weights(), insts, fixed=20, fp=32,
                             log=lambda m: print(m, flush=True))
    tot = {k: sum(r[k] for r in rows) for k in ("matvecs_fixed", "matvecs_mlp", "cost_fixed",
                                                 "cost_mlp", "wall_ms_fixed", "wall_ms_mlp")}
    print("totals: %s" % tot, flush=True)
    for r in rows:
        assert r["err_mlp"] < 1e-5, r
        assert r["err_fixed"] < 1e-5, r
    assert tot["matvecs_mlp"] < tot["matvecs_fixed"]
    assert tot["cost_mlp"] < tot["cost_fixed"]
    assert tot["wall_ms_mlp"] < tot["wall_ms_fixed"]
