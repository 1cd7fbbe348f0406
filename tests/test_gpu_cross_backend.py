"""The two backends side by side (inc/state.h:114-118, inc/gpuimpl.h:
438-462): a State<F, CPUImpl> converted to a State<F, GPUImpl> and back is
bit-identical; both evolve the same state by exp(-(beta + i t) H) with their
own Krylov<Impl> to the same result; the same seed gives the same random
state on both; and the driver_api_check example built for the GPU backend
passes its checks.

Tolerances: FP64 evolutions within 1e-11 of each other (max-abs, unit-norm
states; each within 1e-12 of exact results in its own tests); FP32 within
1e-5; FP64 random states within 1e-13 (the Box-Muller transform rounds
differently on the two sides)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import random_state

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("fp,N,sp,t,beta", [(64, 20, 0.0, 2.0, 0.0), (64, 28, 8.0, 5.0, 0.3),
                                          (32, 28, 8.0, 3.0, 0.0)])
def test_cpu_and_gpu_states_agree_after_evolution(gpu, orc, fp, N, sp, t, beta):
    L = N // 2
    dt = np.complex64 if fp == 32 else np.complex128
    bits, coeff = orc.syk_terms(N, sp, seed=N, fp=fp)
    if fp == 32:
        coeff = coeff.astype(np.complex64).astype(np.complex128)
    psi = random_state(np.random.default_rng(N), L, dt)
    tol = 1e-7 if fp == 32 else 1e-13
    c, g, back = gpu.cross_backend_evolve(L, bits, coeff, psi, t, beta, krydim=20, tol=tol, fp=fp)
    assert np.array_equal(back, psi)
    d = np.abs(c.astype(np.complex128) - g.astype(np.complex128)).max()
    ref, _ = orc.krylov_evolve(L, bits, coeff, psi.astype(np.complex128), t, beta, krydim=20,
                               tol=1e-13)
    e_c = np.abs(c.astype(np.complex128) - ref).max()
    print("fp%d N=%d t=%g beta=%g: |cpu-gpu| %.2e, |cpu-oracle| %.2e" % (fp, N, t, beta, d, e_c),
          flush=True)
    assert d < (1e-5 if fp == 32 else 1e-11)
    assert e_c < (1e-5 if fp == 32 else 1e-11)


def test_same_seed_same_random_state(gpu):
    from reapers_amd import cpu
    L = 16
    st = gpu.DeviceState(L, fp=64)
    st.random(987654321)
    g = st.store()
    st.close()
    c = cpu.random_state(L, 987654321, fp=64)
    d = np.abs(g - c).max()
    print("max |gpu - cpu| random state: %.2e" % d, flush=True)
    assert d < 1e-13


def test_driver_api_check_gpu_build():
    exe = os.path.join(ROOT, "reapers_amd", "examples", "driver_api_check")
    assert os.path.exists(exe), "examples are built by __graft_entry__.build()"
    r = subprocess.run([exe, "16"], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
