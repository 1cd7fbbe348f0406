"""The example program (reapers_amd/examples/syk_block_otoc.cc), a user of
the C++ API written like the reference's production driver
(examples/syk.cc with REAPERS_USE_PARITY_BLOCKS): it must build against
these headers, fail loudly without a GPU, and on the GPU print the same
normalized OTOC as the library's own driver."""
import os
import subprocess

import numpy as np
import pytest

EXE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "reapers_amd",
                   "examples", "syk_block_otoc")


def run(*args):
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=300)
    return r


def parse(out):
    rows = np.array([[float(x) for x in line.split()] for line in out.strip().splitlines()])
    return rows[:, 0], rows[:, 1] + 1j * rows[:, 2]


def test_example_usage(ra):
    assert os.path.exists(EXE)
    r = run()
    assert r.returncode == 2 and "usage" in r.stderr


@pytest.mark.gpu
def test_example_matches_library_driver(gpu):
    N, seed, beta, tmax, nsteps = 14, 12, 2.0, 5.0, 5
    r = run(N, 0, seed, beta, tmax, nsteps, "--fp64")
    assert r.returncode == 0, r.stderr
    t, v = parse(r.stdout)
    tl, vl, _, _ = gpu.syk_block_npoint(N, "otoc", seed=seed, state_seed=seed + 1, beta=beta,
                                        tmax=tmax, nsteps=nsteps, krylov_dim=12, tol=1e-13, fp=64)
    print("example vs driver: %.2e" % np.abs(v - vl).max(), flush=True)
    assert np.allclose(t, tl)
    # six printed digits, the example's Krylov tolerance 1e-7
    assert np.abs(v - vl).max() < 2e-5
    r = run(N, 0, seed, beta, tmax, nsteps, "--trace", "--fp64")
    assert r.returncode == 0, r.stderr
    _, vt = parse(r.stdout)
    _, vlt, _, _ = gpu.syk_block_npoint(N, "otoc-trace", seed=seed, beta=beta, tmax=tmax,
                                        nsteps=nsteps, fp=64)
    assert np.abs(vt - vlt).max() < 2e-5
    # FP32 (its own couplings: the generator's normal draws depend on the
    # precision)
    r = run(N, 0, seed, beta, tmax, nsteps)
    assert r.returncode == 0, r.stderr
    _, v32 = parse(r.stdout)
    _, vl32, _, _ = gpu.syk_block_npoint(N, "otoc", seed=seed, state_seed=seed + 1, beta=beta,
                                         tmax=tmax, nsteps=nsteps, krylov_dim=12, tol=1e-7, fp=32)
    assert np.abs(v32 - vl32).max() < 1e-4
