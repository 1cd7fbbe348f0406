"""The tiled apply's output-tile map and group order (reapers_amd/csrc/rp_tables.h),
checked on the host: a small program builds the complex<float> tables with the
library's own header and reports on them (tests/tile_map_check.cc).

Properties: the map from a workgroup's index (position within its XCD, XCD,
batch) to its output tile is a bijection of the grid, so every output tile is
written exactly once; every group's terms stay contiguous (each high flip part
xh appears once); and the runs of equal XCD class along the group list (each
run is one coset of input tiles fetched into an XCD's L2) are fewer than along
the contiguous order's sorted xh.  For sparse SYK at 30 qubits (2^13 tiles,
one 1024-thread workgroup per CU, 32 per XCD) the runs fall from 117 to 102,
for 24 qubits (2^12 tiles, 64 per XCD) from 20 to 17."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from conftest import random_pauli_sum

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = shutil.which("hipcc") or ("/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc")
                                  else None)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if HIPCC is None:
        pytest.skip("hipcc not found")
    exe = str(tmp_path_factory.mktemp("tilemap") / "tile_map_check")
    cmd = [HIPCC, "-std=c++17", "-O1", "-I" + os.path.join(ROOT, "include"),
           "-I" + os.path.join(ROOT, "reapers_amd", "csrc"),
           os.path.join(ROOT, "tests", "tile_map_check.cc"), "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def run(checker, tmp_path, bits, length, tb, nt, p):
    f = tmp_path / ("bits_%d_%d_%d.bin" % (length, tb, p))
    np.ascontiguousarray(bits, np.uint64).tofile(f)
    r = subprocess.run([checker, str(f), str(length), str(tb), str(nt), str(p)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    import json
    return json.loads(r.stdout)


@pytest.mark.parametrize("N,seed,tb,nt,p,runs,runs_contig", [
    (60, 4000, 13, 1024, 5, 102, 117),
    (48, 2000, 12, 512, 6, 17, 20),
])
def test_sparse_syk_map(checker, orc, tmp_path, N, seed, tb, nt, p, runs, runs_contig):
    bits, _ = orc.syk_terms(N, 8.0, seed=seed, fp=32)
    out = run(checker, tmp_path, bits, N // 2, tb, nt, p)
    assert out["n"] == N // 2 - tb
    assert out["bijective"] and out["contiguous"]
    assert out["runs_contiguous_order"] == runs_contig
    assert out["runs"] == runs


@pytest.mark.parametrize("length,tb,nt,p", [
    (21, 12, 512, 6), (24, 12, 512, 6), (26, 13, 1024, 5), (24, 10, 128, 8), (22, 8, 128, 8),
    (20, 12, 512, 6),
])
def test_random_sums(checker, tmp_path, length, tb, nt, p):
    rng = np.random.default_rng(length * 100 + tb)
    for nterms, max_sites in ((40, None), (300, 6), (900, 10)):
        bits, _ = random_pauli_sum(rng, length, nterms, max_sites=max_sites)
        out = run(checker, tmp_path, bits, length, tb, nt, p)
        assert out["contiguous"]
        if length - tb < p + 3:
            assert out["n"] == 0            # too few tile bits for a map
        if out["n"]:
            assert out["bijective"]
            assert out["runs"] < out["runs_contiguous_order"]
        else:
            assert out["runs"] == out["runs_contiguous_order"]
