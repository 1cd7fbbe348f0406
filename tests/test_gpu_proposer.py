"""The trained Krylov-schedule proposer (reapers_amd/proposer.f32, evaluated
by the HIP inference kernel inside the C++ evolution) against the fixed
schedule (one Krylov dimension for every restart, the reference's
Krylov::evolve) on instances of sizes the training set does not contain
(odd chain lengths; the training set has even ones), at the same tolerance:
it must reach the requested accuracy with less HBM traffic (the quantity its
training labels minimize: applies weighted by the apply's tile groups plus
the Gram-Schmidt and update sweeps) and no more matvecs."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def instances(rng, kinds):
    from reapers_amd import proposer
    return [proposer.make_instance(rng, kind, L) for kind, L in kinds]


@pytest.mark.parametrize("fp", [64, 32])
def test_proposer_gain_on_unseen_instances(gpu, fp):
    from reapers_amd import proposer
    params = proposer.load_weights()
    rng = np.random.default_rng(424242)
    kinds = [("xxz", 15), ("syk", 15), ("xxz", 17), ("syk", 17), ("xxz", 19), ("syk", 19)]
    insts = instances(rng, kinds)
    if fp == 32:
        # FP32 states: tolerances the FP32 rounding floor allows
        for d, *_ in insts:
            d["tol"] = float(10 ** rng.uniform(-7, -6))
    rows = proposer.evaluate(params, insts, fixed=20, fp=fp,
                             log=lambda m: print(m, flush=True))
    for r in rows:
        bound = 10 * r["tol"] if fp == 64 else max(10 * r["tol"], 1e-5)
        assert r["err_fixed"] < bound
        assert r["err_mlp"] < bound, r
    mv = (sum(r["matvecs_fixed"] for r in rows), sum(r["matvecs_mlp"] for r in rows))
    cost = (sum(r["cost_fixed"] for r in rows), sum(r["cost_mlp"] for r in rows))
    wall = (sum(r["wall_ms_fixed"] for r in rows), sum(r["wall_ms_mlp"] for r in rows))
    print("fp%d totals: matvecs fixed %d / proposer %d; HBM passes %.0f / %.0f; "
          "wall %.1f / %.1f ms" % (fp, mv[0], mv[1], cost[0], cost[1], wall[0], wall[1]),
          flush=True)
    assert cost[1] < 0.85 * cost[0]
    assert mv[1] <= mv[0]
