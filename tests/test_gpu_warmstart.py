"""The learned warm start end to end on the GPU: training fields from the batched ensemble,
the U-Net trained by the project's kernels, and held-out geometries solved three ways --
the reference's cold start in FP64 (the arbiter), the relaxed start from rest in FP32 (the
method without the network) and the learned start in FP32."""
import sys

import numpy as np
import pytest

from mflbm import warmstart

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-warmstart] {msg}", flush=True)
    sys.stdout.flush()


def test_learned_warm_start_32():
    say("ensemble data (48 samples of 32^3) and training")
    data = warmstart.generate(range(48), n=32, log=say)
    ws = warmstart.WarmStart(seed=0)
    hist = ws.train(data, epochs=20, lr=1e-3, log=lambda m: None)
    assert hist[-1] < 0.5 * hist[0]
    rows = warmstart.evaluate(ws, range(1000, 1003), n=32, log=say)
    for r in rows:
        # the target: FP32 + warm start within 1 % of the FP64 cold start's permeability
        assert r["k_warm_rel_diff"] < 1e-2 and r["k_relaxed_rel_diff"] < 1e-2, r
        # what the network adds: initial fields closer to the steady state than rest
        assert r["velocity_rel_err"] < 1.0, r
    cold = sum(r["cold_s"] for r in rows)
    warm = sum(r["warm_s"] for r in rows)
    relaxed = sum(r["relaxed_s"] for r in rows)
    steps = {k: sum(r[f"{k}_steps"] for r in rows) for k in ("cold", "relaxed", "warm")}
    say(f"steps cold {steps['cold']} relaxed {steps['relaxed']} warm {steps['warm']}; "
        f"wall cold {cold:.2f} s relaxed {relaxed:.2f} s warm {warm:.2f} s")
    # at most a third of the cold start's wall time (measured: about a tenth)
    assert warm < cold / 3
    # on these geometries the network's fields do not slow the relaxed start down by more
    # than one check interval per case on average
    assert steps["warm"] <= steps["relaxed"] + 100 * len(rows)


def test_fine_grid_starts_reach_the_cold_permeability():
    """The fine-grid path (network on the geometry coarsened by two, the fallback rule, the
    ensemble-kernel starts and the coarse-solve start) on a 32^3 refinement of a 16^3 pack:
    every start reaches the FP64 cold start's permeability."""
    ws = warmstart.WarmStart(seed=0)  # untrained: small fields, exercised for the plumbing
    rows = warmstart.evaluate_fine(ws, [1000], n_coarse=16, log=say, check=100, tol=1e-5, cold_max_steps=60000,
                                   starts=("cold", "relaxed", "warm", "relaxed_ens", "warm_ens", "coarse_ens"))
    r = rows[0]
    say(str({k: v for k, v in r.items() if k.endswith("steps") or "fallback" in k}))
    for name in ("relaxed", "warm", "relaxed_ens", "warm_ens", "coarse_ens"):
        assert r[f"{name}_steps"] > 0
        assert r[f"k_{name}_rel_diff"] < 1e-3, (name, r[f"k_{name}_rel_diff"])
    assert isinstance(r["fallback"], bool) and isinstance(r["coarse_fallback"], bool)
