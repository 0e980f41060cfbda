"""The multiphase warm start end to end on the GPU at 32^3: drainages this solver runs to
their steady saturation as training data, the U-Net phase-field head trained by the
project's kernels, and held-out packs run to the saturation steady test three ways -- the
reference's start in FP64 (the arbiter), the same start in FP32 (the method without the
network) and the learned start in FP32 with the fallback rule."""
import sys

import numpy as np
import pytest

from mflbm import warmstart_mp as W

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-warmstart-mp] {msg}", flush=True)
    sys.stdout.flush()


def test_learned_phase_field_start_32():
    say("16 drainages of 32^3 to steady saturation, training")
    data = W.generate(range(16), n=32, log=say, tol=1e-4, check=1000)
    assert len(data) >= 14
    ws = W.PhaseWarmStart(seed=0)
    hist = ws.train(data, epochs=30, lr=1e-3, log=lambda m: None)
    assert hist[-1] < 0.25 * hist[0]
    rows = W.evaluate(ws, range(1000, 1002), n=32, tol=1e-4, check=1000, log=say)
    for r in rows:
        say(str({k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}))
        # the learned field is the steady one on most pore nodes
        assert r["predicted_phase_agreement"] > 0.8, r
        assert np.isfinite(r["S_warm"]) and np.isfinite(r["S_plain"])
        # at this size the learned start lands on the FP64 cold start's steady saturation
        # (measured 0.3 % and 0.003 %; a few percent is what another drainage path costs)
        assert abs(r["S_warm"] - r["S_cold"]) < 0.02 * r["S_cold"], r
    # what the network gains over the method without it (the reference's start in FP32):
    # fewer steps and less wall time to the steady test, over the held-out packs (measured
    # 50 000 against 99 000 steps)
    warm_steps = sum(r["warm_steps"] for r in rows)
    plain_steps = sum(r["plain_steps"] for r in rows)
    say(f"held-out steps: learned {warm_steps}, reference start {plain_steps}")
    assert warm_steps < plain_steps
    assert sum(r["warm_s"] for r in rows) < sum(r["plain_s"] for r in rows)
