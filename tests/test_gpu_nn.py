"""GPU checks of the warm-start network kernels (csrc/nn.hip) against the FP64 numpy oracle.

Inputs and weights are rounded to bf16 first, so the oracle sees exactly the operand values
the MFMA units multiply; the device accumulates in fp32, which bounds the forward and
per-kernel gradient errors at ~1e-6 relative.  Through the whole U-Net the bf16 storage of
activations and gradients between layers dominates: there the bound is a relative L2 error
of 5e-2 per layer gradient (cosine > 0.995).
"""
import sys

import numpy as np
import pytest

import nn_oracle as N

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-nn] {msg}", flush=True)
    sys.stdout.flush()


def bf16_round(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16  # round to nearest even
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.fixture(scope="module")
def nn():
    from mflbm import nn as mod
    mod.lib()
    return mod


def rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


@pytest.mark.parametrize("cin,cout,shape", [
    (8, 16, (6, 5, 7)), (16, 8, (8, 8, 8)), (24, 40, (4, 6, 8)), (48, 192, (4, 4, 4)),
    # the large-grid launches: 65536 voxels (two voxel tiles per wave with two channel tiles,
    # weight gradient in chunks of 2048), 18432 voxels (three channel tiles over the voxel
    # grid), 16384 voxels with 144 channels (twelve tiles, 192 weight rows)
    (8, 32, (32, 32, 64)), (16, 48, (24, 24, 32)), (16, 144, (16, 32, 32))])
def test_conv_forward_dgrad_wgrad(nn, cin, cout, shape):
    say(f"conv3 cin={cin} cout={cout} shape={shape}")
    rng = np.random.default_rng(cin * 100 + cout)
    x = bf16_round(rng.normal(size=shape + (cin,)))
    conv = nn.Conv(cin, cout, rng)
    w = bf16_round(rng.normal(0, 0.2, (cout, 27 * cin)))
    b = rng.normal(0, 0.1, cout).astype(np.float32).astype(np.float64)
    conv.set_numpy(w, b)
    xd = nn.DevArray.from_numpy(x, "bf16")
    yf = nn.DevArray(shape + (cout,), "f32")
    conv.forward(xd, shape, relu=False, yf=yf)
    ref = N.conv3(x, w, b)
    assert rel(yf.numpy(), ref) < 1e-5
    yb = nn.DevArray(shape + (cout,), "bf16")
    conv.forward(xd, shape, relu=True, yb=yb)
    assert rel(yb.numpy(), np.maximum(ref, 0)) < 1e-2

    dz = bf16_round(rng.normal(size=shape + (cout,)))
    dzd = nn.DevArray.from_numpy(dz, "bf16")
    work = nn.DevArray((nn.lib().mfnn_wgrad_workspace(*shape, conv.kp, cout, nn.wgrad_chunk(*shape)),), "f32")
    dx = nn.DevArray(shape + (cin,), "bf16")
    conv.backward(xd, dzd, shape, work, accumulate=False, dx=dx)
    dw_ref, db_ref = N.conv3_wgrad(x, dz)
    dw_all = conv.dw.numpy()
    assert rel(dw_all[:cout, :27 * cin], dw_ref) < 1e-5
    assert rel(conv.db.numpy(), db_ref) < 1e-5
    assert np.all(dw_all[cout:, :] == 0) and np.all(dw_all[:, 27 * cin:] == 0)
    assert rel(dx.numpy(), N.conv3_dgrad(dz, w, cin)) < 1e-2
    # accumulate adds
    conv.backward(xd, dzd, shape, work, accumulate=True)
    assert rel(conv.dw.numpy()[:cout, :27 * cin], 2 * dw_ref) < 1e-5


def test_unet_gradients_match_oracle(nn):
    """Forward: every layer against the oracle applied to the same (device) input.  Backward:
    the oracle's backward pass run on the device's activations and output gradient, so that
    both sides use the same ReLU gates (a near-zero pre-activation can land on either side
    of zero between an fp32 and an fp64 forward pass, and a flipped gate changes a whole
    gradient entry)."""
    say("unet widths (8,16,16,24) 16^3 gradient check")
    net = nn.UNet(widths=(8, 16, 16, 24), cin=8, cout=8, seed=3)
    rng = np.random.default_rng(4)
    p = {}
    for k, c in net.conv.items():
        w = bf16_round(rng.normal(0, np.sqrt(2.0 / (27 * c.cin)), (c.cout, 27 * c.cin)))
        b = rng.normal(0, 0.05, c.cout).astype(np.float32).astype(np.float64)
        c.set_numpy(w, b)
        p[k] = (w, b)
    shape = (16, 16, 16)
    x = bf16_round(rng.normal(size=shape + (8,)))
    t = rng.normal(size=shape + (8,)).astype(np.float32).astype(np.float64)
    mask = (rng.random(shape) > 0.3).astype(np.float64)
    a = net.forward(nn.DevArray.from_numpy(x, "bf16"))
    part, inv, gy = net.loss_grad(a["y"], nn.DevArray.from_numpy(t, "f32"), nn.DevArray.from_numpy(mask, "f32"))
    net.backward(a, gy)
    nn.sync()
    act = {k: v.numpy().astype(np.float64) for k, v in a.items()}
    inputs = {"e1": "x", "e2": "e1", "e3": "p1", "e4": "e3", "e5": "p2", "e6": "e5", "e7": "p3", "e8": "e7",
              "d9": "c1", "d10": "c2", "d11": "c3", "head": "d11"}
    for k, src in inputs.items():
        ref = N.conv3(act[src], *p[k])
        if k != "head":
            ref = bf16_round(np.maximum(ref, 0.0))
        r = rel(act[k], ref)
        say(f"  forward {k}: rel {r:.2e}")
        assert r < 1e-2, (k, r)
    assert rel(act["p1"], bf16_round(N.pool2(act["e2"]))) < 1e-2
    c0, c1, c2, c3 = net.widths
    assert np.array_equal(act["c1"], np.concatenate([N.up2(act["e8"]), act["e6"]], axis=3))
    loss_ref, gy_ref = N.UNet.loss(act["y"], t, mask)
    assert abs(float(part.numpy().sum()) * inv - loss_ref) <= 1e-4 * loss_ref
    onet = N.UNet((8, 16, 16, 24), 8, 8)
    og = onet.backward(p, act, gy.numpy().astype(np.float64), q=bf16_round)
    gg = net.grads_numpy()
    bad = []
    for k in N.UNet.LAYERS:
        for j, what in ((0, "w"), (1, "b")):
            g_dev, g_ref = gg[k][j].ravel(), og[k][j].ravel()
            cos = float(g_dev @ g_ref / (np.linalg.norm(g_dev) * np.linalg.norm(g_ref) + 1e-30))
            r = float(np.linalg.norm(g_dev - g_ref) / (np.linalg.norm(g_ref) + 1e-30))
            say(f"  {k}.{what}: cos {cos:.6f} rel L2 {r:.3e}")
            if not (cos > 0.99995 and r < 1e-2):
                bad.append((k, what, cos, r))
    assert not bad, bad


def test_training_reduces_loss(nn):
    """Adam on the device fits smooth functions of position given the coordinates as input."""
    say("unet 16^3 fits a smooth target")
    net = nn.UNet(widths=(8, 16, 16, 16), cin=8, cout=8, n_out=4, seed=5)
    shape = (16, 16, 16)
    z, yy, xx = np.meshgrid(*[np.linspace(-1, 1, 16)] * 3, indexing="ij")
    x = np.zeros(shape + (8,))
    x[..., 0], x[..., 1], x[..., 2] = z, yy, xx
    t = np.zeros(shape + (8,))
    t[..., 0] = 0.5 * np.sin(2 * z)
    t[..., 1] = xx * yy
    mask = np.ones(shape)
    xd = nn.DevArray.from_numpy(x, "bf16")
    td, md = nn.DevArray.from_numpy(t, "f32"), nn.DevArray.from_numpy(mask, "f32")
    losses = []
    for it in range(200):
        a = net.forward(xd)
        part, inv, gy = net.loss_grad(a["y"], td, md)
        net.backward(a, gy)
        net.adam(3e-3)
        losses.append(float(part.numpy().sum()) * inv)
    say(f"  loss {losses[0]:.4e} -> {losses[-1]:.4e}")
    assert losses[-1] < 0.1 * losses[0]
    assert np.all(net.conv["head"].w.numpy()[4:8] == 0)


def test_nn_bench_runs(nn):
    """The convolution and training-step benchmark runs end to end at 16^3 and reports every
    layer with positive, finite rates."""
    say("nn_bench 16^3")
    from mflbm import nn_bench
    out = nn_bench.run(n=16, iters=2, train_steps=2, log=lambda s: None)
    assert [r["layer"] for r in out["layers"]] == [name for name, _, _ in nn_bench.LAYERS]
    for r in out["layers"]:
        assert np.isfinite(r["fwd_tflops"]) and r["fwd_tflops"] > 0
        assert np.isfinite(r["wgrad_tflops"]) and r["wgrad_tflops"] > 0
        assert ("dgrad_ms" in r) == (r["layer"] != "e1")
    assert out["train_step_ms"] > 0 and np.isfinite(out["train_tflops"])
