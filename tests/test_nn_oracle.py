"""Pins the FP64 network oracle (oracle/nn_oracle.py) with adjoint identities and central
finite differences; the GPU kernels are then checked against it (tests/test_gpu_nn.py)."""
import numpy as np

import nn_oracle as N


def rand_params(net, rng, scale=0.3):
    p = {}
    for name, (cin, cout) in net.shapes.items():
        p[name] = (rng.normal(0, scale, (cout, 27 * cin)), rng.normal(0, 0.1, cout))
    return p


def test_conv_adjoint_and_wgrad():
    rng = np.random.default_rng(0)
    x = rng.normal(size=(4, 5, 6, 3))
    w = rng.normal(size=(2, 27 * 3))
    dz = rng.normal(size=(4, 5, 6, 2))
    # <conv(x), dz> == <x, dgrad(dz)>  (the data gradient is the adjoint)
    lhs = (N.conv3(x, w) * dz).sum()
    rhs = (x * N.conv3_dgrad(dz, w, 3)).sum()
    assert abs(lhs - rhs) <= 1e-10 * abs(lhs)
    # <conv(x; w), dz> is linear in w with gradient wgrad(x, dz)
    dw, db = N.conv3_wgrad(x, dz)
    assert abs((dw * w).sum() - lhs) <= 1e-10 * abs(lhs)
    assert np.allclose(db, dz.reshape(-1, 2).sum(0))
    # a single tap reproduces a plain shift
    w1 = np.zeros((3, 27 * 3))
    t = 5  # dz=-1, dy=0, dx=+1
    for c in range(3):
        w1[c, t * 3 + c] = 1.0
    assert np.allclose(N.conv3(x, w1), N.shift(x, -1, 0, 1))


def test_pool_up_adjoints():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(4, 6, 8, 3))
    y = rng.normal(size=(2, 3, 4, 3))
    assert abs((N.pool2(x) * y).sum() - (x * N.pool2_bwd(y)).sum()) < 1e-12
    assert abs((N.up2(y) * x).sum() - (y * N.up2_bwd(x)).sum()) < 1e-12
    assert np.allclose(N.pool2(N.up2(y)), y)


def test_unet_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    net = N.UNet((2, 3, 3, 4), cin=3, cout=2)
    p = rand_params(net, rng)
    x = rng.normal(size=(8, 8, 8, 3))
    t = rng.normal(size=(8, 8, 8, 2))
    mask = (rng.random((8, 8, 8)) > 0.3).astype(float)

    def loss(pp):
        a = net.forward(pp, x)
        return N.UNet.loss(a["y"], t, mask)[0]

    a = net.forward(p, x)
    _, gy = N.UNet.loss(a["y"], t, mask)
    g = net.backward(p, a, gy)
    h = 1e-6
    checked = 0
    for name in N.UNet.LAYERS:
        w, b = p[name]
        for _ in range(2):
            i = tuple(rng.integers(0, s) for s in w.shape)
            wp, wm = w.copy(), w.copy()
            wp[i] += h
            wm[i] -= h
            num = (loss({**p, name: (wp, b)}) - loss({**p, name: (wm, b)})) / (2 * h)
            ana = g[name][0][i]
            assert abs(num - ana) <= 1e-6 * max(1e-3, abs(num)) + 1e-9, (name, i, num, ana)
            checked += 1
        j = int(rng.integers(0, b.shape[0]))
        bp, bm = b.copy(), b.copy()
        bp[j] += h
        bm[j] -= h
        num = (loss({**p, name: (w, bp)}) - loss({**p, name: (w, bm)})) / (2 * h)
        assert abs(num - g[name][1][j]) <= 1e-6 * max(1e-3, abs(num)) + 1e-9, (name, "bias", num, g[name][1][j])
    assert checked == 2 * len(N.UNet.LAYERS)
