"""FP64 numpy oracle of the warm-start network (test infrastructure only).

Restates, in float64, what the HIP kernels of mf-lbm-cuda_amd/csrc/nn.hip compute: 3x3x3
zero-padded convolutions on channels-last grids, ReLU, 2x2x2 average pooling, nearest 2x
upsampling, channel concatenation, a masked squared error, and the backward pass of the
U-Net assembled from them (mflbm/nn.py).  The network has no counterpart in the reference;
its check is mathematical: tests/test_nn_oracle.py pins these gradients with central finite
differences.

Weights use the kernels' layout W[n][k], k = tap * cin + ci, tap = (dz+1)*9 + (dy+1)*3 + (dx+1).
"""
from __future__ import annotations

import numpy as np

OFFS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def shift(x, dz, dy, dx):
    """out[z, y, x] = x[z + dz, y + dy, x + dx] (zero outside)."""
    D, H, W = x.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    return xp[1 + dz:1 + dz + D, 1 + dy:1 + dy + H, 1 + dx:1 + dx + W]


def conv3(x, w, b=None):
    """x [D,H,W,cin], w [cout, 27*cin] -> [D,H,W,cout]"""
    cin = x.shape[3]
    cout = w.shape[0]
    wt = w[:, :27 * cin].reshape(cout, 27, cin)
    y = np.zeros(x.shape[:3] + (cout,))
    for t, (dz, dy, dx) in enumerate(OFFS):
        y += shift(x, dz, dy, dx) @ wt[:, t, :].T
    if b is not None:
        y += b
    return y


def conv3_wgrad(x, dz):
    """dW[n][tap*cin + ci] = sum_m dz[m, n] x[m + off(tap), ci]"""
    cin = x.shape[3]
    cout = dz.shape[3]
    dw = np.zeros((cout, 27, cin))
    dzf = dz.reshape(-1, cout)
    for t, (oz, oy, ox) in enumerate(OFFS):
        dw[:, t, :] = dzf.T @ shift(x, oz, oy, ox).reshape(-1, cin)
    return dw.reshape(cout, 27 * cin), dzf.sum(axis=0)


def conv3_dgrad(dz, w, cin):
    """dX[m][ci] = sum_{tap, n} w[n][tap*cin + ci] dz[m - off(tap), n]"""
    cout = dz.shape[3]
    wt = w[:, :27 * cin].reshape(cout, 27, cin)
    dx = np.zeros(dz.shape[:3] + (cin,))
    for t, (oz, oy, ox) in enumerate(OFFS):
        dx += shift(dz, -oz, -oy, -ox) @ wt[:, t, :]
    return dx


def pool2(x):
    D, H, W, C = x.shape
    return x.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).mean(axis=(1, 3, 5))


def pool2_bwd(dy):
    return np.repeat(np.repeat(np.repeat(dy, 2, 0), 2, 1), 2, 2) / 8.0


def up2(x):
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def up2_bwd(dy):
    D, H, W, C = dy.shape
    return dy.reshape(D // 2, 2, H // 2, 2, W // 2, 2, C).sum(axis=(1, 3, 5))


def relu(x):
    return np.maximum(x, 0.0)


class UNet:
    """The four-scale U-Net of mflbm/nn.py (widths c = (c0, c1, c2, c3), cin input and cout
    output channels), with weights given as {name: (W [n][k], b [n])}."""

    LAYERS = ("e1", "e2", "e3", "e4", "e5", "e6", "e7", "e8", "d9", "d10", "d11", "head")

    def __init__(self, widths, cin, cout):
        c0, c1, c2, c3 = widths
        self.cin, self.cout, self.widths = cin, cout, widths
        self.shapes = {"e1": (cin, c0), "e2": (c0, c0), "e3": (c0, c1), "e4": (c1, c1), "e5": (c1, c2),
                       "e6": (c2, c2), "e7": (c2, c3), "e8": (c3, c3), "d9": (c3 + c2, c2),
                       "d10": (c2 + c1, c1), "d11": (c1 + c0, c0), "head": (c0, cout)}

    def forward(self, p, x, q=None):
        """Activations; q (optional) rounds each stored tensor the way the device stores it
        (bf16 after every ReLU convolution and pooling; the head output stays fp32)."""
        q = q or (lambda v: v)
        a = {}
        a["x"] = x
        a["e1"] = q(relu(conv3(x, *p["e1"])))
        a["e2"] = q(relu(conv3(a["e1"], *p["e2"])))
        a["p1"] = q(pool2(a["e2"]))
        a["e3"] = q(relu(conv3(a["p1"], *p["e3"])))
        a["e4"] = q(relu(conv3(a["e3"], *p["e4"])))
        a["p2"] = q(pool2(a["e4"]))
        a["e5"] = q(relu(conv3(a["p2"], *p["e5"])))
        a["e6"] = q(relu(conv3(a["e5"], *p["e6"])))
        a["p3"] = q(pool2(a["e6"]))
        a["e7"] = q(relu(conv3(a["p3"], *p["e7"])))
        a["e8"] = q(relu(conv3(a["e7"], *p["e8"])))
        a["c1"] = np.concatenate([up2(a["e8"]), a["e6"]], axis=3)
        a["d9"] = q(relu(conv3(a["c1"], *p["d9"])))
        a["c2"] = np.concatenate([up2(a["d9"]), a["e4"]], axis=3)
        a["d10"] = q(relu(conv3(a["c2"], *p["d10"])))
        a["c3"] = np.concatenate([up2(a["d10"]), a["e2"]], axis=3)
        a["d11"] = q(relu(conv3(a["c3"], *p["d11"])))
        a["y"] = conv3(a["d11"], *p["head"])
        return a

    @staticmethod
    def loss(y, t, mask):
        d = (y - t) * mask[..., None]
        n = y.size
        return float((d * d).sum() / n), 2.0 * d / n

    def backward(self, p, a, gy, q=None):
        """Gradients {name: (dW, db)} of the loss whose output gradient is gy; q (optional)
        rounds each stored gradient tensor the way the device stores it (bf16)."""
        q = q or (lambda v: v)
        g = {}
        c0, c1, c2, c3 = self.widths

        def layer(name, x, dz, gate=None, need_dx=True):
            dw, db = conv3_wgrad(x, dz)
            g[name] = (dw, db)
            if not need_dx:
                return None
            dx = conv3_dgrad(dz, p[name][0], x.shape[3])
            return q(dx * (gate > 0) if gate is not None else dx)

        gy = q(gy)
        gd11 = layer("head", a["d11"], gy, a["d11"])
        gc3 = layer("d11", a["c3"], gd11)
        gd10 = q(up2_bwd(gc3[..., :c1])) * (a["d10"] > 0)
        ge2_a = gc3[..., c1:]
        gc2 = layer("d10", a["c2"], gd10)
        gd9 = q(up2_bwd(gc2[..., :c2])) * (a["d9"] > 0)
        ge4_a = gc2[..., c2:]
        gc1 = layer("d9", a["c1"], gd9)
        ge8 = q(up2_bwd(gc1[..., :c3])) * (a["e8"] > 0)
        ge6_a = gc1[..., c3:]
        ge7 = layer("e8", a["e7"], ge8, a["e7"])
        ge6 = q((ge6_a + q(pool2_bwd(layer("e7", a["p3"], ge7)))) * (a["e6"] > 0))
        ge5 = layer("e6", a["e5"], ge6, a["e5"])
        ge4 = q((ge4_a + q(pool2_bwd(layer("e5", a["p2"], ge5)))) * (a["e4"] > 0))
        ge3 = layer("e4", a["e3"], ge4, a["e3"])
        ge2 = q((ge2_a + q(pool2_bwd(layer("e3", a["p1"], ge3)))) * (a["e2"] > 0))
        ge1 = layer("e2", a["e1"], ge2, a["e1"])
        layer("e1", a["x"], ge1, need_dx=False)
        return g
