"""Throughput of the BF16-MFMA convolution kernels and of one U-Net training step.

    python -m mflbm.nn_bench [--n 64] [--iters 20]

For every layer of the warm-start U-Net at n^3 (widths 16/32/64/128, 8 input and 8 output
channels): the forward convolution (`mfnn_conv3`), the weight gradient (`mfnn_conv3_wgrad`)
and the data gradient (`mfnn_conv3` on the flipped weights, gated by the layer input), each
timed over `iters` launches between device synchronisations.  Algorithmic flops of each are
2 * voxels * 27 * cin * cout (the padding of the tap and output dimensions not counted).  Then
whole training steps (forward, masked squared error, backward, Adam) of the network on one
n^3 sample.  Inputs are seeded random values.  Prints one JSON line; rates are in TFLOP/s
against the 2500 TFLOP/s dense BF16 MFMA peak of the MI355X.
"""
import argparse
import json
import sys
import time

import numpy as np

from mflbm import nn

PEAK_BF16_TFLOPS = 2500.0

# layer -> (input activation, grid level); level l has n / 2^l nodes per side
LAYERS = [("e1", "x", 0), ("e2", "e1", 0), ("e3", "p1", 1), ("e4", "e3", 1), ("e5", "p2", 2), ("e6", "e5", 2),
          ("e7", "p3", 3), ("e8", "e7", 3), ("d9", "c1", 2), ("d10", "c2", 1), ("d11", "c3", 0), ("head", "d11", 0)]


def _timed(fn, iters):
    nn.sync()
    t0 = time.perf_counter()
    for _ in range(iters):
        fn()
    nn.sync()
    return (time.perf_counter() - t0) / iters * 1e3


def run(n=64, iters=20, train_steps=10, widths=(16, 32, 64, 128), log=print):
    if n % 8:
        raise ValueError("n must be a multiple of 8")
    rng = np.random.default_rng(7)
    net = nn.UNet(widths=widths, cin=8, cout=8)
    x = nn.DevArray.from_numpy(rng.standard_normal((n, n, n, 8)).astype(np.float32), "bf16")
    target = nn.DevArray.from_numpy(rng.standard_normal((n, n, n, 8)).astype(np.float32), "f32")
    mask = nn.DevArray.from_numpy((rng.random((n, n, n)) < 0.4).astype(np.float32), "f32")
    a = net.forward(x)
    work = net._workspace(n, n, n)
    rows = []
    total_flops = 0.0
    for name, src, lev in LAYERS:
        c = net.conv[name]
        s = n >> lev
        shape = (s, s, s)
        inp = a[src]
        flops = 2.0 * s ** 3 * 27 * c.cin * c.cout
        relu = name != "head"
        if name == "head":
            yf = a[name]
            fwd = _timed(lambda: c.forward(inp, shape, relu, yf=yf), iters)
        else:
            yb = a[name]
            fwd = _timed(lambda: c.forward(inp, shape, relu, yb=yb), iters)
        dz = nn.DevArray.from_numpy(rng.standard_normal(shape + (c.cout,)).astype(np.float32), "bf16")
        wgr = _timed(lambda: c.backward(inp, dz, shape, work, 0), iters)
        row = {"layer": name, "grid": s, "cin": c.cin, "cout": c.cout, "gflop": flops / 1e9,
               "fwd_ms": fwd, "fwd_tflops": flops / fwd / 1e9, "wgrad_ms": wgr, "wgrad_tflops": flops / wgr / 1e9}
        total_flops += 2 * flops
        if name != "e1":  # the network takes no data gradient of its input
            dx = nn.DevArray(shape + (c.cin,), "bf16")
            both = _timed(lambda: c.backward(inp, dz, shape, work, 0, dx=dx, gate=inp), iters)
            dgr = max(both - wgr, 1e-6)
            row.update({"dgrad_ms": dgr, "dgrad_tflops": flops / dgr / 1e9})
            total_flops += flops
        rows.append(row)
        log(f"{name:5s} {s:4d}^3 {c.cin:4d}->{c.cout:4d}  fwd {fwd:8.3f} ms {row['fwd_tflops']:7.1f} TF/s  "
            f"wgrad {wgr:8.3f} ms {row['wgrad_tflops']:7.1f} TF/s" +
            (f"  dgrad {row['dgrad_ms']:8.3f} ms {row['dgrad_tflops']:7.1f} TF/s" if "dgrad_ms" in row else ""))

    def step():
        act = net.forward(x)
        _, _, gy = net.loss_grad(act["y"], target, mask)
        net.backward(act, gy)
        net.adam(1e-4)

    step()
    ms = _timed(step, train_steps)
    conv_ms = sum(r["fwd_ms"] + r["wgrad_ms"] + r.get("dgrad_ms", 0.0) for r in rows)
    best = max(rows, key=lambda r: r["gflop"])
    out = {"metric": "unet_train_step", "n": n, "widths": list(widths), "params": net.n_params(),
           "wgrad_chunk": [nn.WGRAD_CHUNK, nn.WGRAD_CHUNK_SMALL],
           "train_step_ms": ms, "train_tflops": total_flops / ms / 1e9,
           "train_frac": total_flops / ms / 1e9 / PEAK_BF16_TFLOPS, "conv_ms_sum": conv_ms,
           "largest_layer": {"layer": best["layer"], "fwd_tflops": best["fwd_tflops"],
                             "fwd_frac": best["fwd_tflops"] / PEAK_BF16_TFLOPS},
           "peak_tflops": PEAK_BF16_TFLOPS, "layers": rows}
    log(f"training step {ms:.3f} ms, {out['train_tflops']:.1f} TFLOP/s ({out['train_frac']:.3f} of peak)")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--train-steps", type=int, default=10)
    ap.add_argument("--chunk", type=int, default=nn.WGRAD_CHUNK,
                    help="voxels per weight-gradient chunk on grids of 65536 voxels and more")
    ap.add_argument("--small-chunk", type=int, default=nn.WGRAD_CHUNK_SMALL,
                    help="voxels per weight-gradient chunk on smaller grids")
    args = ap.parse_args(argv)
    nn.WGRAD_CHUNK, nn.WGRAD_CHUNK_SMALL = args.chunk, args.small_chunk
    out = run(args.n, args.iters, args.train_steps, log=lambda s: print(s, file=sys.stderr, flush=True))
    print(json.dumps(out))


if __name__ == "__main__":
    main()
