#!/usr/bin/env python3
"""Benchmark of the MI355X colour-gradient LBM step.

Workload (default): 256^3 two-colour D3Q19 MRT drainage, FP64 — seeded sphere pack
(radius 8, porosity 0.35, 8 empty buffer layers at inlet and outlet), contact angle 30
deg, viscosity ratio 0.1, Zou-He pressure inlet/outlet along the flow axis, periodic
transverse axis, walls on the other.  Inputs are resident on the device before timing.

One step = collide/recolour/stream + boundary conditions + colour-gradient chain
(the reference's main_iteration_kernel_GPU without its device-to-host copy).
MLUPS follows the reference's definition (src/main.cpp:194,270): all lattice nodes,
solids included, per wall-clock second.

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "mf-lbm-cuda_amd"))

from mflbm import cases, config, gpu  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(geo, params, steps):
    """Time the OpenMP FP64 oracle on this host over `steps` steps of the same workload."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # the checker doubles as the reported CPU baseline

    threads = O.threads()
    o = O.Oracle(params, geo)
    o.setup()
    o.run(1, 2)  # warm caches / first-touch
    t0 = time.perf_counter()
    o.run(3, steps)
    dt = time.perf_counter() - t0
    nodes = params["nx"] * params["ny"] * params["nz"]
    o.close()
    return {"value": nodes * steps / dt / 1e6, "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"{steps} steps of the same {params['nx']}x{params['ny']}x{params['nz']} workload, "
                      f"OpenMP FP64 oracle, {threads} threads on {cpu_model()}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=400)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--n", type=int, default=256, help="cubic lattice edge")
    ap.add_argument("--cpu-steps", type=int, default=8, help="steps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--probe-mib", type=int, default=2048, help="plain-copy calibration size (0 = skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU z-slab "
              "decomposition is a design note, not built.", file=sys.stderr)
        sys.exit(2)

    prec = gpu.FP64 if args.precision == "fp64" else gpu.FP32
    s = 8 if prec == gpu.FP64 else 4
    n = args.n
    log(f"generating {n}^3 sphere pack")
    geo, ctrl = cases.config2(n=(n, n, n))
    params = config.derive(ctrl)
    log("creating device state")
    sim = gpu.Solver(params, geo, prec)
    sim.setup()
    counts = sim.counts()
    n_fluid = counts["fluid"]
    n_all = n * n * n
    log(f"fluid nodes {n_fluid} ({n_fluid / n_all:.3f}), solid-boundary {counts['solid_boundary']}")

    sim.run(1, args.warmup)
    ntime0 = 1 + args.warmup
    sim.sync()
    t0 = time.perf_counter()
    sim.run(ntime0, args.steps)
    sim.sync()
    dt = time.perf_counter() - t0
    mlups = n_all * args.steps / dt / 1e6
    mflups = n_fluid * args.steps / dt / 1e6
    log(f"{args.steps} steps in {dt:.3f} s: {mlups:.1f} MLUPS, {mflups:.1f} fluid-MLUPS")

    # per-kernel device times (events on the solver's stream)
    kt = sim.profile(ntime0 + args.steps, args.profile_steps)
    coll_ms = (kt["collide_odd_ms"] + kt["collide_even_ms"]) / max(1, kt["n_odd"] + kt["n_even"])
    step_ms = sum(kt[k] for k in ("collide_odd_ms", "collide_even_ms", "boundary_ms", "phi_extrap_ms", "normals_ms",
                                  "normal_extrap_ms", "curvature_ms")) / max(1, kt["n_odd"] + kt["n_even"])
    # algorithmic bytes of one collide launch: 38 slots read + 38 written + phi written per
    # fluid node, plus one mask byte per lattice node
    coll_bytes = n_fluid * (76 * s + s) + n_all * 1
    achieved = coll_bytes / (coll_ms * 1e-3) / 1e9
    step_bytes = n_fluid * 78 * s + n_all
    log(f"collide {coll_ms:.3f} ms ({achieved:.0f} GB/s algorithmic), step (kernels) {step_ms:.3f} ms")

    probe = None
    if args.probe_mib > 0:
        nbytes = args.probe_mib << 20
        pms = sim.bandwidth_probe(nbytes, repeats=10)
        probe = {"bytes": nbytes, "ms": round(pms, 4), "gbs": round(2 * nbytes / (pms * 1e-3) / 1e9, 1)}
        log(f"plain copy of {args.probe_mib} MiB: {pms:.3f} ms = {probe['gbs']} GB/s (read+write)")

    cpu = None
    if args.cpu_steps > 0:
        log(f"CPU baseline: {args.cpu_steps} oracle steps")
        cpu = cpu_baseline(geo, params, args.cpu_steps)
        log(f"CPU baseline {cpu['value']:.2f} MLUPS on {cpu['cores']} threads")

    out = {
        "metric": "MLUPS", "value": round(mlups, 3), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == gpu.FP64 else "f32",
        "data": "synthetic: seeded sphere pack",
        "config": {"workload": f"colour-gradient drainage {n}^3, D3Q19x2 MRT, AA streaming, "
                               f"{'FP64' if prec == gpu.FP64 else 'FP32'}",
                   "lattice": [n, n, n], "porosity": round(1 - float(geo.mean()), 4), "fluid_nodes": int(n_fluid)},
        "fluid_mlups": round(mflups, 3),
        "step_roofline_frac": round(step_bytes / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "kernel_ms": {k: round(v / max(1, kt["n_odd"] + kt["n_even"]), 4) for k, v in kt.items() if k.endswith("_ms")},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_collide (odd+even average)"},
        "copy_probe": probe,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
