"""FP32-versus-FP64 precision study of the colour-gradient step on the device.

Runs the FP32 build and the FP64 build (bit-identical to the FP64 oracle, see the GPU parity
tests) side by side on the same workload and compares, every `interval` steps, the monitor
quantities the reference derives (src/Monitor.cpp:17-273), all reduced in FP64 on the device:

* saturation S = sum(vol1) / (sum(vol1) + sum(vol2));
* fluid fluxes fl1, fl2 (w-flux summed over slices, averaged over nz) and the capillary
  number Ca = (fl1_avg / A_xy) * nu1 / gamma (src/Monitor.cpp:170-171);
* total mass M = sum(mass1 + mass2), and each colour's mass on its own.

The phase fluxes under the same driving force stand for the relative permeabilities
(k_i = nu_i u_i / grad p), so their relative errors are the relative-permeability errors.

    python -m mflbm.precision --case grf_drainage512 --steps 50000 --interval 1000
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import cases, config, gpu


def summary(mon: dict, params: dict) -> dict:
    vol1, vol2 = float(np.sum(mon["vol1"])), float(np.sum(mon["vol2"]))
    nz = len(mon["fl1"])
    fl1 = float(np.sum(mon["fl1"])) / nz
    fl2 = float(np.sum(mon["fl2"])) / nz
    axy = float(params["nx"] * params["ny"])
    nu1 = 1.0 / params["la_nui1"]
    return {"S": vol1 / max(vol1 + vol2, 1e-300), "fl1": fl1, "fl2": fl2,
            "Ca": fl1 / axy * nu1 / params["lbm_gamma"],
            "mass": float(np.sum(mon["mass1"]) + np.sum(mon["mass2"])),
            "mass1": float(np.sum(mon["mass1"])), "mass2": float(np.sum(mon["mass2"]))}


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def run(case="grf_drainage512", steps=50000, interval=1000, n=None, log=print) -> dict:
    make = {"drainage256": cases.drainage256, "grf_drainage512": cases.grf_drainage512}[case]
    geo, ctrl = make(n=n) if n is not None else make()
    params = config.derive(ctrl)
    log(f"{case}: lattice {params['nx']}x{params['ny']}x{params['nz']}, porosity {1 - geo.mean():.3f}")
    sims = {"fp64": gpu.Solver(params, geo, gpu.FP64), "fp32": gpu.Solver(params, geo, gpu.FP32)}
    for s in sims.values():
        s.setup()
    series = []
    t_run = {"fp64": 0.0, "fp32": 0.0}
    m0 = {k: summary(s.monitor(), params)["mass"] for k, s in sims.items()}
    t = 1
    while t <= steps:
        nstep = min(interval, steps - t + 1)
        for k, s in sims.items():
            s.sync()
            t0 = time.perf_counter()
            s.run(t, nstep)
            s.sync()
            t_run[k] += time.perf_counter() - t0
        t += nstep
        a = summary(sims["fp64"].monitor(), params)
        b = summary(sims["fp32"].monitor(), params)
        row = {"step": t - 1, "S64": a["S"], "S32": b["S"], "dS": b["S"] - a["S"],
               "fl1_rel": rel(b["fl1"], a["fl1"]), "fl2_rel": rel(b["fl2"], a["fl2"]),
               "Ca64": a["Ca"], "Ca32": b["Ca"], "Ca_rel": rel(b["Ca"], a["Ca"]),
               "mass_drift64": (a["mass"] - m0["fp64"]) / m0["fp64"],
               "mass_drift32": (b["mass"] - m0["fp32"]) / m0["fp32"],
               "mass_rel": rel(b["mass"], a["mass"]),
               "mass1_rel": rel(b["mass1"], a["mass1"]), "mass2_rel": rel(b["mass2"], a["mass2"])}
        series.append(row)
        log(f"step {row['step']}: S64 {row['S64']:.6f} S32 {row['S32']:.6f} dS {row['dS']:+.2e} "
            f"fl1 {row['fl1_rel']:.2e} fl2 {row['fl2_rel']:.2e} mass {row['mass_rel']:.2e} "
            f"mass1 {row['mass1_rel']:.2e} mass2 {row['mass2_rel']:.2e}")
    for s in sims.values():
        s.close()
    nodes = params["nx"] * params["ny"] * params["nz"]
    last = series[-1]
    return {
        "case": case, "lattice": [params["nx"], params["ny"], params["nz"]], "steps": steps,
        "interval": interval, "reference": "FP64 build on the device (bit-identical to the FP64 oracle)",
        "final": {k: last[k] for k in ("S64", "S32", "dS", "fl1_rel", "fl2_rel", "Ca64", "Ca32", "Ca_rel",
                                       "mass_drift64", "mass_drift32", "mass_rel", "mass1_rel", "mass2_rel")},
        "max_over_run": {"abs_dS": max(abs(r["dS"]) for r in series),
                         "mass_rel": max(r["mass_rel"] for r in series),
                         "mass1_rel": max(r["mass1_rel"] for r in series),
                         "mass2_rel": max(r["mass2_rel"] for r in series)},
        "mlups": {k: nodes * steps / v / 1e6 for k, v in t_run.items()},
        "series": series,
    }


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--case", choices=["drainage256", "grf_drainage512"], default="grf_drainage512")
    ap.add_argument("--steps", type=int, default=50000)
    ap.add_argument("--interval", type=int, default=1000)
    ap.add_argument("--out", default=None, help="write the JSON result here as well")
    args = ap.parse_args(argv)
    res = run(args.case, args.steps, args.interval, log=lambda m: print(m, file=sys.stderr, flush=True))
    text = json.dumps(res)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(text + "\n")
    print(json.dumps({k: v for k, v in res.items() if k != "series"}), flush=True)


if __name__ == "__main__":
    main()
