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
import csv
import glob
import json
import os
import platform
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "mf-lbm-cuda_amd"))

from mflbm import cases, config, gpu  # noqa: E402

HBM_PEAK_GBS = 8000.0


T_START = time.perf_counter()


def log(*a):
    print(f"[bench {time.perf_counter() - T_START:7.1f}s]", *a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline_child(args):
    """Child process of the CPU baseline: times the OpenMP FP64 oracle on this host over
    `--cpu-steps` steps of the same workload (the reference's timing: steady clock around
    the step loop only, src/main.cpp:142-199) and prints one JSON line."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle as O  # the oracle doubles as the reported CPU baseline

    geo = np.load(args.cpu_child, allow_pickle=False)
    n = geo.shape[0]
    _, ctrl = cases.drainage256(n=(n, n, n), geometry=False)
    params = config.derive(ctrl)
    threads = O.threads()
    o = O.Oracle(params, geo)
    o.setup()
    o.run(1, 2)  # first touch of every page by the threads that own it
    t0 = time.perf_counter()
    o.run(3, args.cpu_steps)
    dt = time.perf_counter() - t0
    o.close()
    print(json.dumps({"seconds": dt, "threads": threads}), flush=True)


def cpu_baseline(geo, steps):
    """The CPU baseline in a child process whose OpenMP runtime starts with the placement
    the reference's plan states: OMP_PROC_BIND=close, OMP_PLACES=cores, all host threads
    (OMP_NUM_THREADS as set on the host)."""
    tmp = tempfile.mkdtemp(prefix="mflbm_cpu_", dir="/tmp")
    geo_path = os.path.join(tmp, "geo.npy")
    np.save(geo_path, geo)
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child", geo_path,
                            "--cpu-steps", str(steps)], env=env, capture_output=True, text=True, timeout=600)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed (rc {r.returncode}): {r.stderr[-400:]}")
    res = json.loads(r.stdout.strip().splitlines()[-1])
    n = geo.shape[0]
    nodes = geo.size
    return {"value": nodes * steps / res["seconds"] / 1e6, "unit": "MLUPS", "cores": res["threads"], "kind": "port",
            "seconds": round(res["seconds"], 2),
            "sample": f"{steps} steps (3 .. {steps + 2}) of the same {n}x{n}x{n} workload, OpenMP FP64 oracle, "
                      f"{res['threads']} threads (OMP_PROC_BIND=close, OMP_PLACES=cores) on {cpu_model()}"}


PROBE_KERNEL = "k_probe_copy"
PMC_STEPS = 4  # steps 1..4: two odd and two even collisions


def kernel_group(name):
    """Short name of a step kernel from its demangled name (None for other kernels)."""
    if "k_probe_copy_blocked" in name:
        return "probe_blocked"
    if PROBE_KERNEL in name:
        return "probe"
    for key, short in (("k_collide<double, true", "collide_odd"), ("k_collide<float, true", "collide_odd"),
                       ("k_collide<double, false", "collide_even"), ("k_collide<float, false", "collide_even"),
                       ("k_gradient<", "gradient"), ("k_grad_normals<", "grad_normals"),
                       ("k_grad_force<", "grad_force")):
        if key in name:
            return short
    return None


def pmc_child(args):
    """Run under `rocprofv3 --pmc`: a few plain steps of the workload and one probe copy."""
    prec = gpu.FP64 if args.precision == "fp64" else gpu.FP32
    geo = np.load(args.pmc_child, allow_pickle=False)
    n = geo.shape[0]
    _, ctrl = cases.drainage256(n=(n, n, n), geometry=False)
    sim = gpu.Solver(config.derive(ctrl), geo, prec)
    sim.setup()
    for t in range(1, PMC_STEPS + 1):
        sim.step(t)
    sim.bandwidth_probe(args.probe_mib << 20, repeats=1)
    sim.sync()
    sim.close()
    print("pmc child done", flush=True)


def counter_pass(counters, geo_path, args, timeout=240):
    """Values of the given counters per dispatch from one rocprofv3 --pmc pass, grouped by
    kernel: {group: {counter: [value per dispatch]}}."""
    rp = shutil.which("rocprofv3") or "/opt/rocm/bin/rocprofv3"
    out = tempfile.mkdtemp(prefix="mflbm_pmc_", dir="/tmp")
    cmd = [rp, "--pmc", *counters, "-d", out, "-o", "pmc", "--output-format", "csv", "--",
           sys.executable, os.path.abspath(__file__), "--pmc-child", geo_path, "--precision", args.precision,
           "--probe-mib", str(args.probe_mib)]
    env = dict(os.environ, TMPDIR="/tmp")
    log(f"rocprofv3 --pmc {' '.join(counters)}")
    proc = subprocess.Popen(cmd, cwd="/tmp", env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                            start_new_session=True, text=True)
    try:
        text, _ = proc.communicate(timeout=timeout)
    except subprocess.TimeoutExpired:
        os.killpg(proc.pid, signal.SIGKILL)
        proc.communicate()
        shutil.rmtree(out, ignore_errors=True)
        raise RuntimeError(f"rocprofv3 --pmc {counters} timed out")
    # rocprofv3 itself may end with SIGSEGV after a complete run; the child's last line and
    # the result file decide
    files = glob.glob(os.path.join(out, "**", "*counter_collection.csv"), recursive=True)
    if "pmc child done" not in text or not files:
        shutil.rmtree(out, ignore_errors=True)
        raise RuntimeError(f"rocprofv3 --pmc {counters} failed (rc {proc.returncode}): {text[-400:]}")
    vals = {}
    with open(files[0], newline="") as fh:
        for row in csv.DictReader(fh):
            g = kernel_group(row["Kernel_Name"])
            if g:
                vals.setdefault(g, {}).setdefault(row["Counter_Name"], []).append(float(row["Counter_Value"]))
    shutil.rmtree(out, ignore_errors=True)
    return vals


SQ_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_LDS_BANK_CONFLICT", "SQ_INSTS_VMEM_RD",
               "SQ_BUSY_CYCLES", "SQ_WAVE_CYCLES")


def sq_counters(args):
    """Instruction mix of the step kernels: one rocprofv3 --pmc pass of eight SQ counters over
    the counter child (steps 1-4 of the workload), printed as per-dispatch means and per-wave
    figures."""
    geo, _ = cases.drainage256(n=(args.n, args.n, args.n))
    tmp = tempfile.mkdtemp(prefix="mflbm_geo_", dir="/tmp")
    try:
        path = os.path.join(tmp, "geometry.npy")
        np.save(path, geo)
        vals = counter_pass(list(SQ_COUNTERS), path, args)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    out = {}
    for g, d in vals.items():
        m = {c: float(np.mean(v)) for c, v in d.items()}
        w = max(1.0, m.get("SQ_WAVES", 1.0))
        out[g] = {"per_dispatch": m,
                  "per_wave": {c.replace("SQ_INSTS_", "").lower(): round(m[c] / w, 1)
                               for c in ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_INSTS_LDS", "SQ_INSTS_VMEM_RD") if c in m}}
    print(json.dumps({"precision": args.precision, "lattice": [args.n] * 3, "sq": out}), flush=True)


GROUPS = ("collide_odd", "collide_even", "gradient", "grad_normals", "grad_force", "probe", "probe_blocked")


def measure_traffic(geo, args):
    """HBM bytes per dispatch of the step kernels.

    Two measures, each from rocprofv3 --pmc passes over a child running the same workload:
    * `requests`: the L2-to-memory read requests by size (32, 64, 128 B) and write requests
      (32, 64 B), summed over all L2 channels, times their sizes;
    * `fetch_write`: the FETCH_SIZE and WRITE_SIZE derived counters, scaled so that the
      grid-stride plain copy of known size reads and writes exactly its size.
    Both plain copies are reported so that each measure can be checked against a known size."""
    tmp = tempfile.mkdtemp(prefix="mflbm_geo_", dir="/tmp")
    geo_path = os.path.join(tmp, "geo.npy")
    np.save(geo_path, geo)
    try:
        rd = counter_pass(["TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum",
                           "TCC_EA0_RDREQ_sum"], geo_path, args)
        wr = counter_pass(["TCC_EA0_WRREQ_sum", "TCC_EA0_WRREQ_64B_sum"], geo_path, args)
        f = counter_pass(["FETCH_SIZE"], geo_path, args)
        w = counter_pass(["WRITE_SIZE"], geo_path, args)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    med = lambda v: float(np.median(v))
    nbytes = args.probe_mib << 20
    res = {"copy_bytes": nbytes, "requests": {}, "fetch_write": {}}
    for g in GROUPS:
        if g in rd and g in wr:
            r = rd[g]
            rb = 32 * med(r["TCC_EA0_RDREQ_32B_sum"]) + 64 * med(r["TCC_EA0_RDREQ_64B_sum"]) + \
                128 * med(r["TCC_EA0_RDREQ_128B_sum"])
            w64 = med(wr[g]["TCC_EA0_WRREQ_64B_sum"])
            wb = 64 * w64 + 32 * (med(wr[g]["TCC_EA0_WRREQ_sum"]) - w64)
            res["requests"][g] = {"fetch": round(rb), "write": round(wb), "total": round(rb + wb)}
    sf = nbytes / (med(f["probe"]["FETCH_SIZE"]) * 1024.0)
    sw = nbytes / (med(w["probe"]["WRITE_SIZE"]) * 1024.0)
    res["fetch_write"]["scale"] = {"fetch": round(sf, 4), "write": round(sw, 4)}
    for g in GROUPS:
        if g in f and g in w:
            fb = med(f[g]["FETCH_SIZE"]) * 1024.0 * sf
            wb = med(w[g]["WRITE_SIZE"]) * 1024.0 * sw
            res["fetch_write"][g] = {"fetch": round(fb), "write": round(wb), "total": round(fb + wb)}
    for kind in ("requests", "fetch_write"):
        d = res[kind]
        if "gradient" not in d and "grad_normals" in d and "grad_force" in d:  # the two-pass chain
            d["gradient"] = {k: d["grad_normals"][k] + d["grad_force"][k] for k in ("fetch", "write", "total")}
        log(f"traffic per dispatch ({kind}, GB): " + ", ".join(
            f"{g} {d[g]['fetch'] / 1e9:.3f}+{d[g]['write'] / 1e9:.3f}" for g in GROUPS if g in d))
    return res


def measure_precision(params, geo, n_all, prec, steps, warmup, profile_steps):
    """Step rate and per-kernel times of the same workload in another precision (a fresh
    handle: set-up, `warmup` untimed steps, `steps` timed ones), for the line's second-build
    entry; the collide's fraction of the HBM peak on its algorithmic bytes as for the main
    figure."""
    s = 8 if prec == gpu.FP64 else 4
    sim = gpu.Solver(params, geo, prec)
    try:
        sim.setup()
        n_fluid = sim.counts()["fluid"]
        sim.run(1, warmup)
        t1 = 1 + warmup
        sim.sync()
        t0 = time.perf_counter()
        sim.run(t1, steps)
        sim.sync()
        dt = time.perf_counter() - t0
        kt = sim.profile(t1 + steps, profile_steps)
        chain = sim.gradient_path()
    finally:
        sim.close()
    ns = max(1, kt["n_odd"] + kt["n_even"])
    odd = kt["collide_odd_ms"] / max(1, kt["n_odd"])
    even = kt["collide_even_ms"] / max(1, kt["n_even"])
    coll = 0.5 * (odd + even)
    return {"dtype": "f64" if prec == gpu.FP64 else "f32", "steps": steps, "warmup": warmup,
            "ms_per_step": round(dt / steps * 1e3, 4), "mlups": round(n_all * steps / dt / 1e6, 3),
            "fluid_glups": round(n_fluid * steps / dt / 1e9, 4),
            "collide_roofline_frac": round(n_fluid * 77 * s / (coll * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "step_roofline_frac": round((n_fluid * 78 * s + n_all) / (dt / steps) / 1e9 / HBM_PEAK_GBS, 4),
            "gradient_chain": chain,
            "kernel_ms": {"collide_odd": round(odd, 4), "collide_even": round(even, 4),
                          "boundary": round(kt["boundary_ms"] / ns, 4),
                          "gradient": round(kt["gradient_ms"] / ns, 4)}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--n", type=int, default=256, help="cubic lattice edge")
    ap.add_argument("--cpu-steps", type=int, default=100,
                    help="steps of the CPU baseline sample, about 30 s on 16 host cores at 256^3 (0 = skip)")
    ap.add_argument("--profile-steps", type=int, default=20)
    ap.add_argument("--late-steps", type=int, default=4000,
                    help="further drainage steps before a second, later measurement (0 = skip)")
    ap.add_argument("--probe-mib", type=int, default=1024, help="plain-copy calibration size (0 = skip)")
    ap.add_argument("--traffic", action=argparse.BooleanOptionalAction, default=True,
                    help="measure HBM bytes per kernel with two rocprofv3 --pmc passes")
    ap.add_argument("--other-steps", type=int, default=1000,
                    help="timed steps of the same workload in the other precision (0 = skip)")
    ap.add_argument("--sq-counters", action="store_true",
                    help="print the step kernels' SQ instruction counts (one rocprofv3 --pmc pass) and exit")
    ap.add_argument("--pmc-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--cpu-child", default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.pmc_child:
        pmc_child(args)
        return
    if args.cpu_child:
        cpu_baseline_child(args)
        return
    if args.sq_counters:
        sq_counters(args)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU z-slab "
              "decomposition is a design note, not built.", file=sys.stderr)
        sys.exit(2)

    prec = gpu.FP64 if args.precision == "fp64" else gpu.FP32
    s = 8 if prec == gpu.FP64 else 4
    n = args.n
    log(f"generating {n}^3 sphere pack")
    geo, ctrl = cases.drainage256(n=(n, n, n))
    params = config.derive(ctrl)

    traffic = None
    if args.traffic and args.probe_mib > 0:
        try:  # before this process touches the GPU
            traffic = measure_traffic(geo, args)
        except (RuntimeError, OSError, KeyError, IndexError) as e:
            log(f"traffic measurement skipped: {e}")

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

    # per-kernel device times: HIP events on the solver's stream, around each kernel group
    kt = sim.profile(ntime0 + args.steps, args.profile_steps)
    n_steps = max(1, kt["n_odd"] + kt["n_even"])
    odd_ms = kt["collide_odd_ms"] / max(1, kt["n_odd"])
    even_ms = kt["collide_even_ms"] / max(1, kt["n_even"])
    coll_ms = 0.5 * (odd_ms + even_ms)
    grad_ms = kt["gradient_ms"] / n_steps
    bnd_ms = kt["boundary_ms"] / n_steps
    step_ms = coll_ms + bnd_ms + grad_ms
    # algorithmic bytes of one collide launch: 38 slots read + 38 written + phi written per
    # fluid node (77 s); the force records it reads are derivable from phi and not counted
    coll_bytes = n_fluid * 77 * s
    achieved = coll_bytes / (coll_ms * 1e-3) / 1e9
    # one step: 76 s of populations + phi written once and read once per fluid node, plus
    # one byte of node type per lattice node, over the timed wall-clock step
    step_bytes = n_fluid * 78 * s + n_all
    step_frac = step_bytes / (dt / args.steps) / 1e9 / HBM_PEAK_GBS
    log(f"collide odd {odd_ms:.3f} ms, even {even_ms:.3f} ms ({achieved:.0f} GB/s algorithmic), "
        f"boundary {bnd_ms:.3f} ms, gradient {grad_ms:.3f} ms, step (kernels) {step_ms:.3f} ms")

    # Later in the drainage: the gradient chain skips work away from interfaces, so its cost
    # grows with the invaded volume.  Advance further and measure again.
    late = None
    if args.late_steps > 0:
        t_late = ntime0 + args.steps + args.profile_steps
        sim.run(t_late, args.late_steps)
        t_late += args.late_steps
        nl = 200
        sim.sync()
        t0 = time.perf_counter()
        sim.run(t_late, nl)
        sim.sync()
        dl = time.perf_counter() - t0
        kl = sim.profile(t_late + nl, args.profile_steps)
        nls = max(1, kl["n_odd"] + kl["n_even"])
        late = {"after_step": t_late - 1, "ms_per_step": round(dl / nl * 1e3, 4),
                "mlups": round(n_all * nl / dl / 1e6, 3),
                "kernel_ms": {"collide_odd": round(kl["collide_odd_ms"] / max(1, kl["n_odd"]), 4),
                              "collide_even": round(kl["collide_even_ms"] / max(1, kl["n_even"]), 4),
                              "boundary": round(kl["boundary_ms"] / nls, 4),
                              "gradient": round(kl["gradient_ms"] / nls, 4)}}
        log(f"after step {t_late - 1}: {late['mlups']:.1f} MLUPS, gradient {late['kernel_ms']['gradient']:.3f} ms")

    # the same kernels on a colour field with no uniform region (random phi in [-1, 1] on
    # every node, populations re-initialised from it): the gradient chain's worst case,
    # where no plane takes the exact uniform-region shortcut
    rng = np.random.default_rng(1)
    sim.upload("phi", rng.uniform(-1.0, 1.0, size=sim.shape(4)))
    sim.init_equilibrium()
    sim.color_gradient()
    sim.run(1, 4)
    km = sim.profile(5, args.profile_steps)
    nm = max(1, km["n_odd"] + km["n_even"])
    mixed = {"collide_odd": round(km["collide_odd_ms"] / max(1, km["n_odd"]), 4),
             "collide_even": round(km["collide_even_ms"] / max(1, km["n_even"]), 4),
             "boundary": round(km["boundary_ms"] / nm, 4), "gradient": round(km["gradient_ms"] / nm, 4)}
    log(f"mixed colour field: gradient {mixed['gradient']:.3f} ms, collide odd {mixed['collide_odd']:.3f} / "
        f"even {mixed['collide_even']:.3f} ms")

    probe = None
    if args.probe_mib > 0:
        nbytes = args.probe_mib << 20
        pms = sim.bandwidth_probe(nbytes, repeats=10)
        probe = {"bytes": nbytes, "ms": round(pms, 4), "gbs": round(2 * nbytes / (pms * 1e-3) / 1e9, 1)}
        log(f"plain copy of {args.probe_mib} MiB: {pms:.3f} ms = {probe['gbs']} GB/s (read+write)")
    chain = sim.gradient_path()
    sim.close()

    other = None
    if args.other_steps > 0:
        oprec = gpu.FP32 if prec == gpu.FP64 else gpu.FP64
        log(f"the same workload in {'FP32' if oprec == gpu.FP32 else 'FP64'}: {args.other_steps} steps")
        other = measure_precision(params, geo, n_all, oprec, args.other_steps, args.warmup, args.profile_steps)
        log(f"{other['dtype']}: {other['mlups']:.1f} MLUPS, {other['ms_per_step']:.4f} ms per step, "
            f"collide {other['collide_roofline_frac']:.3f} of the HBM peak")

    cpu = None
    if args.cpu_steps > 0:
        log(f"CPU baseline: {args.cpu_steps} oracle steps")
        cpu = cpu_baseline(geo, args.cpu_steps)
        log(f"CPU baseline {cpu['value']:.2f} MLUPS on {cpu['cores']} threads ({cpu['seconds']} s)")

    coll_traffic = None
    if traffic and "collide_odd" in traffic["requests"] and "collide_even" in traffic["requests"]:
        t = traffic["requests"]
        coll_traffic = round(0.5 * (t["collide_odd"]["total"] + t["collide_even"]["total"]))
    out = {
        "metric": "MLUPS", "value": round(mlups, 3), "unit": "MLUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dt / args.steps * 1e3, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64" if prec == gpu.FP64 else "f32",
        "data": "synthetic: seeded sphere pack",
        "config": {"workload": f"colour-gradient drainage {n}^3, D3Q19x2 MRT, AA streaming, "
                               f"{'FP64' if prec == gpu.FP64 else 'FP32'}",
                   "lattice": [n, n, n], "porosity": round(1 - float(geo.mean()), 4), "fluid_nodes": int(n_fluid)},
        "fluid_mlups": round(mflups, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": coll_traffic,
                     "kernel": "k_collide (mean of the odd and even launches)",
                     "algorithmic_bytes": int(coll_bytes)},
        "step_roofline": {"bytes": int(step_bytes), "ms": round(dt / args.steps * 1e3, 4),
                          "frac": round(step_frac, 4),
                          "definition": "(N_fluid * 78 * s + N_all * 1 B) / timed ms_per_step / 8 TB/s"},
        "fluid_glups": round(n_fluid * args.steps / dt / 1e9, 4),
        "gradient_chain": chain,
        "kernel_ms": {"collide_odd": round(odd_ms, 4), "collide_even": round(even_ms, 4),
                      "boundary": round(bnd_ms, 4), "gradient": round(grad_ms, 4), "step": round(step_ms, 4)},
        "late": late,
        "kernel_ms_mixed_field": mixed,
        "traffic_bytes": traffic,
        "copy_probe": probe,
        "cpu_baseline": cpu,
        "other_precision": other,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
