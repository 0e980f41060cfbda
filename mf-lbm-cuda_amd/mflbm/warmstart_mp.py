"""Learned warm start of a drainage to its steady saturation: the multiphase half of the warm
start (mflbm.warmstart holds the single-phase half).

The reference starts a drainage (initial_fluid_distribution_option 1) with fluid 1 below
interface_z0 only (src/Init_multiphase.cpp:299-358) and ends it with its saturation monitor:
the run is steady once |S(t) - S(t - T)| / S falls below convergence_criteria for a monitor
interval T (steady_state_option 3, src/Monitor.cpp:253-271), S the fluid-1 fraction of the
pore space between the excluded inlet and outlet layers.  Reaching that state is the slow
part of a drainage: the invading fluid creeps through the pores for 10^5 steps and more.

Here a U-Net (mflbm.nn, trained on the device by the project's own kernels) maps the pore
geometry -- solid mask, signed distance to the solid, position along the flow axis -- to
the steady phase field, and the run starts from that field, sharpened to +-1 (the
populations at rest, as the reference's start), which the solver then relaxes to its own
steady state under the same monitor.  The reference's start stays the arbiter (FP64) and
the fallback.

Data: drainages this solver runs from the reference's start to their steady saturation in
FP32, on seeded sphere packs (porosity and radius drawn per seed, empty layers at the inlet
and outlet), with the physics of PHYSICS.  Held-out geometries are other seeds; `--fine`
evaluates on their two-fold refinements (network on the coarsened grid, its field
interpolated back; the pressure drop scaled so that it keeps its ratio to the capillary
entry pressures, which halve with the pore sizes).

Fallback: a warm start whose saturation still changes by more than `fallback_residual` per
step (relative, over the PROBE_WINDOW steps ending at step `fallback_step`) is abandoned for
the reference's start; the steps already taken count against it.

    python -m mflbm.warmstart_mp --n 64 --train 32 --eval 4 --epochs 40
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import cases, config, gpu, nn
from .warmstart import coarsen, refine, signed_distance

BUFFER = 8           # empty layers at the inlet and at the outlet (64^3)
# at 64^3: the wetting of the 256^3 workload, equal viscosities (tau 0.7) and a pressure drop
# that sweeps the pack in ~10^5 steps.  (With fluid 2 ten times more viscous, drops that
# converge in that time drive the invading fluid's channels out of the lattice's stable
# range once it breaks through: 5 of 36 packs blew up, in FP64 as in FP32.)
PHYSICS = dict(theta=30.0, ratio=1.0, gamma=0.01, body_force=4e-3)
PROBE_WINDOW = 100


def drainage_case(seed, n=64, physics=None):
    """Seeded sphere pack (porosity U[0.32, 0.40], radius U[5, 8]) with BUFFER empty layers
    at each end of the flow axis, fluid 1 injected through a Zou-He pressure inlet.  (Packs
    of radius 4.5 under this pressure drop blew up -- the flow through their narrowest throats
    leaves the low-Mach range of the lattice.)"""
    ph = dict(PHYSICS, **(physics or {}))
    rng = np.random.default_rng(20_000 + seed)
    porosity = float(rng.uniform(0.32, 0.40))
    radius = float(rng.uniform(5.0, 8.0))
    geo = cases.sphere_geometry((n, n, n), radius, porosity, seed, buffer=BUFFER)
    return geo, _control(n, BUFFER, ph["body_force"], ph), {"seed": seed, "porosity": porosity, "radius": radius}


def _control(n, buf, body_force, ph):
    nu1 = cases.nu_from_tau(0.7)
    ctrl = cases.drainage_control(n=(n, n, n), theta=ph["theta"], nu1=nu1, nu2=nu1 / ph["ratio"], gamma=ph["gamma"],
                                  body_force=body_force, interface_z0=float(buf))
    ctrl.update(n_exclude_inlet=buf, n_exclude_outlet=buf)
    return config.derive(ctrl)


def fine_case(seed, n_coarse, physics=None):
    """The two-fold refinement of drainage_case(seed, n_coarse).  Pore sizes double, so the
    capillary entry pressures halve: the pressure drop is halved too (the body force that
    sets it, spread over twice the length, divided by four)."""
    ph = dict(PHYSICS, **(physics or {}))
    gc, _, meta = drainage_case(seed, n_coarse, physics)
    n = 2 * n_coarse
    return refine(gc), _control(n, 2 * BUFFER, ph["body_force"] / 4.0, ph), dict(meta, radius=2 * meta["radius"])


def saturation(sim, params):
    """S over the slices n_exclude_inlet + 1 .. nz - n_exclude_outlet (src/Monitor.cpp:88-110)"""
    m = sim.monitor()
    k0, k1 = params["n_exclude_inlet"], params["nz"] - params["n_exclude_outlet"]
    v1, v2 = m["vol1"][k0:k1].sum(), m["vol2"][k0:k1].sum()
    return float(v1 / (v1 + v2))


def start_field(sim, params, phi=None):
    """The reference's initial state (option 1: fluid 1 below interface_z0, inlet ghost
    layers at phi_inlet) with, if given, the interior phase field (nz, ny, nx) replaced;
    populations at rest with the D3Q19 weights (src/Init_multiphase.cpp:389-443)."""
    sim.init_phi(1, params["interface_z0"])
    if phi is not None:
        full = sim.download("phi")
        full[4:-4, 4:-4, 4:-4] = phi
        sim.upload("phi", full)
    sim.init_equilibrium()
    sim.color_gradient()


def run_to_steady_saturation(sim, params, phi=None, tol=1e-4, check=2000, max_steps=400000, probe=None, log=None,
                             info=None):
    """Run from start_field(phi) until |S(t) - S(t - check)| / S < tol.  probe = (step,
    residual): give up at `step`, returning (steps, None), if the relative saturation
    change per step over the PROBE_WINDOW steps ending there exceeds `residual`.
    Returns (steps, S); S is NaN for a run that blew up (it stops at the first NaN)."""
    start_field(sim, params, phi)
    t = 1
    if probe:
        pre = probe[0] - PROBE_WINDOW
        sim.run(1, pre)
        s0 = saturation(sim, params)
        sim.run(pre + 1, PROBE_WINDOW)
        s1 = saturation(sim, params)
        residual = abs(s1 - s0) / max(s1, 1e-12) / PROBE_WINDOW
        if info is not None:
            info["residual"] = residual
        if log:
            log(f"  step {probe[0]}: saturation {s1:.5f}, residual {residual:.3e} per step")
        if residual > probe[1]:
            return probe[0], None
        t = probe[0] + 1
    prev = saturation(sim, params)
    while t <= max_steps:
        sim.run(t, check)
        t += check
        s = saturation(sim, params)
        if log and (t - 1) % (25 * check) == 0:
            log(f"  step {t - 1}: saturation {s:.5f}")
        if not np.isfinite(s):
            return t - 1, s
        if abs(s - prev) / max(s, 1e-12) < tol:
            return t - 1, s
        prev = s
    return t - 1, prev


def fluid1_flux(sim):
    """fluid-1 axial flux summed over the slices (src/Monitor.cpp:60-67): the steady
    relative permeability of the invading phase up to the constant driving factor"""
    return float(np.sum(sim.monitor()["fl1"]))


def steady_field(geo, params, precision=gpu.FP32, **kw):
    """(steps, S, interior phi (nz, ny, nx) float32) of the reference's start run to steady"""
    sim = gpu.Solver(params, geo, precision)
    steps, s = run_to_steady_saturation(sim, params, **kw)
    phi = sim.download("phi")[4:-4, 4:-4, 4:-4].astype(np.float32)
    sim.close()
    return steps, s, phi


def generate(seeds, n=64, log=print, **kw):
    """[(geometry, steady phi, params, meta)]: the reference's start of each seed run to its
    steady saturation (FP32), one solve after the other on one host thread."""
    data = []
    for seed in seeds:
        geo, params, meta = drainage_case(seed, n)
        t0 = time.perf_counter()
        steps, s, phi = steady_field(geo, params, log=lambda m: log(f"sample {seed}:{m}"), **kw)
        meta.update(steps=steps, saturation=s, seconds=time.perf_counter() - t0)
        log(f"sample {seed}: steady S {s:.4f} after {steps} steps ({meta['seconds']:.1f} s)")
        if np.isfinite(s) and np.isfinite(phi).all():
            data.append((geo, phi, params, meta))
        else:
            log(f"sample {seed} dropped: the run blew up")
    return data


def solid_of(geo):
    """the solver's solid nodes: the geometry plus its x-wall planes"""
    s = geo.astype(bool).copy()
    s[:, :, 0] = True
    s[:, :, -1] = True
    return s


def features(geo, channels=8):
    """network input: solid mask, signed distance / 8 (clipped), position along the flow
    axis (0 .. 1), the inlet buffer"""
    solid = solid_of(geo)
    nz = geo.shape[0]
    x = np.zeros(geo.shape + (channels,), np.float32)
    x[..., 0] = solid
    x[..., 1] = np.clip(signed_distance(solid) / 8.0, -1.0, 1.0)
    x[..., 2] = ((np.arange(nz) + 0.5) / nz)[:, None, None]
    x[:BUFFER, :, :, 3] = 1.0
    return x


class PhaseWarmStart:
    """U-Net from the pore geometry to the steady phase field of the drainage."""

    def __init__(self, widths=(16, 32, 64, 128), seed=0):
        self.widths = tuple(widths)
        self.net = nn.UNet(widths=widths, cin=8, cout=8, n_out=1, seed=seed)

    def save(self, path):
        np.savez(path, widths=np.array(self.widths), **self.net.state())

    @classmethod
    def load(cls, path):
        with np.load(path) as z:
            ws = cls(widths=tuple(int(c) for c in z["widths"]))
            ws.net.load_state({k: z[k] for k in z.files if k.endswith(".w") or k.endswith(".b")})
        return ws

    def _tensors(self, geo, phi=None):
        x = nn.DevArray.from_numpy(features(geo), "bf16")
        if phi is None:
            return x, None, None
        t = np.zeros(geo.shape + (8,), np.float32)
        t[..., 0] = phi
        mask = (~solid_of(geo)).astype(np.float32)
        return x, nn.DevArray.from_numpy(t, "f32"), nn.DevArray.from_numpy(mask, "f32")

    def train(self, data, epochs=40, lr=1e-3, log=print):
        tensors = [self._tensors(g, phi) for g, phi, _, _ in data]
        hist = []
        rng = np.random.default_rng(0)
        for ep in range(epochs):
            tot = 0.0
            for i in rng.permutation(len(tensors)):
                x, t, m = tensors[i]
                a = self.net.forward(x)
                part, inv, gy = self.net.loss_grad(a["y"], t, m)
                self.net.backward(a, gy)
                self.net.adam(lr)
                tot += float(part.numpy().sum()) * inv
            hist.append(tot / len(tensors))
            log(f"epoch {ep + 1}: loss {hist[-1]:.4e}")
        return hist

    def raw(self, geo):
        x, _, _ = self._tensors(geo)
        return self.net.forward(x)["y"].numpy()[..., 0]

    @staticmethod
    def sharpen(y, geo):
        """+-1 field from the network's output y: fluid 1 at the pore nodes of the highest y,
        as many as the output's own saturation estimate mean((1 + clip(y)) / 2) implies
        (thresholding at 0 would give every uncertain node to the majority phase), -1 at
        solid nodes"""
        pore = ~solid_of(geo)
        yp = y[pore]
        s_est = float(np.mean((1.0 + np.clip(yp, -1.0, 1.0)) * 0.5))
        k = int(round(s_est * yp.size))
        thr = np.inf if k <= 0 else np.partition(yp, yp.size - k)[yp.size - k]
        return np.where(pore & (y >= thr), 1.0, -1.0)

    def predict(self, geo):
        """the phase field to start from: the network's, sharpened"""
        return self.sharpen(self.raw(geo), geo)

    def predict_fine(self, geo):
        """the phase field of a geometry twice as fine as the network's grid: the network on
        the coarsened geometry, its field interpolated back (trilinear), then sharpened"""
        from scipy import ndimage
        y = ndimage.zoom(self.raw(coarsen(geo)), 2, order=1, grid_mode=True, mode="nearest")
        return self.sharpen(y, geo)


def field_agreement(phi_a, phi_b, geo):
    """fraction of the pore nodes whose phases (sign of phi) agree"""
    pore = ~solid_of(geo)
    return float(((phi_a > 0) == (phi_b > 0))[pore].mean())


def _case_of(s, n, fine):
    return fine_case(s, n) if fine else drainage_case(s, n)


def evaluate_cold(s, n=64, fine=False, tol=1e-4, check=2000, max_steps=400000, log=print):
    """(row, phase sign (nz, ny, nx) int8) of held-out seed s: the reference's start in FP64,
    the arbiter, run to the steady saturation."""
    geo, params, meta = _case_of(s, n, fine)
    row = dict(meta, n=geo.shape[0])
    sim = gpu.Solver(params, geo, gpu.FP64)
    t0 = time.perf_counter()
    row["cold_steps"], row["S_cold"] = run_to_steady_saturation(sim, params, tol=tol, check=check,
                                                                max_steps=max_steps, log=log)
    row["cold_s"] = time.perf_counter() - t0
    row["fl1_cold"] = fluid1_flux(sim)
    sign = np.sign(sim.download("phi")[4:-4, 4:-4, 4:-4]).astype(np.int8)
    sim.close()
    log(f"held-out {s} ({geo.shape[0]}^3): FP64 reference start {row['cold_steps']} steps "
        f"{row['cold_s']:.1f} s, S {row['S_cold']:.5f}")
    if not np.isfinite(row["S_cold"]):
        row["blew_up"] = True
    return row, sign


def evaluate_rest(ws, row, cold_sign, fine=False, precision=gpu.FP32, tol=1e-4, check=2000, max_steps=400000,
                  fallback_step=500, fallback_residual=1e-4, log=print):
    """Completes a row of evaluate_cold: the same start in `precision` (the method without
    the network) and the learned start in `precision` with the fallback rule."""
    s = row["seed"]
    geo, params, _ = _case_of(s, row["n"] // 2 if fine else row["n"], fine)
    kw = dict(tol=tol, check=check, max_steps=max_steps)
    sim = gpu.Solver(params, geo, precision)
    t0 = time.perf_counter()
    row["plain_steps"], row["S_plain"] = run_to_steady_saturation(sim, params, **kw)
    row["plain_s"] = time.perf_counter() - t0
    row["fl1_plain"] = fluid1_flux(sim)
    t0 = time.perf_counter()
    phi0 = ws.predict_fine(geo) if fine else ws.predict(geo)
    row["predict_s"] = time.perf_counter() - t0
    row["predicted_phase_agreement"] = field_agreement(phi0, cold_sign, geo)
    info = {}
    steps, sat = run_to_steady_saturation(sim, params, phi0, probe=(fallback_step, fallback_residual), info=info,
                                          **kw)
    row["warm_residual_at_probe"] = info.get("residual")
    row["fallback"] = sat is None
    if sat is None:
        lost = steps
        steps, sat = run_to_steady_saturation(sim, params, **kw)
        steps += lost
    row["warm_s"] = time.perf_counter() - t0
    row["warm_steps"], row["S_warm"] = steps, sat
    row["fl1_warm"] = fluid1_flux(sim)
    sim.close()
    for k in ("plain", "warm"):
        row[f"S_{k}_rel_diff"] = abs(row[f"S_{k}"] - row["S_cold"]) / row["S_cold"]
        row[f"fl1_{k}_rel_diff"] = abs(row[f"fl1_{k}"] - row["fl1_cold"]) / abs(row["fl1_cold"])
    log(f"held-out {s}: FP32 reference start {row['plain_steps']} steps {row['plain_s']:.1f} s, "
        f"S {row['S_plain']:.5f}; learned start {row['warm_steps']} steps {row['warm_s']:.1f} s, "
        f"S {row['S_warm']:.5f} (fallback {row['fallback']}, phases {row['predicted_phase_agreement']:.3f} "
        f"as the FP64 steady state)")
    return row


def evaluate(ws, seeds, n=64, fine=False, precision=gpu.FP32, tol=1e-4, check=2000, max_steps=400000,
             fallback_step=500, fallback_residual=1e-4, log=print):
    """Per held-out geometry: the reference's start in FP64 (the arbiter), the same start in
    `precision` (the method without the network) and the learned start in `precision` with
    the fallback rule, each run to the steady saturation."""
    rows = []
    for s in seeds:
        row, sign = evaluate_cold(s, n, fine, tol, check, max_steps, log)
        if not row.get("blew_up"):
            evaluate_rest(ws, row, sign, fine, precision, tol, check, max_steps, fallback_step, fallback_residual, log)
        rows.append(row)
    return rows


def summarise(rows):
    rows = [r for r in rows if not r.get("blew_up")]
    if not rows:
        return {"blew_up": True}
    tot = {k: float(sum(r[k] for r in rows)) for k in ("cold_steps", "plain_steps", "warm_steps", "cold_s", "plain_s",
                                                        "warm_s")}
    return {"warm_vs_cold_wall": tot["warm_s"] / tot["cold_s"], "plain_vs_cold_wall": tot["plain_s"] / tot["cold_s"],
            "warm_vs_plain_steps": tot["warm_steps"] / tot["plain_steps"],
            "warm_vs_plain_wall": tot["warm_s"] / tot["plain_s"],
            "max_S_warm_rel_diff": float(max(r["S_warm_rel_diff"] for r in rows)),
            "max_S_plain_rel_diff": float(max(r["S_plain_rel_diff"] for r in rows)),
            "max_fl1_warm_rel_diff": float(max(r["fl1_warm_rel_diff"] for r in rows)),
            "max_fl1_plain_rel_diff": float(max(r["fl1_plain_rel_diff"] for r in rows)),
            "fallbacks": int(sum(r["fallback"] for r in rows))}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--train", type=int, default=32)
    ap.add_argument("--eval", type=int, default=4)
    ap.add_argument("--first-eval-seed", type=int, default=1000, help="held-out seeds start here")
    ap.add_argument("--epochs", type=int, default=40)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--fine", action="store_true", help="evaluate on geometries twice as fine as --n")
    ap.add_argument("--tol", type=float, default=1e-4)
    ap.add_argument("--check", type=int, default=2000)
    ap.add_argument("--save", default=None)
    ap.add_argument("--load", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--max-steps", type=int, default=400000)
    ap.add_argument("--stage", choices=["all", "cold", "rest"], default="all",
                    help="cold: only the FP64 reference starts, saved to --out and <out>.phase<seed>; "
                         "rest: the other starts from those files (--cold), so that a large evaluation "
                         "can run as two jobs")
    ap.add_argument("--cold", default=None, help="--stage rest: the --out file of the cold stage")
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    res = {"n": args.n, "physics": PHYSICS, "tol": args.tol, "check": args.check}
    check = 2 * args.check if args.fine else args.check
    seeds = range(args.first_eval_seed, args.first_eval_seed + args.eval)
    if args.stage == "cold":
        rows = []
        for s in seeds:
            row, sign = evaluate_cold(s, args.n, args.fine, args.tol, check, args.max_steps, log)
            with open(f"{args.out}.phase{s}", "wb") as fh:  # a file object: no suffix appended
                np.savez_compressed(fh, sign=sign)
            rows.append(row)
        res.update(stage="cold", eval=rows)
        _write(res, args.out)
        return
    if args.load:
        ws = PhaseWarmStart.load(args.load)
    else:
        t0 = time.perf_counter()
        data = generate(range(args.train), args.n, log, tol=args.tol, check=args.check)
        res["data_seconds"] = time.perf_counter() - t0
        res["train_steady_steps_mean"] = float(np.mean([m["steps"] for _, _, _, m in data]))
        ws = PhaseWarmStart()
        t0 = time.perf_counter()
        hist = ws.train(data, args.epochs, args.lr, log)
        res.update(train_seconds=time.perf_counter() - t0, loss_first=hist[0], loss_last=hist[-1],
                   params=ws.net.n_params())
        res["train_phase_agreement"] = [field_agreement(ws.predict(g), phi, g) for g, phi, _, _ in data[:4]]
        if args.save:
            ws.save(args.save)
        del data
    if args.stage == "rest":
        with open(args.cold) as fh:
            rows = json.loads(fh.read())["eval"]
        for row in rows:
            if not row.get("blew_up"):
                with np.load(f"{args.cold}.phase{row['seed']}") as z:
                    sign = z["sign"]
                evaluate_rest(ws, row, sign, args.fine, tol=args.tol, check=check, max_steps=args.max_steps, log=log)
    else:
        rows = evaluate(ws, seeds, args.n, args.fine, tol=args.tol, check=check, max_steps=args.max_steps, log=log)
    res.update(summarise(rows), fine=args.fine, eval=rows)
    _write(res, args.out)


def _write(res, path):
    text = json.dumps(res)
    if path:
        with open(path, "w") as fh:
            fh.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
