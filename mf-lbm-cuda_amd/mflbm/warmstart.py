"""Learned warm start of the flow solve.

The reference starts every run from rest (u = 0, src/Init_multiphase.cpp:569-572) and
builds the populations from (rho, u, v, w, phi) with the D3Q19 weights 1/3, 1/18, 1/36
(src/Init_multiphase.cpp:389-443).  Here the start has two parts:

1. Relaxed initialisation (mflbm_init_relaxed).  The MRT collision relaxes the fourth-order
   moment e2 to 0, so its rest state is rho (4/21, 13/126, 1/63), not the D3Q19 weights
   (e2 = 3 rho).  At fluid nodes the first collision removes the difference locally; in the
   bounce-back links it returns to the fluid as a mass pulse of about -0.048 (axis links)
   and +0.012 (diagonal links) per wall link, which then has to spread through the pore
   space by pressure diffusion -- the slowest process of a single-phase start.  Starting
   at the collision's own rest weights removes the pulse.  The mean density is set from
   the cold start's total mass (relaxed_mean_density): the same amount of fluid, so the
   same steady state.
2. Learned fields.  A 3-D convolutional U-Net (mflbm.nn, trained on the device by the
   project's own kernels) maps the pore geometry -- solid mask and signed distance to the
   solid -- to the steady velocity and pressure (density) fields, which enter through the
   same initialisation.  The predicted density perturbation is shifted to zero mean over the
   fluid, so the mass stays that of the cold start.  With Guo forcing the populations of the
   first step are pre-collision values whose momentum is rho u - F / 2, so F / 2 is taken
   off the predicted velocity.

The cold-start solver stays the arbiter: every start runs to the same steady criterion.  `evaluate`
reports, per held-out geometry, the reference's cold start (FP64), the relaxed start from
rest (the method without the network) and the learned warm start, the last two in the
chosen precision (FP32 by default).

Data: single-phase body-force-driven flow (phi = 1 everywhere, body force 1e-6 along the flow
axis, periodic along flow and one transverse axis, walls on the other) through seeded sphere
packs, porosity drawn from [0.30, 0.50] and radius from [4, 10]; the samples are run to their
steady states together by the solver's batched ensemble.  Targets are the velocities scaled by nu / F (a length
squared, independent of the driving force) and the density perturbation scaled by 1 / (3 F)
(a length), each normalised by one constant of the training set.

Steady criterion: the relative L2 change of the velocity field over `check` steps falls
below `tol` (default 1e-5 per 100 steps, about 1e-7 per step).

Fine grids (`--fine`): the network runs on the geometry coarsened by two (solid where at
least half of a 2x2x2 block is solid) and its fields are interpolated back, the velocity
times 4 and the density perturbation times 2 (the targets are a length squared and a length
in lattice units, and the lattice unit halves).  The held-out fine geometries are the
two-fold refinements of held-out coarse packs.  A warm start whose velocity still changes by
more than `fallback_residual` per step (relative L2, over the 100 steps ending at step
`fallback_step`) is abandoned for the relaxed start from rest (the method without the
network), and the steps already taken count against it.

    python -m mflbm.warmstart --train 48 --eval 8 --epochs 30
    python -m mflbm.warmstart --n 128 --fine --train 96 --eval 2 --epochs 20
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import cases, config, geometry, gpu, nn

BODY_FORCE = 1e-6


def single_phase_case(seed, n=64):
    rng = np.random.default_rng(10_000 + seed)
    porosity = float(rng.uniform(0.30, 0.50))
    radius = float(rng.uniform(4.0, 10.0))
    geo = geometry.sphere_pack((n, n, n), radius, porosity, seed, periodic=(True, True, False))
    ctrl = cases.periodic_control(n=(n, n, n), body_force=BODY_FORCE, option=1, interface_z0=float(n + 10))
    return geo, config.derive(ctrl), {"seed": seed, "porosity": porosity, "radius": radius}


def solid_with_walls(geo):
    """the solver's solid nodes: the geometry plus the x-wall planes it adds"""
    s = geo.astype(bool).copy()
    s[:, :, 0] = True
    s[:, :, -1] = True
    return s


def signed_distance(solid):
    """distance to the nearest solid node in the pore (> 0), minus the distance to the
    nearest pore node in the solid; periodic images along z and y are included."""
    from scipy import ndimage
    pad = ((16, 16), (16, 16), (0, 0))
    s = np.pad(solid.astype(bool), pad, mode="wrap")
    d_out = ndimage.distance_transform_edt(~s)
    d_in = ndimage.distance_transform_edt(s)
    sd = (d_out - d_in)[16:-16, 16:-16, :]
    return sd


def features(solid, channels=8):
    """network input: channel 0 = solid mask, channel 1 = signed distance / 8 (clipped)"""
    x = np.zeros(solid.shape + (channels,), np.float32)
    x[..., 0] = solid
    x[..., 1] = np.clip(signed_distance(solid) / 8.0, -1.0, 1.0)
    return x


D3Q19 = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 1, 0), (1, -1, 0),
         (-1, -1, 0), (1, 0, 1), (-1, 0, 1), (1, 0, -1), (-1, 0, -1), (0, 1, 1), (0, -1, 1), (0, 1, -1), (0, -1, -1)]


def relaxed_mean_density(geo):
    """Mean fluid density of the steady state the reference's cold start reaches, for a
    start at the collision's rest weights.  Mass is conserved: the cold start holds
    N_fluid * 1 in the fluid plus the D3Q19 weight of each bounce-back link (the population
    a solid neighbour returns on the second step); at rest the same links hold rho times
    the relaxed weights.  So rho = (N + sum_links w) / (N + sum_links r)."""
    solid = solid_with_walls(geo)
    fluid = ~solid
    nf = int(fluid.sum())
    w_std = r_rel = 0.0
    for e in D3Q19:
        # fluid x whose neighbour x - e is solid: one live bounce-back slot
        links = int((fluid & np.roll(solid, shift=(e[2], e[1], e[0]), axis=(0, 1, 2))).sum())
        axis = sum(abs(c) for c in e) == 1
        w_std += links * (1.0 / 18.0 if axis else 1.0 / 36.0)
        r_rel += links * (13.0 / 126.0 if axis else 1.0 / 63.0)
    return (nf + w_std) / (nf + r_rel)


PROBE_WINDOW = 100


def run_to_steady(sim, params, init=None, relaxed=False, tol=1e-5, check=100, max_steps=60000, probe=None,
                  log=None, info=None):
    """Initialise from init = (rho, u, v, w) on the s1 grid (None = the reference's rest
    state) with the D3Q19 or, relaxed, the collision's rest weights, and step until the
    relative L2 change of the velocity over `check` steps is below tol.
    probe = (step, residual): give up, returning (steps, None, None), if the residual at
    `step` -- the relative velocity change per step over the PROBE_WINDOW steps ending
    there -- is above `residual` (stored in info["residual"]).  The steady test resumes
    from the probe step.
    Returns (steps, velocity (3, nz, ny, nx), density (nz, ny, nx))."""
    sim.init_phi(params["initial_fluid_distribution_option"], params["interface_z0"])
    sim.init_equilibrium(*(init or (None, None, None, None)), relaxed=relaxed)
    sim.color_gradient()
    sim.velocity_change()  # the change is measured on the device from here on
    t = 1
    while t <= max_steps:
        if probe and t - 1 < probe[0] <= t - 1 + check and probe[0] - PROBE_WINDOW >= t - 1:
            pre = probe[0] - PROBE_WINDOW - (t - 1)
            if pre:
                sim.run(t, pre)
            sim.velocity_change()
            sim.run(t + pre, PROBE_WINDOW)
            residual = sim.velocity_change() / PROBE_WINDOW
            t = probe[0] + 1
            if info is not None:
                info["residual"] = residual
            if log:
                log(f"  step {probe[0]}: residual {residual:.3e} per step")
            if residual > probe[1]:
                return t - 1, None, None
            continue
        sim.run(t, check)
        t += check
        change = sim.velocity_change()
        if log and (t - 1) % (50 * check) == 0:
            log(f"  step {t - 1}: velocity change {change:.3e}")
        if probe and t - 1 == probe[0] and change > probe[1]:
            return t - 1, None, None
        if change < tol:
            break
    rho, u, v, w = sim.macro()
    vel = np.stack([u[1:-1, 1:-1, 1:-1], v[1:-1, 1:-1, 1:-1], w[1:-1, 1:-1, 1:-1]])
    return t - 1, vel, rho[1:-1, 1:-1, 1:-1]


def permeability(vel, params):
    """k = nu <w> / F (Darcy velocity over the whole cross-section)"""
    nu = 1.0 / params["la_nui1"]
    return float(vel[2].mean() * nu / BODY_FORCE)


def coarsen(geo):
    """solid where at least half of each 2x2x2 block is solid (exact for refine(g))"""
    nz, ny, nx = geo.shape
    b = geo.astype(np.int16).reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).sum(axis=(1, 3, 5))
    return (b >= 4).astype(np.int8)


def refine(geo):
    """each node becomes a 2x2x2 block"""
    return np.asarray(geo, np.int8).repeat(2, 0).repeat(2, 1).repeat(2, 2)


def fine_case(seed, n_coarse):
    """the two-fold refinement of single_phase_case(seed, n_coarse), same body force"""
    gc, _, meta = single_phase_case(seed, n_coarse)
    n = 2 * n_coarse
    ctrl = cases.periodic_control(n=(n, n, n), body_force=BODY_FORCE, option=1, interface_z0=float(n + 10))
    return refine(gc), config.derive(ctrl), dict(meta, radius=2 * meta["radius"])


def to_s1(field):
    """interior (nz, ny, nx) -> s1 grid with a periodic one-node halo"""
    return np.pad(np.asarray(field, np.float64), 1, mode="wrap")


def rest_state(geo):
    """(rho, u, v, w) of the relaxed start from rest"""
    rho = np.full(tuple(s + 2 for s in geo.shape), relaxed_mean_density(geo))
    return rho, None, None, None


def generate(seeds, n=64, log=print, chunk=512):
    """[(geometry, target (4, nz, ny, nx), params, meta)] with target = velocity * nu / F and
    the zero-mean density perturbation / (3 F) over the fluid.  The steady fields come from
    the batched ensemble (mflbm.ensemble: all samples of a chunk stepped together on the
    device from the relaxed start)."""
    from . import ensemble
    seeds = list(seeds)
    data = []
    for c0 in range(0, len(seeds), chunk):
        part = seeds[c0:c0 + chunk]
        t0 = time.perf_counter()
        geos, metas = ensemble.sphere_pack_ensemble(part, n)
        _, params, _ = single_phase_case(part[0], n)
        ens = ensemble.Ensemble(params, geos)
        for taken in range(0, 60000, 2000):  # in slices, to report progress
            out = ens.solve(check=100, tol=1e-5, max_steps=2000)
            if (out["steps"] > 0).all():
                break
            log(f"  {taken + 2000} steps: {(out['steps'] > 0).sum()} of {len(part)} domains steady")
        fields = ens.fields()
        rho_ref = ens.rho_ref.copy()
        ens.close()
        nu = 1.0 / params["la_nui1"]
        for d, (geo, meta) in enumerate(zip(geos, metas)):
            fluid = ~solid_with_walls(geo)
            vel = fields[d, :3].astype(np.float64)
            rho = rho_ref[d] + fields[d, 3].astype(np.float64)
            drho = np.where(fluid, rho - rho[fluid].mean(), 0.0)
            meta.update(steps=int(out["steps"][d]), k=float(out["permeability"][d]))
            target = np.concatenate([vel * nu / BODY_FORCE, (drho / (3.0 * BODY_FORCE))[None]]).astype(np.float32)
            data.append((geo, target, params, meta))
        log(f"samples {part[0]}..{part[-1]}: steady in {np.abs(out['steps']).mean():.0f} steps on average, "
            f"{time.perf_counter() - t0:.1f} s")
    return data


class WarmStart:
    """rho_weight: weight of the density channel in the loss relative to the velocity ones
    (the pressure field is what shortens the start: on the FP64 oracle the exact steady
    velocity alone gains nothing over the relaxed start, the exact density alone most of
    what both give)."""

    def __init__(self, widths=(16, 32, 64, 128), seed=0, rho_weight=16.0):
        self.widths = tuple(widths)
        self.net = nn.UNet(widths=widths, cin=8, cout=8, n_out=4, seed=seed)
        self.scale = np.ones(4)
        self.rho_weight = rho_weight

    def save(self, path):
        np.savez(path, scale=self.scale, widths=np.array(self.widths), rho_weight=self.rho_weight, **self.net.state())

    @classmethod
    def load(cls, path):
        with np.load(path) as z:
            ws = cls(widths=tuple(int(c) for c in z["widths"]), rho_weight=float(z["rho_weight"]))
            ws.scale = z["scale"].copy()
            ws.net.load_state({k: z[k] for k in z.files if k.endswith(".w") or k.endswith(".b")})
        return ws

    def _tensors(self, geo, target=None):
        x = nn.DevArray.from_numpy(features(geo), "bf16")
        if target is None:
            return x, None, None
        t = np.zeros(geo.shape + (8,), np.float32)
        t[..., :4] = np.moveaxis(target, 0, -1) / self.scale
        mask = (~solid_with_walls(geo)).astype(np.float32)
        return x, nn.DevArray.from_numpy(t, "f32"), nn.DevArray.from_numpy(mask, "f32")

    def train(self, data, epochs=30, lr=1e-3, log=print):
        fl = [~solid_with_walls(g) for g, _, _, _ in data]
        self.scale = np.array([np.mean([np.abs(t[c][f]).mean() for (_, t, _, _), f in zip(data, fl)])
                               for c in range(3)] + [np.mean([t[3][f].std() for (_, t, _, _), f in zip(data, fl)])])
        self.scale[:3] = self.scale[2]  # one velocity scale, set by the flow-direction mean
        self.scale[3] /= np.sqrt(self.rho_weight)
        self.scale = np.maximum(self.scale, 1e-12)
        tensors = []
        for i, (g, t, _, _) in enumerate(data):
            tensors.append(self._tensors(g, t))
            if (i + 1) % 32 == 0:
                log(f"  network inputs: {i + 1} of {len(data)}")
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

    def predict(self, geo, params):
        """(rho, u, v, w) on the s1 grid for mflbm_init_relaxed"""
        x, _, _ = self._tensors(geo)
        y = self.net.forward(x)["y"].numpy()
        nu = 1.0 / params["la_nui1"]
        fluid = ~solid_with_walls(geo)
        vel = [np.where(fluid, y[..., c] * self.scale[c] * BODY_FORCE / nu, 0.0) for c in range(3)]
        vel[2] = vel[2] - 0.5 * BODY_FORCE * fluid  # pre-collision momentum of the first step
        drho = y[..., 3] * self.scale[3] * 3.0 * BODY_FORCE
        rbar = relaxed_mean_density(geo)
        rho = np.where(fluid, rbar + drho - drho[fluid].mean(), rbar)
        return tuple(to_s1(a) for a in (rho, *vel))


    def predict_fine(self, geo, params):
        """(rho, u, v, w) on the s1 grid of a geometry twice as fine as the network's grid"""
        from scipy import ndimage
        y = self.net.forward(self._tensors(coarsen(geo))[0])["y"].numpy()
        up = lambda a: ndimage.zoom(a, 2, order=1, grid_mode=True, mode="nearest")
        nu = 1.0 / params["la_nui1"]
        fluid = ~solid_with_walls(geo)
        vel = [np.where(fluid, up(y[..., c]) * (4.0 * self.scale[c] * BODY_FORCE / nu), 0.0) for c in range(3)]
        vel[2] = vel[2] - 0.5 * BODY_FORCE * fluid
        drho = up(y[..., 3]) * (2.0 * self.scale[3] * 3.0 * BODY_FORCE)
        rbar = relaxed_mean_density(geo)
        rho = np.where(fluid, rbar + drho - drho[fluid].mean(), rbar)
        return tuple(to_s1(a) for a in (rho, *vel))


def ensemble_run(params, geo, fields=None, tol=5e-5, check=500, max_steps=400000, probe=None, info=None):
    """One domain through the single-phase ensemble kernel (FP32 deviations), from rest or
    from fields (4, nz, ny, nx) = u, v, w, rho - rho_ref; probe as in run_to_steady.
    Returns (steps, permeability or None after a failed probe, wall seconds without set-up)."""
    from . import ensemble
    ens = ensemble.Ensemble(params, geo[None])
    t0 = time.perf_counter()
    taken = 0
    if fields is not None:
        ens.init_fields(fields)
    if probe:
        pre = probe[0] - PROBE_WINDOW
        if pre > 0:
            ens.solve(check=pre, tol=-1.0, max_steps=pre)
        out = ens.solve(check=PROBE_WINDOW, tol=-1.0, max_steps=PROBE_WINDOW)
        taken = probe[0]
        residual = float(out["change"][0]) / PROBE_WINDOW
        if info is not None:
            info["residual"] = residual
        if residual > probe[1]:
            ens.close()
            return taken, None, time.perf_counter() - t0
    out = ens.solve(check=check, tol=tol, max_steps=max_steps - taken)
    seconds = time.perf_counter() - t0
    ens.close()
    return abs(int(out["steps"][0])), float(out["permeability"][0]), seconds


def coarse_start(geo, n_coarse, seed, tol=5e-5, check=500):
    """Fields (4, nz, ny, nx) for ensemble_run from the steady state of the geometry coarsened
    by two, solved from rest by the ensemble kernel and interpolated back with the same
    length scaling as the network's prediction (velocity x4, density perturbation x2)."""
    from scipy import ndimage
    from . import ensemble
    gc = coarsen(geo)
    _, pc, _ = single_phase_case(seed, n_coarse)
    ens = ensemble.Ensemble(pc, gc[None])
    ens.solve(check=check, tol=tol, max_steps=400000)
    fc = ens.fields()[0].astype(np.float64)
    ens.close()
    up = lambda a: ndimage.zoom(a, 2, order=1, grid_mode=True, mode="nearest")
    fluid = ~solid_with_walls(geo)
    u = [np.where(fluid, up(fc[c]) * 4.0, 0.0) for c in range(3)]
    d = up(fc[3]) * 2.0
    d = np.where(fluid, d - d[fluid].mean(), 0.0)
    return np.stack(u + [d]).astype(np.float32)


def evaluate_fine(ws, seeds, n_coarse=128, precision=gpu.FP32, log=print, check=500, tol=5e-5,
                  cold_max_steps=400000, fallback_step=500, fallback_residual=1e-4, starts=("cold", "relaxed", "warm")):
    """Held-out geometries twice as fine as the network's grid: the reference's FP64 cold
    start (the arbiter), the relaxed start from rest and the learned warm start in `precision`,
    with the fallback rule on the warm start."""
    rows = []
    for s in seeds:
        geo, params, meta = fine_case(s, n_coarse)
        row = dict(meta, n=geo.shape[0])
        fluid = ~solid_with_walls(geo)
        ref = None
        if "cold" in starts:
            sim = gpu.Solver(params, geo, gpu.FP64)
            t0 = time.perf_counter()
            row["cold_steps"], cold, rho_cold = run_to_steady(sim, params, tol=tol, check=check,
                                                              max_steps=cold_max_steps, log=log)
            row["cold_s"] = time.perf_counter() - t0
            sim.close()
            row["k_cold"] = permeability(cold, params)
            ref = (cold, rho_cold)
            log(f"held-out {s} ({geo.shape[0]}^3): cold {row['cold_steps']} steps {row['cold_s']:.1f} s, "
                f"k {row['k_cold']:.5f}")
        sim = gpu.Solver(params, geo, precision)
        for name in [n for n in ("relaxed", "warm") if n in starts]:
            t0 = time.perf_counter()
            if name == "warm":
                init = ws.predict_fine(geo, params)
                row["predict_s"] = time.perf_counter() - t0
                if ref is not None:
                    pr = [a[1:-1, 1:-1, 1:-1] for a in init]
                    vel_p = np.stack(pr[1:])
                    vel_p[2] += 0.5 * BODY_FORCE * fluid
                    row["velocity_rel_err"] = float(np.linalg.norm((vel_p - ref[0])[:, fluid]) /
                                                    np.linalg.norm(ref[0][:, fluid]))
                    d_ref = ref[1][fluid] - ref[1][fluid].mean()
                    d_pre = pr[0][fluid] - pr[0][fluid].mean()
                    row["density_rel_err"] = float(np.linalg.norm(d_pre - d_ref) / max(np.linalg.norm(d_ref), 1e-300))
                    del pr, vel_p
                info = {}
                steps, vel, _ = run_to_steady(sim, params, init, relaxed=True, tol=tol, check=check,
                                              max_steps=cold_max_steps, probe=(fallback_step, fallback_residual), log=log,
                                              info=info)
                row["residual_at_probe"] = info.get("residual")
                row["fallback"] = vel is None
                if vel is None:
                    lost = steps
                    steps, vel, _ = run_to_steady(sim, params, rest_state(geo), relaxed=True, tol=tol, check=check,
                                                  max_steps=cold_max_steps, log=log)
                    steps += lost
            else:
                steps, vel, _ = run_to_steady(sim, params, rest_state(geo), relaxed=True, tol=tol, check=check,
                                              max_steps=cold_max_steps, log=log)
            row[f"{name}_steps"] = steps
            row[f"{name}_s"] = time.perf_counter() - t0
            row[f"k_{name}"] = permeability(vel, params)
            if ref is not None:
                row[f"k_{name}_rel_diff"] = abs(row[f"k_{name}"] - row["k_cold"]) / abs(row["k_cold"])
            log(f"held-out {s}: {name} {steps} steps {row[f'{name}_s']:.1f} s, k {row[f'k_{name}']:.5f}"
                + (f", fallback {row['fallback']}" if name == "warm" else ""))
        sim.close()
        # the same starts through the single-phase ensemble kernel (phi = 1 throughout), and
        # a start from the steady state of the coarse geometry (solved by the same kernel)
        for name in [n for n in ("relaxed_ens", "warm_ens", "coarse_ens") if n in starts]:
            if name == "coarse_ens":
                t0 = time.perf_counter()
                fields = coarse_start(geo, n_coarse, s)
                t_pre = time.perf_counter() - t0
                info = {}
                steps, k, sec = ensemble_run(params, geo, fields, tol, check, cold_max_steps,
                                             probe=(fallback_step, fallback_residual), info=info)
                row["coarse_residual_at_probe"] = info.get("residual")
                row["coarse_fallback"] = k is None
                if k is None:
                    lost, lost_s = steps, sec
                    steps, k, sec = ensemble_run(params, geo, None, tol, check, cold_max_steps)
                    steps, sec = steps + lost, sec + lost_s
                row["coarse_prepare_s"] = t_pre
                sec += t_pre
            elif name == "warm_ens":
                t0 = time.perf_counter()
                pr = [a[1:-1, 1:-1, 1:-1] for a in ws.predict_fine(geo, params)]
                fields = np.stack([pr[1], pr[2], pr[3], np.where(fluid, pr[0] - relaxed_mean_density(geo), 0.0)])
                t_pred = time.perf_counter() - t0
                del pr
                info = {}
                steps, k, sec = ensemble_run(params, geo, fields, tol, check, cold_max_steps,
                                             probe=(fallback_step, fallback_residual), info=info)
                row["ens_residual_at_probe"] = info.get("residual")
                row["ens_fallback"] = k is None
                if k is None:
                    lost, lost_s = steps, sec
                    steps, k, sec = ensemble_run(params, geo, None, tol, check, cold_max_steps)
                    steps, sec = steps + lost, sec + lost_s
                sec += t_pred
            else:
                steps, k, sec = ensemble_run(params, geo, None, tol, check, cold_max_steps)
            row[f"{name}_steps"], row[f"{name}_s"], row[f"k_{name}"] = steps, sec, k
            if "k_cold" in row:
                row[f"k_{name}_rel_diff"] = abs(k - row["k_cold"]) / abs(row["k_cold"])
            log(f"held-out {s}: {name} {steps} steps {sec:.2f} s, k {k:.5f}"
                + (f", fallback {row['ens_fallback']}" if name == "warm_ens" else "")
                + (f", fallback {row['coarse_fallback']}" if name == "coarse_ens" else ""))
        rows.append(row)
    return rows


def evaluate(ws, seeds, n=64, precision=gpu.FP32, log=print):
    """Per held-out geometry: the reference's cold start in FP64 (the arbiter), the relaxed
    start from rest and the learned warm start, both in `precision`."""
    rows = []
    for s in seeds:
        geo, params, meta = single_phase_case(s, n)
        row = dict(meta)
        sim = gpu.Solver(params, geo, gpu.FP64)
        t0 = time.perf_counter()
        row["cold_steps"], cold, rho_cold = run_to_steady(sim, params)
        row["cold_s"] = time.perf_counter() - t0
        sim.close()
        k_cold = permeability(cold, params)
        row["k_cold"] = k_cold
        # accuracy of the predicted fields against the steady state (fluid nodes)
        fluid = ~solid_with_walls(geo)
        pr = [a[1:-1, 1:-1, 1:-1] for a in ws.predict(geo, params)]
        vel_p = np.stack(pr[1:])
        vel_p[2] += 0.5 * BODY_FORCE * fluid
        row["velocity_rel_err"] = float(np.linalg.norm((vel_p - cold)[:, fluid]) / np.linalg.norm(cold[:, fluid]))
        d_ref = rho_cold[fluid] - rho_cold[fluid].mean()
        d_pre = pr[0][fluid] - pr[0][fluid].mean()
        row["density_rel_err"] = float(np.linalg.norm(d_pre - d_ref) / max(np.linalg.norm(d_ref), 1e-300))
        sim = gpu.Solver(params, geo, precision)
        for name in ("relaxed", "warm"):
            t0 = time.perf_counter()
            init = ws.predict(geo, params) if name == "warm" else rest_state(geo)
            steps, vel, _ = run_to_steady(sim, params, init, relaxed=True)
            row[f"{name}_steps"] = steps
            row[f"{name}_s"] = time.perf_counter() - t0
            row[f"k_{name}"] = permeability(vel, params)
            row[f"k_{name}_rel_diff"] = abs(row[f"k_{name}"] - k_cold) / abs(k_cold)
        sim.close()
        rows.append(row)
        log(f"held-out {s}: field errors u {row['velocity_rel_err']:.3f} rho {row['density_rel_err']:.3f}; "
            f"cold {row['cold_steps']} steps {row['cold_s']:.2f} s, relaxed {row['relaxed_steps']} "
            f"steps {row['relaxed_s']:.2f} s, warm {row['warm_steps']} steps {row['warm_s']:.2f} s; k {k_cold:.4f} / "
            f"{row['k_relaxed']:.4f} / {row['k_warm']:.4f}")
    return rows


def summarise(rows):
    tot = {k: float(sum(r[k] for r in rows)) for k in ("cold_steps", "relaxed_steps", "warm_steps", "cold_s",
                                                        "relaxed_s", "warm_s")}
    return {"mean_cold_steps": tot["cold_steps"] / len(rows), "mean_relaxed_steps": tot["relaxed_steps"] / len(rows),
            "mean_warm_steps": tot["warm_steps"] / len(rows),
            "warm_vs_cold_wall": tot["warm_s"] / tot["cold_s"], "relaxed_vs_cold_wall": tot["relaxed_s"] / tot["cold_s"],
            "warm_vs_relaxed_steps": tot["warm_steps"] / tot["relaxed_steps"],
            "max_k_warm_rel_diff": float(max(r["k_warm_rel_diff"] for r in rows)),
            "max_k_relaxed_rel_diff": float(max(r["k_relaxed_rel_diff"] for r in rows)),
            "mean_velocity_rel_err": float(np.mean([r["velocity_rel_err"] for r in rows])),
            "mean_density_rel_err": float(np.mean([r["density_rel_err"] for r in rows]))}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--train", type=int, default=48)
    ap.add_argument("--eval", type=int, default=8)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--precision", choices=("fp32", "fp64"), default="fp32", help="of the relaxed and warm solves")
    ap.add_argument("--rho-weight", type=float, default=16.0, help="loss weight of the density channel")
    ap.add_argument("--out", default=None, help="write the JSON summary here as well")
    ap.add_argument("--widths", default="16,32,64,128", help="channels of the four scales")
    ap.add_argument("--fine", action="store_true", help="evaluate on geometries twice as fine as --n")
    ap.add_argument("--save", default=None, help="write the trained network (.npz) here")
    ap.add_argument("--load", default=None, help="skip data and training: evaluate this network (.npz)")
    ap.add_argument("--starts", default="cold,relaxed,warm", help="fine evaluation: which starts to run")
    ap.add_argument("--eval-seeds", default=None, help="comma-separated held-out seeds (default 1000, 1001, ...)")
    ap.add_argument("--cold-max-steps", type=int, default=400000)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    eval_seeds = ([int(v) for v in args.eval_seeds.split(",")] if args.eval_seeds
                  else list(range(1000, 1000 + args.eval)))
    prec = gpu.FP32 if args.precision == "fp32" else gpu.FP64
    if args.load:
        ws = WarmStart.load(args.load)
        log(f"network parameters: {ws.net.n_params()} (loaded)")
        rows = evaluate_fine(ws, eval_seeds, args.n, prec, log, starts=tuple(args.starts.split(",")),
                             cold_max_steps=args.cold_max_steps)
        text = json.dumps({"n": args.n, "fine": True, "params": ws.net.n_params(), "eval": rows})
        if args.out:
            with open(args.out, "w") as fh:
                fh.write(text + "\n")
        print(text, flush=True)
        return
    t0 = time.perf_counter()
    data = generate(range(args.train), args.n, log)
    t_data = time.perf_counter() - t0
    ws = WarmStart(widths=tuple(int(c) for c in args.widths.split(",")), rho_weight=args.rho_weight)
    log(f"network parameters: {ws.net.n_params()}")
    t0 = time.perf_counter()
    hist = ws.train(data, args.epochs, args.lr, log)
    t_train = time.perf_counter() - t0
    if args.save:
        ws.save(args.save)
    # fit on a few training samples, in the same terms as the held-out field errors
    fit = []
    for geo, target, params, _ in data[:4]:
        fluid = ~solid_with_walls(geo)
        pr = [a[1:-1, 1:-1, 1:-1] for a in ws.predict(geo, params)]
        nu = 1.0 / params["la_nui1"]
        vel = target[:3] * BODY_FORCE / nu
        vel_p = np.stack(pr[1:])
        vel_p[2] += 0.5 * BODY_FORCE * fluid
        d_ref = target[3][fluid] * 3.0 * BODY_FORCE
        d_pre = pr[0][fluid] - pr[0][fluid].mean()
        fit.append((float(np.linalg.norm((vel_p - vel)[:, fluid]) / np.linalg.norm(vel[:, fluid])),
                    float(np.linalg.norm(d_pre - d_ref) / np.linalg.norm(d_ref))))
    log("training-sample field errors (u, rho): " + ", ".join(f"({a:.3f}, {b:.3f})" for a, b in fit))
    del data
    res = {"n": args.n, "train_samples": args.train, "epochs": args.epochs, "params": ws.net.n_params(),
           "precision": args.precision, "data_seconds": t_data, "train_seconds": t_train, "loss_first": hist[0],
           "loss_last": hist[-1], "train_fit": fit}
    if args.fine:
        rows = evaluate_fine(ws, eval_seeds, args.n, prec, log, starts=tuple(args.starts.split(",")),
                             cold_max_steps=args.cold_max_steps)
        res.update(fine=True, eval=rows)
    else:
        rows = evaluate(ws, eval_seeds, args.n, prec, log)
        res.update(**summarise(rows), eval=rows)
    text = json.dumps(res)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
