"""Batched single-phase permeability ensemble (mflbm_ensemble_* of include/mflbm.h).

n independent domains (walls on both x faces, periodic in y and z, body force along z) are
stepped together on the device, one launch per step over all of them, each until its
velocity field is steady (relative L2 change over `check` steps below `tol`).  Every domain
starts from rest at the collision's rest weights and the mean density of the reference's cold
start, so it reaches the cold start's steady state.

    ens = Ensemble(params, geometries)          # geometries: (n, nz, ny, nx) int8
    out = ens.solve(check=100, tol=1e-5, max_steps=60000)
    fld = ens.fields(0, n)                      # (n, 4, nz, ny, nx) float32: u, v, w, rho - rho_ref
"""
from __future__ import annotations

import ctypes as C
import multiprocessing as mp

import numpy as np

from . import gpu


class EnsembleError(RuntimeError):
    pass


class Ensemble:
    def __init__(self, params: dict, geometries: np.ndarray):
        L = gpu.load()
        geo = np.ascontiguousarray(geometries, dtype=np.int8)
        if geo.ndim != 4:
            raise ValueError("geometries must be (n_domains, nz, ny, nx)")
        self.n, nz, ny, nx = geo.shape
        p = dict(params)
        if (int(p["nx"]), int(p["ny"]), int(p["nz"])) != (nx, ny, nz):
            raise ValueError("params and geometries disagree on the domain size")
        self.shape = (nz, ny, nx)
        self.nu = 1.0 / p["la_nui1"]
        self.force = p["force_z"]
        cp = gpu.make_params(p, gpu.FP32)
        self.rho_ref = np.zeros(self.n)
        self.h = C.c_void_p()
        rc = L.mflbm_ensemble_create(C.byref(cp), geo.ctypes.data_as(C.POINTER(C.c_int8)), self.n,
                                     self.rho_ref.ctypes.data_as(C.POINTER(C.c_double)), C.byref(self.h))
        if rc != 0:
            msg = L.mflbm_ensemble_last_error(self.h).decode() if self.h else "ensemble create failed"
            self.close()
            raise EnsembleError(msg)

    def _check(self, rc):
        if rc != 0:
            raise EnsembleError(gpu.load().mflbm_ensemble_last_error(self.h).decode())

    def fluid_nodes(self) -> int:
        return int(gpu.load().mflbm_ensemble_fluid_nodes(self.h))

    def solve(self, check=100, tol=1e-5, max_steps=60000) -> dict:
        """Step every domain towards its steady state, at most max_steps more steps.  Returns
        per-domain arrays: steps taken so far (negative while not yet steady), permeability
        (nu <w> / F over the whole domain), change (the last relative change) and the device
        time of this call's steps.  Calling again continues."""
        steps = np.zeros(self.n, np.int32)
        perm = np.zeros(self.n)
        change = np.zeros(self.n)
        ms = C.c_double()
        self._check(gpu.load().mflbm_ensemble_solve(
            self.h, int(check), float(tol), int(max_steps), steps.ctypes.data_as(C.POINTER(C.c_int32)),
            perm.ctypes.data_as(C.POINTER(C.c_double)), change.ctypes.data_as(C.POINTER(C.c_double)), C.byref(ms)))
        return {"steps": steps, "permeability": perm, "change": change, "device_ms": ms.value}

    def fields(self, first=0, count=None) -> np.ndarray:
        """(count, 4, nz, ny, nx) float32: u, v, w and rho - rho_ref, zero at solids."""
        count = self.n - first if count is None else count
        out = np.zeros((count, 4) + self.shape, np.float32)
        self._check(gpu.load().mflbm_ensemble_fields(self.h, int(first), int(count),
                                                    out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def init_fields(self, fields, first=0):
        """Start domains first .. from fields (count, 4, nz, ny, nx): u, v, w and rho - rho_ref
        (the layout of fields()); only before their first step."""
        a = np.ascontiguousarray(fields, dtype=np.float32)
        if a.ndim == 4:
            a = a[None]
        if a.shape[1:] != (4,) + self.shape:
            raise ValueError(f"fields must be (count, 4) + {self.shape}")
        self._check(gpu.load().mflbm_ensemble_init_fields(self.h, int(first), int(a.shape[0]),
                                                         a.ctypes.data_as(C.POINTER(C.c_float))))

    def velocity(self, first=0, count=None) -> np.ndarray:
        return self.fields(first, count)[:, :3]

    def close(self):
        if getattr(self, "h", None):
            gpu.load().mflbm_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _case(args):
    from . import warmstart
    seed, n = args
    geo, _, meta = warmstart.single_phase_case(seed, n)
    return geo, meta


def sphere_pack_ensemble(seeds, n=64, workers=8):
    """Geometries and metadata of the single-phase cases of mflbm.warmstart for `seeds`,
    built in a process pool (numpy, host)."""
    seeds = list(seeds)
    if workers > 1 and len(seeds) > 8:
        with mp.get_context("fork").Pool(workers) as pool:
            res = pool.map(_case, [(s, n) for s in seeds], chunksize=4)
    else:
        res = [_case((s, n)) for s in seeds]
    return np.stack([g for g, _ in res]), [m for _, m in res]


def main(argv=None):
    """Ensemble of n single-phase domains of 64^3, porosity U[0.30, 0.50], radius U[4, 10], body force 1e-6,
    each to a relative velocity change below 1e-7 per step (tol per `check` steps)."""
    import argparse
    import json
    import sys
    import time

    from . import warmstart
    ap = argparse.ArgumentParser(description=main.__doc__)
    ap.add_argument("--domains", type=int, default=2048)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--check", type=int, default=100)
    ap.add_argument("--tol", type=float, default=1e-5, help="relative change per `check` steps")
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    log = lambda m: print(m, file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    geos, metas = sphere_pack_ensemble(range(args.domains), args.n, args.workers)
    t_geo = time.perf_counter() - t0
    log(f"{args.domains} geometries in {t_geo:.1f} s")
    _, p, _ = warmstart.single_phase_case(0, args.n)
    t0 = time.perf_counter()
    ens = Ensemble(p, geos)
    t_create = time.perf_counter() - t0
    nf = ens.fluid_nodes()
    log(f"ensemble created in {t_create:.1f} s: {nf} fluid nodes")
    t0 = time.perf_counter()
    device_ms, done = 0.0, 0
    while done < args.max_steps:  # in chunks, so that progress is visible
        out = ens.solve(args.check, args.tol, min(2000, args.max_steps - done))
        device_ms += out["device_ms"]
        done += min(2000, args.max_steps - done)
        active = int((out["steps"] < 0).sum())
        log(f"step {done}: {args.domains - active} of {args.domains} domains steady")
        if active == 0:
            break
    out["device_ms"] = device_ms
    t_solve = time.perf_counter() - t0
    steps = out["steps"]
    fl = np.array([int((g[:, :, 1:-1] == 0).sum()) for g in geos])  # fluid nodes (x faces are walls)
    node_steps = float(np.sum(np.abs(steps).astype(np.float64) * fl))
    res = {
        "domains": args.domains, "n": args.n, "fluid_nodes": nf, "geometry_seconds": t_geo,
        "create_seconds": t_create, "solve_seconds": t_solve, "device_ms": out["device_ms"],
        "steady": int((steps > 0).sum()), "steps_min": int(np.abs(steps).min()), "steps_mean": float(np.abs(steps).mean()),
        "steps_max": int(np.abs(steps).max()),
        "fluid_glups_device": node_steps / (out["device_ms"] * 1e-3) / 1e9,
        # 19 populations read and written (4 B), plus the 38.5 B short neighbour row on odd steps
        "hbm_gbs_algorithmic": node_steps * (152.0 + 19.25) / (out["device_ms"] * 1e-3) / 1e9,
        "permeability_min": float(out["permeability"].min()), "permeability_max": float(out["permeability"].max()),
        "rho_ref_range": [float(ens.rho_ref.min()), float(ens.rho_ref.max())],
    }
    v = ens.velocity(0, 1)
    res["velocity_domain0_finite"] = bool(np.isfinite(v).all())
    ens.close()
    text = json.dumps(res)
    if args.out:
        with open(args.out, "w") as fh:
            fh.write(text + "\n")
    print(text, flush=True)


if __name__ == "__main__":
    main()
