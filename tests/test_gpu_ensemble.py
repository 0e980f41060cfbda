"""The batched single-phase ensemble on the GPU against single-domain runs of the step path
and against the FP64 oracle."""
import sys

import numpy as np
import pytest

import oracle as O
from mflbm import cases, config, ensemble, gpu, warmstart

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-ensemble] {msg}", flush=True)
    sys.stdout.flush()


def params(n):
    nx, ny, nz = n
    return config.derive(cases.periodic_control(n=n, body_force=1e-6, option=1, interface_z0=float(nz + 10)))


def geometries(n, seeds, radius=3.0, porosity=0.6):
    nx, ny, nz = n
    return np.stack([cases.geometry.sphere_pack((nz, ny, nx), radius, porosity, s, periodic=(True, True, False))
                     for s in seeds])


def single_run(p, geo, check, tol, max_steps, precision=gpu.FP32):
    """The same relaxed start through the two-colour step path (mflbm_init_relaxed)."""
    s = gpu.Solver(p, geo, precision)
    init = warmstart.rest_state(geo)
    steps, vel, rho = warmstart.run_to_steady(s, p, init, relaxed=True, tol=tol, check=check, max_steps=max_steps)
    s.close()
    return steps, vel, rho


def test_rest_density_matches_the_host_formula():
    n = (20, 16, 24)
    geos = geometries(n, [1, 2, 3])
    ens = ensemble.Ensemble(params(n), geos)
    for d in range(3):
        assert abs(ens.rho_ref[d] - warmstart.relaxed_mean_density(geos[d])) < 1e-13
    ens.close()


def test_ensemble_matches_single_domain_runs():
    """Three domains of different geometry (and convergence time) in one ensemble against
    three separate FP32 runs of the two-colour path from the same relaxed start."""
    n = (20, 16, 24)
    p = params(n)
    geos = geometries(n, [4, 7, 9])
    say("ensemble of 3 domains 20x16x24")
    ens = ensemble.Ensemble(p, geos)
    out = ens.solve(check=100, tol=1e-5, max_steps=20000)
    fields = ens.fields()
    vel = fields[:, :3]
    rho_ref = ens.rho_ref.copy()
    ens.close()
    assert (out["steps"] > 0).all(), out
    for d in range(3):
        say(f"single run of domain {d}")
        steps, v1, r1 = single_run(p, geos[d], 100, 1e-5, 20000)
        fluid = ~warmstart.solid_with_walls(geos[d])
        assert np.abs(rho_ref[d] + fields[d, 3][fluid] - r1[fluid]).max() < 1e-7
        assert abs(steps - out["steps"][d]) <= 100, (steps, out["steps"][d])
        err = np.linalg.norm(vel[d] - v1) / np.linalg.norm(v1)
        say(f"domain {d}: steps {out['steps'][d]} vs {steps}, velocity rel L2 {err:.2e}")
        assert err < 1e-4
        k1 = warmstart.permeability(v1, p)
        assert abs(out["permeability"][d] / k1 - 1) < 1e-4


def test_ensemble_permeability_matches_fp64_oracle_cold_start():
    """One 16^3 domain: the ensemble's steady permeability against the FP64 oracle run from
    the reference's own initial state."""
    n = (16, 16, 16)
    p = params(n)
    geo = cases.geometry.sphere_pack((16, 16, 16), 3.0, 0.6, 4, periodic=(True, True, False))
    ens = ensemble.Ensemble(p, geo[None])
    out = ens.solve(check=100, tol=1e-5, max_steps=20000)
    ens.close()
    o = O.Oracle(p, geo)
    o.setup()
    prev, t = None, 1
    while t <= 20000:
        o.run(t, 100)
        t += 100
        _, u, v, w = o.macro(zero_solid_phi=False)
        vel = np.stack([u, v, w])[:, 1:-1, 1:-1, 1:-1]
        if prev is not None and np.linalg.norm(vel - prev) / np.linalg.norm(vel) < 1e-5:
            break
        prev = vel
    k_oracle = warmstart.permeability(vel, p)
    say(f"ensemble k {out['permeability'][0]:.6f} in {out['steps'][0]} steps, oracle {k_oracle:.6f} in {t - 1}")
    assert abs(out["permeability"][0] / k_oracle - 1) < 1e-3
    assert 3 * out["steps"][0] <= t - 1


def test_converged_domains_stay_frozen():
    n = (20, 16, 24)
    p = params(n)
    geos = geometries(n, [4, 7])
    ens = ensemble.Ensemble(p, geos)
    a = ens.solve(check=100, tol=1e-5, max_steps=20000)
    v1 = ens.velocity()
    b = ens.solve(check=100, tol=1e-5, max_steps=20000)  # nothing left to step
    v2 = ens.velocity()
    ens.close()
    assert np.array_equal(v1, v2)
    assert (a["steps"] > 0).all() and np.array_equal(a["steps"], b["steps"])
    assert np.array_equal(a["permeability"], b["permeability"]) and b["device_ms"] == 0.0


def test_solve_in_chunks_equals_one_call():
    n = (20, 16, 24)
    p = params(n)
    geos = geometries(n, [4, 7, 9])
    one = ensemble.Ensemble(p, geos)
    a = one.solve(check=100, tol=1e-5, max_steps=20000)
    va = one.velocity()
    one.close()
    two = ensemble.Ensemble(p, geos)
    for _ in range(40):
        b = two.solve(check=100, tol=1e-5, max_steps=300)
    vb = two.velocity()
    two.close()
    assert np.array_equal(a["steps"], b["steps"])
    assert np.array_equal(va, vb)


def test_init_fields_starts_from_the_given_state():
    """mflbm_ensemble_init_fields: the stored state is the relaxed equilibrium of the given
    fields (checked through the moments the steady test reads), and a start from the steady
    fields of a converged ensemble reaches the same state, in no more steps."""
    n = (20, 16, 24)
    p = params(n)
    geos = geometries(n, [4, 7])
    a = ensemble.Ensemble(p, geos)
    out_a = a.solve(check=100, tol=1e-5, max_steps=20000)
    fa = a.fields()
    a.close()
    b = ensemble.Ensemble(p, geos)
    init = fa.copy()
    # fields() holds the steady test's velocity, the momentum less F / 2 along z
    init[:, 2] += np.where(fa[:, 2] != 0, 0.5 * p["force_z"], 0.0).astype(np.float32)
    b.init_fields(init)
    f0 = b.fields()
    for d in range(2):
        fluid = ~warmstart.solid_with_walls(geos[d])
        rho = b.rho_ref[d] + fa[d, 3].astype(np.float64)
        # x momentum = rho u_x (no force along x); density deviation exact
        assert np.abs(f0[d, 0][fluid] - rho[fluid] * fa[d, 0][fluid]).max() < 1e-6 * np.abs(fa[d, 0]).max()
        assert np.abs(f0[d, 3][fluid] - fa[d, 3][fluid]).max() < 1e-6
    out_b = b.solve(check=100, tol=1e-5, max_steps=20000)
    fb = b.fields()
    b.close()
    say(f"steps from rest {out_a['steps']}, from the steady fields {out_b['steps']}")
    assert (out_b["steps"] > 0).all() and (out_b["steps"] <= out_a["steps"]).all()
    for d in range(2):
        err = np.linalg.norm(fb[d, :3] - fa[d, :3]) / np.linalg.norm(fa[d, :3])
        assert err < 1e-4, err
        assert abs(out_b["permeability"][d] / out_a["permeability"][d] - 1) < 1e-4


def test_config4_2048_domains_at_real_size():
    """Config 4 at its real size: 2048 seeded 64^3 sphere packs (porosity U[0.30, 0.50],
    radius U[4, 10], body force 1e-6) in one ensemble, each to a relative velocity change
    below 1e-7 per step.  Every domain must become steady with a positive permeability, and
    three domains spread over the batch, solved again on their own, must give the same step
    counts, permeabilities and fields bit for bit (the domains do not interact)."""
    n, count, check, tol = 64, 2048, 100, 1e-5
    say(f"config 4: {count} geometries of {n}^3")
    geos, _ = ensemble.sphere_pack_ensemble(range(count), n, workers=16)
    _, p, _ = warmstart.single_phase_case(0, n)
    ens = ensemble.Ensemble(p, geos)
    done = 0
    while done < 60000:
        out = ens.solve(check, tol, 2000)
        done += 2000
        active = int((out["steps"] < 0).sum())
        say(f"step {done}: {count - active} of {count} domains steady")
        if active == 0:
            break
    assert (out["steps"] > 0).all()
    assert np.isfinite(out["permeability"]).all() and (out["permeability"] > 0).all()
    sample = [0, 1023, 2047]
    fields = {d: ens.fields(first=d, count=1)[0] for d in sample}
    ens.close()
    for d in sample:
        one = ensemble.Ensemble(p, geos[d:d + 1])
        o = one.solve(check, tol, 60000)
        f = one.fields()[0]
        one.close()
        say(f"domain {d}: {o['steps'][0]} steps alone, {out['steps'][d]} in the batch")
        assert o["steps"][0] == out["steps"][d]
        assert o["permeability"][0] == out["permeability"][d]
        assert np.array_equal(f, fields[d])
