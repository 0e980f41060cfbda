"""The relaxed start of the warm-start path (mflbm_init_relaxed + relaxed_mean_density),
pinned on the FP64 oracle.

The collision relaxes the fourth-order moment to 0, so its rest state is rho (4/21, 13/126,
1/63) rather than the D3Q19 weights (1/3, 1/18, 1/36).  A start at those rest weights with the
mean density fixed by the cold start's mass must reach the cold start's steady state: the same
mean density and permeability.
"""
import numpy as np

import lbm_textbook as TB
import oracle as O
from mflbm import cases, config, warmstart

REST = np.array([4 / 21] + [13 / 126] * 6 + [1 / 63] * 12)


def test_relaxed_weights_are_the_collision_rest_state():
    # the populations of the moment-space rest equilibrium (rho, e = -11 rho, 0, ...)
    np.testing.assert_allclose(TB.equilibrium_moments_state(1.0), REST, rtol=0, atol=1e-15)
    p = config.derive(cases.periodic_control(n=(8, 8, 8)))
    p.update(lbm_gamma=0.0, force_z=0.0)
    for rho in (1.0, 0.97):
        _, o1, o2 = O.collide_node(p, rho * REST, np.zeros(19), (0, 0, 0), 0.0, 0.0)
        np.testing.assert_allclose(o1, rho * REST, rtol=0, atol=1e-15)
        assert np.all(o2 == 0)


def test_oracle_relaxed_init_at_rest_and_with_velocity():
    n = (6, 5, 4)
    p = config.derive(cases.periodic_control(n=n, option=1, interface_z0=100.0))
    o = O.Oracle(p, np.zeros(n[::-1], np.int8))
    o.init_phi(1, 100.0)
    rho = np.full(o.shape(1), 0.98)
    o.init_pdf(rho, relaxed=True)
    f = o.array("pdf")
    np.testing.assert_allclose(f[0], 0.98 * REST[:, None, None, None] * np.ones(o.shape(1)), rtol=0, atol=1e-16)
    assert np.all(f[1] == 0)
    # the velocity terms are those of the D3Q19 equilibrium: relaxed - D3Q19 = rho (r_q - w_q)
    rng = np.random.default_rng(3)
    uvw = [1e-3 * rng.standard_normal(o.shape(1)) for _ in range(3)]
    o.init_pdf(rho, *uvw, relaxed=True)
    relaxed = o.array("pdf")[0].copy()
    o.init_pdf(rho, *uvw)
    plain = o.array("pdf")[0].copy()
    np.testing.assert_allclose(relaxed - plain, (0.98 * (REST - TB.W))[:, None, None, None] * np.ones(o.shape(1)),
                               rtol=0, atol=1e-15)


def test_flat_walls_leave_the_mean_density_unchanged():
    # per wall node one axis link and four diagonal links: 1/18 + 4/36 = 13/126 + 4/63
    geo = np.zeros((10, 9, 8), np.int8)
    assert abs(warmstart.relaxed_mean_density(geo) - 1.0) < 1e-15


def _run_to_steady(o, tol=1e-5, check=100, max_steps=20000):
    prev, t = None, 1
    while t <= max_steps:
        o.run(t, check)
        t += check
        rho, u, v, w = o.macro(zero_solid_phi=False)
        vel = np.stack([u, v, w])[:, 1:-1, 1:-1, 1:-1]
        if prev is not None and np.linalg.norm(vel - prev) / np.linalg.norm(vel) < tol:
            break
        prev = vel
    return t - 1, rho, w


def test_relaxed_start_reaches_the_cold_steady_state_sooner():
    """16^3 single-phase sphere pack: the reference's cold start against the relaxed start."""
    n = 16
    geo = cases.geometry.sphere_pack((n, n, n), 3.0, 0.6, 4, periodic=(True, True, False))
    p = config.derive(cases.periodic_control(n=(n, n, n), body_force=1e-6, option=1, interface_z0=float(n + 10)))
    fluid = np.pad(~warmstart.solid_with_walls(geo), 1)
    rho_bar = warmstart.relaxed_mean_density(geo)
    out = {}
    for relaxed in (False, True):
        o = O.Oracle(p, geo)
        o.init_phi(1, float(n + 10))
        o.init_pdf(np.full(o.shape(1), rho_bar) if relaxed else None, relaxed=relaxed)
        o.color_gradient()
        out[relaxed] = _run_to_steady(o)
        o.close()
    (s_cold, r_cold, w_cold), (s_rel, r_rel, w_rel) = out[False], out[True]
    # mass balance predicts the cold start's steady mean density
    assert abs(r_cold[fluid].mean() - rho_bar) < 1e-6
    assert abs(r_rel[fluid].mean() - rho_bar) < 1e-6
    k_cold, k_rel = w_cold[1:-1, 1:-1, 1:-1].mean(), w_rel[1:-1, 1:-1, 1:-1].mean()
    assert abs(k_rel / k_cold - 1) < 1e-4
    assert 3 * s_rel <= s_cold, (s_rel, s_cold)


def test_coarsen_inverts_refine_and_keeps_the_majority():
    from mflbm import warmstart
    rng = np.random.default_rng(3)
    g = (rng.random((8, 6, 10)) < 0.4).astype(np.int8)
    f = warmstart.refine(g)
    assert f.shape == (16, 12, 20) and np.array_equal(warmstart.coarsen(f), g)
    # a block with 3 of 8 solid nodes coarsens to pore, 4 of 8 to solid
    b = np.zeros((2, 2, 2), np.int8)
    b.flat[:3] = 1
    assert warmstart.coarsen(b)[0, 0, 0] == 0
    b.flat[3] = 1
    assert warmstart.coarsen(b)[0, 0, 0] == 1


def test_fine_case_is_the_refined_coarse_case():
    from mflbm import warmstart
    gc, pc, mc = warmstart.single_phase_case(1000, 16)
    gf, pf, mf = warmstart.fine_case(1000, 16)
    assert np.array_equal(warmstart.coarsen(gf), gc) and gf.shape == (32, 32, 32)
    assert mf["radius"] == 2 * mc["radius"] and pf["force_z"] == pc["force_z"] and pf["nz"] == 32
