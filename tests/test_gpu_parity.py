"""GPU parity: the HIP step (through the C ABI) against the FP64 oracle on the same inputs.

FP64 handles must agree bit for bit on every field the reference keeps (the 38
population slots including ghosts, phi, normals, curvature at fluid nodes) — the
required bound is a relative L-infinity of 1e-12, bit equality is stronger.  Monitor
sums are reduced in a different order on the device; they must agree to 1e-12 of the
summed magnitudes.

Every test runs on both gradient chains the library can choose: the one-pass z-march
(the default on large domains) and the two-pass 3-D tiles (the default on small ones).
"""
import os
import sys

import numpy as np
import pytest

import oracle as O
from helpers import assert_same, periodic_drop, rel_linf, small_drainage
from mflbm import cases, config, gpu

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


PATH = {"march": "0", "tiles": "5"}
_chain = {}


@pytest.fixture(autouse=True, params=sorted(PATH))
def gradient_chain(request, monkeypatch):
    """Forces the handle's gradient chain (MFLBM_GRAD_SHAPE is read at creation)."""
    monkeypatch.setenv("MFLBM_GRAD_SHAPE", PATH[request.param])
    _chain["name"] = request.param
    yield request.param


def pair(geo, params, precision=gpu.FP64):
    o = O.Oracle(params, geo)
    g = gpu.Solver(params, geo, precision)
    assert g.gradient_path() == _chain["name"]
    return o, g


def fluid_interior_mask(o):
    walls = o.int_array("walls")[1:-1, 1:-1, 1:-1]  # s2 -> s1 footprint
    m = np.zeros(o.shape(1), bool)
    m[1:-1, 1:-1, 1:-1] = walls[1:-1, 1:-1, 1:-1] == 0
    return m


def compare_geometry(o, g):
    assert_same(g.download("walls").astype(np.int32), o.int_array("walls"), what="walls")
    assert_same(g.download("walls_type").astype(np.int32), o.int_array("walls_type"), what="walls_type")
    for n in ("s_nx", "s_ny", "s_nz"):
        assert_same(g.download(n), o.array(n), what=n)


def compare_state(o, g, tol=0.0, tag=""):
    assert_same(g.download("pdf"), o.array("pdf"), tol, what=f"{tag} pdf")
    assert_same(g.download("phi"), o.array("phi"), tol, what=f"{tag} phi")
    open_z = o.p["kper"] == 0 and o.p["wall_z_min"] == 0 and o.p["wall_z_max"] == 0
    if open_z and o.p["inlet_bc"] == 1:
        assert_same(g.download("w_in"), o.array("w_in"), tol, what=f"{tag} w_in")
    if open_z and o.p["outlet_bc"] == 1:
        assert_same(g.download("convec"), o.array("convec"), tol, what=f"{tag} convective outlet state")
    for n in ("cn_x", "cn_y", "cn_z", "c_norm"):
        assert_same(g.download(n), o.array(n), tol, what=f"{tag} {n}")
    m = fluid_interior_mask(o)
    assert_same(g.download("curv")[m], o.array("curv")[m], tol, what=f"{tag} curv(fluid)")


CASES = {
    "drainage_yper": lambda: small_drainage(n=(24, 24, 32), seed=7, jper=1),
    "drainage_ywalls": lambda: small_drainage(n=(20, 24, 28), seed=11, jper=0, theta=30.0, nu_ratio=0.1),
    "drop_periodic_yz": lambda: periodic_drop(n=(24, 20, 22), radius=6.0, geo_seed=3, body_force=1e-5),
    # velocity inlet (src/main_iteration_GPU.cu:1008-1077) and convective outlet (:1245-1357)
    "velocity_in_convective_out": lambda: small_drainage(n=(20, 18, 28), seed=5, jper=0, inlet=1, outlet=1),
    "velocity_in_zouhe_out_yper": lambda: small_drainage(n=(22, 20, 26), seed=8, jper=1, inlet=1, outlet=2),
    "no_inlet_bc_convective_out": lambda: small_drainage(n=(20, 20, 24), seed=9, jper=1, inlet=0, outlet=1),
    # porous plate (src/main_iteration_GPU.cu:1744-1882) across the sphere pack / in the inlet buffer
    "plate_blocks_fluid1": lambda: small_drainage(n=(20, 22, 28), seed=6, jper=1, plate=1, plate_z=14),
    "plate_blocks_fluid2": lambda: small_drainage(n=(18, 20, 26), seed=4, jper=0, plate=2, plate_z=2),
    # sizes that no workgroup shape divides, and an axis so short that the inlet and outlet
    # columns share their threads
    "awkward_sizes": lambda: small_drainage(n=(13, 11, 9), seed=4, radius=2.5, porosity=0.5, jper=1, buffer=2),
    "five_planes": lambda: small_drainage(n=(12, 10, 5), seed=2, radius=2.0, porosity=0.6, jper=1, buffer=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_setup_and_steps_bitwise(name):
    geo, p = CASES[name]()
    say(f"case={name} shape={geo.shape} fp64 setup")
    o, g = pair(geo, p)
    compare_geometry(o, g)
    assert o.counts()[0] >= 0
    o.setup()
    g.setup()
    compare_state(o, g, tag=f"{name} init")
    for t in range(1, 5):
        say(f"case={name} step {t}")
        o.step(t)
        g.step(t)
        compare_state(o, g, tag=f"{name} step{t}")
    say(f"case={name} graph run 5..40")
    o.run(5, 36)
    g.run(5, 36)
    compare_state(o, g, tag=f"{name} step40")


def test_long_link_waves_bitwise():
    """Two solid plates across a 256 x 256 plane between fully fluid layers, each with two
    holes near opposite corners: the two holes of a plate are consecutive fluid nodes, so
    the wave holding them links to fluid cells a whole plane apart, more than the 16-bit
    neighbour rows span, and reads the 32-bit rows while every other wave decodes the
    16-bit ones.  Both kinds match the oracle bit for bit."""
    geo, p = small_drainage(n=(256, 256, 12), seed=3, jper=1, buffer=6)
    geo[3] = 1
    geo[3, 1, 1] = geo[3, 254, 254] = 0
    geo[8] = 1
    geo[8, 2, 3] = geo[8, 253, 250] = 0
    say("case=long_link_waves 256x256x12 fp64")
    o, g = pair(geo, p)
    assert g.counts()["long_link_waves"] >= 1
    o.setup()
    g.setup()
    for t in range(1, 5):
        o.step(t)
        g.step(t)
    compare_state(o, g, tag="long links step4")
    o.run(5, 8)
    g.run(5, 8)
    compare_state(o, g, tag="long links step12")


@pytest.mark.parametrize("name", list(CASES))
def test_fused_gradient_matches_split(name):
    """The fused gradient-chain kernel and the four split kernels give identical steps.

    No download between the steps, so every collision reads force records written by
    the kernel under test."""
    geo, p = CASES[name]()
    say(f"case={name} fused vs split gradient, 30 steps")
    a = gpu.Solver(p, geo, gpu.FP64)
    b = gpu.Solver(p, geo, gpu.FP64, split_gradient=True)
    a.setup()
    b.setup()
    a.run(1, 30)
    b.run(1, 30)
    assert_same(a.download("pdf"), b.download("pdf"), what=f"{name} pdf")
    for n in ("phi", "cn_x", "cn_y", "cn_z", "c_norm", "curv"):
        assert_same(a.download(n), b.download(n), what=f"{name} {n}")


def test_mixed_phase_field_bitwise():
    """A random colour field (no uniform region anywhere) against the oracle: every
    gradient stage takes its full computation, none the uniform-region shortcut."""
    geo, p = CASES["drainage_ywalls"]()
    say("case=drainage_ywalls random phi, 20 steps")
    o, g = pair(geo, p)
    o.init_phi(1, p["interface_z0"])
    g.init_phi(1, p["interface_z0"])
    rng = np.random.default_rng(12)
    ph = rng.uniform(-1.0, 1.0, size=o.shape(4))
    o.array("phi")[...] = ph
    g.upload("phi", ph)
    o.init_pdf()
    g.init_equilibrium()
    o.color_gradient()
    g.color_gradient()
    compare_state(o, g, tag="random phi init")
    o.run(1, 20)
    g.run(1, 20)
    compare_state(o, g, tag="random phi step20")


def test_relaxed_init_bitwise():
    """mflbm_init_relaxed (the warm start's initialisation: rest weights 4/21, 13/126, 1/63)
    from random density and velocity fields, then 20 steps, against the oracle."""
    geo, p = CASES["drop_periodic_yz"]()
    say("case=drop_periodic_yz relaxed init from random rho/u, 20 steps")
    o, g = pair(geo, p)
    o.init_phi(5, p["interface_z0"])
    g.init_phi(5, p["interface_z0"])
    rng = np.random.default_rng(21)
    shp = o.shape(1)
    rho = 1.0 + 1e-3 * rng.standard_normal(shp)
    uvw = [1e-3 * rng.standard_normal(shp) for _ in range(3)]
    o.init_pdf(rho, *uvw, relaxed=True)
    g.init_equilibrium(rho, *uvw, relaxed=True)
    o.color_gradient()
    g.color_gradient()
    compare_state(o, g, tag="relaxed init")
    o.run(1, 20)
    g.run(1, 20)
    compare_state(o, g, tag="relaxed step20")


@pytest.mark.parametrize("open_z", [True, False])
def test_random_initial_distribution_bitwise(open_z):
    """Option 6 (mflbm_init_phi_random): the seeded random phase field, the populations
    built on it and 30 steps of its break-up agree bit for bit with the oracle."""
    n = (20, 18, 24)
    nx, ny, nz = n
    if open_z:
        ctrl = cases.drainage_control(n=n, interface_z0=3.0)
        geo = cases.sphere_geometry(n, 3.0, 0.7, 4, buffer=3)
    else:
        ctrl = cases.periodic_control(n=n, body_force=1e-5, option=6)
        geo = np.zeros((nz, ny, nx), np.int8)
    ctrl.update(initial_fluid_distribution_option=6, target_fluid1_saturation=0.35, random_seed=2024)
    p = config.derive(ctrl)
    say(f"case=option 6 open_z={open_z}")
    o, g = pair(geo, p)
    o.setup()
    g.setup()
    assert_same(g.download("phi"), o.array("phi"), what="initial phi")
    compare_state(o, g, tag="setup")
    o.run(1, 30)
    g.run(1, 30)
    compare_state(o, g, tag="step 30")


def test_fused_gradient_short_z_runs(monkeypatch, gradient_chain):
    """Workgroups marching 5-plane z-runs (many run boundaries) still match the oracle."""
    if gradient_chain != "march":
        pytest.skip("z-runs belong to the march")
    monkeypatch.setenv("MFLBM_GRAD_KCHUNK", "5")
    geo, p = CASES["drainage_ywalls"]()
    say("case=drainage_ywalls kchunk=5, 24 steps")
    o, g = pair(geo, p)
    o.setup()
    g.setup()
    o.run(1, 24)
    g.run(1, 24)
    compare_state(o, g, tag="kchunk5 step24")


def test_fused_gradient_fp32_matches_split():
    geo, p = CASES["drainage_yper"]()
    say("case=drainage_yper fp32 fused vs split, 30 steps")
    a = gpu.Solver(p, geo, gpu.FP32)
    b = gpu.Solver(p, geo, gpu.FP32, split_gradient=True)
    a.setup()
    b.setup()
    a.run(1, 30)
    b.run(1, 30)
    # FP32 is built with contraction allowed, so the two paths may round differently
    assert_same(a.download("pdf"), b.download("pdf"), 1e-5, what="fp32 pdf")
    assert_same(a.download("phi"), b.download("phi"), 1e-5, what="fp32 phi")


def test_config1_1000_steps():
    """64^3 drainage of the first workload, 1000 steps, relative L-inf <= 1e-12 (bitwise expected)."""
    geo, ctrl = cases.drainage64()
    p = config.derive(ctrl)
    say("case=drainage64 64^3 fp64 1000 steps")
    o, g = pair(geo, p)
    o.setup()
    g.setup()
    o.run(1, 1000)
    g.run(1, 1000)
    for n in ("pdf", "phi"):
        r = rel_linf(g.download(n), o.array(n))
        assert r <= 1e-12, f"{n}: relative Linf {r:.3e}"
    compare_state(o, g, tag="drainage64 step1000")
    rho_o, u_o, v_o, w_o = o.macro(zero_solid_phi=False)
    rho_g, u_g, v_g, w_g = g.macro()
    for a, b, n in ((rho_g, rho_o, "rho"), (u_g, u_o, "u"), (v_g, v_o, "v"), (w_g, w_o, "w")):
        assert rel_linf(a, b) <= 1e-12, n


def test_monitor_matches_oracle():
    geo, p = small_drainage(n=(24, 24, 32), seed=5, jper=1)
    say("case=monitor fp64")
    o, g = pair(geo, p)
    o.setup()
    g.setup()
    o.run(1, 20)
    g.run(1, 20)
    rho, u, v, w = o.macro(zero_solid_phi=False)
    ref = o.monitor_sums(rho, u, v, w)
    got = g.monitor()
    mag = {"fl1": np.abs(w).sum(), "fl2": np.abs(w).sum(), "vol1": rho.size, "vol2": rho.size,
           "mass1": np.abs(rho).sum(), "mass2": np.abs(rho).sum(), "pre": np.abs(rho).sum()}
    for k in ("fl1", "fl2", "vol1", "vol2", "mass1", "mass2", "pre"):
        err = np.abs(got[k] - ref[k]).max()
        assert err <= 1e-12 * mag[k], f"{k}: {err:.3e}"
    assert got["umax2"] == ref["umax2"]
    assert abs(got["usq1"] - ref["usq1"]) <= 1e-12 * max(1.0, abs(ref["usq1"]))
    assert abs(got["usq2"] - ref["usq2"]) <= 1e-12 * max(1.0, abs(ref["usq2"]))


def test_fp32_tracks_fp64():
    """FP32 populations/collision against the FP64 oracle: saturation and mass stay close."""
    geo, p = small_drainage(n=(24, 24, 32), seed=9, jper=1)
    say("case=fp32 vs fp64 oracle")
    o = O.Oracle(p, geo)
    g = gpu.Solver(p, geo, gpu.FP32)
    o.setup()
    g.setup()
    o.run(1, 200)
    g.run(1, 200)
    rho, u, v, w = o.macro(zero_solid_phi=False)
    ref = o.monitor_sums(rho, u, v, w)
    got = g.monitor()
    s_ref = ref["vol1"].sum() / (ref["vol1"].sum() + ref["vol2"].sum())
    s_got = got["vol1"].sum() / (got["vol1"].sum() + got["vol2"].sum())
    assert abs(s_got - s_ref) < 1e-3, (s_got, s_ref)
    m_ref = ref["mass1"].sum() + ref["mass2"].sum()
    m_got = got["mass1"].sum() + got["mass2"].sum()
    assert abs(m_got - m_ref) / m_ref < 1e-4
    # each colour's mass on its own: the FP32 reading of a colour-2 trace must neither
    # create nor remove fluid 2
    d1 = abs(got["mass1"].sum() / ref["mass1"].sum() - 1.0)
    d2 = abs(got["mass2"].sum() / ref["mass2"].sum() - 1.0)
    say(f"fp32 vs fp64 oracle after 200 steps: saturation {abs(s_got - s_ref):.2e}, mass {abs(m_got / m_ref - 1):.2e}, "
        f"mass1 {d1:.2e}, mass2 {d2:.2e}")
    assert d1 < 1e-4 and d2 < 1e-4, (d1, d2)
    assert rel_linf(g.download("phi"), o.array("phi")) < 0.05


def test_fp32_low_force_single_phase_velocity():
    """FP32 storage keeps the deviation of the total population from the rest state, so
    a flow of u ~ 1e-6 (body force 1e-6) is resolved, not lost under the float spacing of
    O(0.1) populations.  From the relaxed start (no start-up pulse), 600 steps."""
    from mflbm import warmstart
    n = 24
    geo = cases.geometry.sphere_pack((n, n, n), 3.0, 0.6, 4, periodic=(True, True, False))
    p = config.derive(cases.periodic_control(n=(n, n, n), body_force=1e-6, option=1, interface_z0=float(n + 10)))
    say("case=fp32 single phase, body force 1e-6, relaxed start, 600 steps")
    o = O.Oracle(p, geo)
    g = gpu.Solver(p, geo, gpu.FP32)
    rho0 = np.full(o.shape(1), warmstart.relaxed_mean_density(geo))
    o.init_phi(1, float(n + 10))
    g.init_phi(1, float(n + 10))
    o.init_pdf(rho0, relaxed=True)
    g.init_equilibrium(rho0, relaxed=True)
    o.color_gradient()
    g.color_gradient()
    o.run(1, 600)
    g.run(1, 600)
    rho_o, u_o, v_o, w_o = o.macro(zero_solid_phi=False)
    rho_g, u_g, v_g, w_g = g.macro()
    vo = np.stack([u_o, v_o, w_o])[:, 1:-1, 1:-1, 1:-1]
    vg = np.stack([u_g, v_g, w_g])[:, 1:-1, 1:-1, 1:-1]
    err = np.linalg.norm(vg - vo) / np.linalg.norm(vo)
    fl = rho_g != 0
    derr = np.abs(rho_g[fl] - rho_o[fl]).max()
    say(f"fp32 velocity rel L2 {err:.2e}, density max abs {derr:.2e}")
    assert err < 1e-4, err
    assert derr < 1e-8, derr


def test_fp32_pdf_round_trip():
    """The FP32 population field crosses the ABI in the two-colour layout: download,
    upload and download again returns the same values."""
    geo, p = CASES["drainage_yper"]()
    say("case=drainage_yper fp32 pdf download/upload")
    g = gpu.Solver(p, geo, gpu.FP32)
    g.setup()
    g.run(1, 10)
    a = g.download("pdf")
    g.upload("pdf", a)
    b = g.download("pdf")
    assert np.abs(b.astype(np.float64) - a).max() <= 2e-8


@pytest.mark.parametrize("name, kw", [
    ("velocity_in_convective_out_300", dict(n=(20, 18, 28), seed=5, jper=0, inlet=1, outlet=1, ca=0.1)),
    ("plate_at_inlet_plane_300", dict(n=(18, 20, 24), seed=3, jper=1, plate=2, plate_z=1)),
    ("plate_at_outlet_plane_300", dict(n=(18, 18, 24), seed=2, jper=0, inlet=1, outlet=1, plate=1, plate_z=24)),
])
def test_open_boundaries_300_steps_bitwise(name, kw):
    """Longer runs of the open boundaries through the step graphs: the injected front
    moves into the pack and the outlet state turns over."""
    geo, p = small_drainage(**kw)
    say(f"case={name} 300 steps")
    o, g = pair(geo, p)
    o.setup()
    g.setup()
    o.run(1, 300)
    g.run(1, 300)
    compare_state(o, g, tag=name)


def test_fp32_velocity_inlet_tracks_fp64():
    """FP32 storage with the velocity inlet and convective outlet: the injected saturation
    and the velocity field follow the FP64 run."""
    geo, p = small_drainage(n=(20, 18, 28), seed=5, jper=0, inlet=1, outlet=1, ca=0.1)
    say("fp32 vs fp64, velocity inlet + convective outlet, 400 steps")
    a = gpu.Solver(p, geo, gpu.FP64)
    b = gpu.Solver(p, geo, gpu.FP32)
    for s in (a, b):
        s.setup()
        s.run(1, 400)
    _, ua, va, wa = a.macro()
    _, ub, vb, wb = b.macro()
    err = np.sqrt(((ua - ub) ** 2 + (va - vb) ** 2 + (wa - wb) ** 2).sum() / (ua ** 2 + va ** 2 + wa ** 2).sum())
    ma, mb = a.monitor(), b.monitor()
    sa = ma["vol1"].sum() / (ma["vol1"].sum() + ma["vol2"].sum())
    sb = mb["vol1"].sum() / (mb["vol1"].sum() + mb["vol2"].sum())
    say(f"velocity rel L2 {err:.2e}, saturation {sa:.6f} vs {sb:.6f}")
    assert err < 1e-3 and abs(sa - sb) < 1e-4
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_velocity_change_matches_host_norms(precision):
    """mflbm_velocity_change against the same norms of mflbm_macro's velocity on the host."""
    from mflbm import warmstart
    geo, p, _ = warmstart.single_phase_case(3, 32)
    s = gpu.Solver(p, geo, precision)
    s.setup()
    assert s.velocity_change() == -1.0
    vel = lambda: np.stack(s.macro()[1:])
    u0 = vel()
    s.run(1, 60)
    u1 = vel()
    c = s.velocity_change()
    ref = np.linalg.norm(u1 - u0) / np.linalg.norm(u1)
    say(f"velocity change {c:.15e} host {ref:.15e}")
    assert c > 0 and abs(c - ref) <= 1e-10 * ref
    s.run(61, 40)
    u2 = vel()
    c = s.velocity_change()
    assert abs(c - np.linalg.norm(u2 - u1) / np.linalg.norm(u2)) <= 1e-10 * c
    s.upload("pdf", s.download("pdf"))
    assert s.velocity_change() == -1.0
    s.close()


@pytest.mark.parametrize("model", [1, 3, 4])
def test_collision_models_bitwise(model):
    """The reference's other collision models (its compile-time `mrt` switch: 1 MRT
    bounce-back optimised, 3 SRT, 4 MRT advection optimised), 40 steps of a drainage with
    unequal viscosities, bit for bit against the oracle."""
    geo, p = CASES["drainage_ywalls"]()
    p = dict(p, mrt=model)
    say(f"case=drainage_ywalls mrt={model}, 40 steps")
    o, g = pair(geo, p)
    o.setup()
    g.setup()
    o.run(1, 40)
    g.run(1, 40)
    compare_state(o, g, tag=f"mrt{model} step40")
