"""The random initial phase distribution (initial_fluid_distribution_option 6,
src/Init_multiphase.cpp:344-350): phi = 1, or -1 where the node's uniform deviate exceeds
target_fluid1_saturation.  The reference seeds rand() from time(NULL); this build draws a
counter-based deviate per node (splitmix64 of the seed and the node's s1 index).

Pinned by: the generator's published first outputs for seed 0, an independent numpy
restatement of it, the phase fraction against the target, reproducibility per seed."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from helpers import DRIVER, write_case
from mflbm import cases, config

M = np.uint64
GAMMA = M(0x9E3779B97F4A7C15)


def splitmix64(seed, n):
    """splitmix64 output of state seed + (n + 1) * gamma (numpy, wrapping uint64)"""
    with np.errstate(over="ignore"):
        z = M(seed) + (np.asarray(n, dtype=M) + M(1)) * GAMMA
        z = (z ^ (z >> M(30))) * M(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> M(27))) * M(0x94D049BB133111EB)
        return z ^ (z >> M(31))


def test_generator_matches_published_splitmix64_outputs():
    # splitmix64 (Steele, Lea, Flood 2014) from state 0: its first three outputs
    assert [int(v) for v in splitmix64(0, np.arange(3))] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4,
                                                            0x06C45D188009454F]
    lib = O.lib()
    for seed in (0, 1, 12345, 2**63 + 7):
        n = np.arange(0, 5000, 7, dtype=np.uint64)
        want = (splitmix64(seed, n) >> M(11)).astype(np.float64) * 2.0**-53
        got = np.array([lib.oracle_node_uniform(seed, int(k)) for k in n])
        assert np.array_equal(got, want), seed


def random_case(n=(12, 10, 14), sa=0.35, seed=5, open_z=True):
    ctrl = cases.drainage_control(n=n, interface_z0=3.0)
    ctrl.update(initial_fluid_distribution_option=6, target_fluid1_saturation=sa, random_seed=seed)
    if not open_z:
        ctrl.update(kper=1, inlet_BC=0, outlet_BC=0)
    nx, ny, nz = n
    return np.zeros((nz, ny, nx), np.int8), config.derive(ctrl)


@pytest.mark.parametrize("open_z", [True, False])
def test_oracle_option6_field(open_z):
    geo, p = random_case(open_z=open_z)
    nx, ny, nz = p["nx"], p["ny"], p["nz"]
    o = O.Oracle(p, geo)
    o.init_phi_random(p["sa_target"], p["random_seed"])
    phi = o.array("phi")  # s4 layout, (z, y, x) with 4 ghost layers
    # numpy restatement over [0, n+1]^3, x fastest, s1 index i + (nx+2)(j + (ny+2)k)
    k, j, i = np.meshgrid(np.arange(nz + 2), np.arange(ny + 2), np.arange(nx + 2), indexing="ij")
    idx = (i + (nx + 2) * (j + (ny + 2) * k)).astype(np.uint64)
    u = (splitmix64(p["random_seed"], idx) >> M(11)).astype(np.float64) * 2.0**-53
    want = np.where(u > p["sa_target"], -1.0, 1.0)
    if open_z:
        want[0] = p["phi_inlet"]  # the inlet ghost layers k <= 0 take phi_inlet
    assert np.array_equal(phi[3:nz + 5, 3:ny + 5, 3:nx + 5], want)
    if open_z:
        assert np.all(phi[:4] == p["phi_inlet"])


def test_option6_fraction_and_reproducibility():
    geo, p = random_case(n=(40, 40, 40), sa=0.3, open_z=False)
    o = O.Oracle(p, geo)
    o.init_phi_random(0.3, 11)
    a = o.array("phi")[4:44, 4:44, 4:44].copy()
    frac = float((a > 0).mean())
    assert abs(frac - 0.3) < 4 * np.sqrt(0.3 * 0.7 / a.size), frac
    o.init_phi_random(0.3, 11)
    assert np.array_equal(o.array("phi")[4:44, 4:44, 4:44], a)
    o.init_phi_random(0.3, 12)
    b = o.array("phi")[4:44, 4:44, 4:44]
    assert 0.3 < float((a != b).mean()) < 0.5  # independent draws differ at 2 s (1 - s) = 42 %
    o.init_phi_random(0.0, 11)
    assert np.all(o.array("phi")[4:44, 4:44, 4:44] == -1.0)
    o.init_phi_random(1.0, 11)
    assert np.all(o.array("phi")[4:44, 4:44, 4:44] == 1.0)


def test_driver_accepts_option6_and_rejects_others(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("mflbm_run is missing: build with __graft_entry__.build()")
    ctrl = cases.drainage_control(n=(12, 10, 14), interface_z0=3.0)
    geo = np.zeros((14, 10, 12), np.int8)
    for option, ok in ((6, True), (7, False), (0, False)):
        root = tmp_path / f"o{option}"
        c = dict(ctrl, initial_fluid_distribution_option=option, target_fluid1_saturation=0.4)
        write_case(str(root), c, geo)
        r = subprocess.run([DRIVER, "--check"], cwd=root, capture_output=True, text=True, timeout=60)
        assert (r.returncode == 0) == ok, r.stdout + r.stderr
        if not ok:
            assert "initial_fluid_distribution_option" in r.stderr
    root = tmp_path / "sa"
    write_case(str(root), dict(ctrl, initial_fluid_distribution_option=6, target_fluid1_saturation=1.5), geo)
    r = subprocess.run([DRIVER, "--check"], cwd=root, capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "target_fluid1_saturation" in r.stderr
