"""Pins the oracle's velocity inlet, convective outlet and porous plate against what the
physics fixes independently of their code:

- the inlet profile is the series solution of laminar flow in a rectangular duct: it solves
  the Poisson equation whose source the mean velocity fixes, has that mean, and vanishes
  at the duct walls;
- a single-phase duct driven by the velocity inlet and drained by the convective outlet
  carries the injected mass flux through every plane, with the inlet profile downstream;
- a porous plate stops the colour it blocks (zero flux through every plane) and lets the
  other colour through, which then streams as if the plate plane were not there (the flux
  of a pressure-driven duct grows by nz / (nz - 1)).

References: src/Misc.cpp:387-419, src/Init_multiphase.cpp:258-285 (profile),
src/main_iteration_GPU.cu:1008-1077 (inlet), :1245-1357 (outlet), :1744-1882 (plate).
"""
import math

import numpy as np
import pytest

import oracle as O
from mflbm import cases, config

N = (12, 12, 32)


def duct(inlet, outlet, ca=0.0033, plate=0, plate_z=0, body_force=0.0, n=N):
    """Empty duct with walls in x and y, all fluid 1 (interface beyond the outlet)."""
    ctrl = cases.drainage_control(n=n, interface_z0=n[2] + 10.0, jper=0, sa_inject=1.0, body_force=body_force)
    ctrl.update(inlet_BC=inlet, outlet_BC=outlet, capillary_number=ca, porous_plate_cmd=plate,
                Z_porous_plate=plate_z)
    p = config.derive(ctrl)
    o = O.Oracle(p, np.zeros(n[::-1], np.int8))
    o.setup()
    return o, p


def plane_flux(o):
    rho, _, _, w = o.macro(zero_solid_phi=False)
    return (rho * w)[1:-1, 1:-1, 1:-1].sum(axis=(1, 2))


def duct_constant(a, b, terms=1000):
    """1 - 192 a / (pi^5 b) sum_n odd tanh(n pi b / 2a) / n^5: mean / (G a^2 / 3 mu)."""
    s = sum(math.tanh(0.5 * n * math.pi * b / a) / n ** 5 for n in range(1, terms + 1, 2))
    return 1.0 - 192.0 / math.pi ** 5 * (a / b) * s


@pytest.mark.parametrize("n", [(12, 12, 8), (20, 12, 8), (10, 26, 8)])
def test_inlet_profile_is_the_duct_solution(n):
    o, p = duct(1, 1, ca=0.01, n=n)
    w = o.array("w_in")  # (Y1, X1)
    nx, ny = n[0], n[1]
    U = p["uin_avg"]
    assert U > 0
    # zero on the wall nodes x = 1, nx and y = 1, ny and on the ghost frame
    inner = np.zeros_like(w, bool)
    inner[2:ny, 2:nx] = True
    assert (w[~inner] == 0).all() and (w[inner] > 0).all()
    # symmetric about the duct centre
    core = w[2:ny, 2:nx]
    assert np.allclose(core, core[::-1, :], rtol=1e-12, atol=0) and np.allclose(core, core[:, ::-1], rtol=1e-12, atol=0)
    # Poisson equation: the five-point Laplacian of the profile at interior nodes equals
    # -3 U / (c a^2) up to the stencil's O(h^2) error on a smooth field
    a, b = 0.5 * (nx - 2), 0.5 * (ny - 2)
    src = -3.0 * U / (duct_constant(a, b) * a * a)
    lap = (w[3:ny - 1, 2:nx - 2] + w[3:ny - 1, 4:nx] + w[2:ny - 2, 3:nx - 1] + w[4:ny, 3:nx - 1]
           - 4 * w[3:ny - 1, 3:nx - 1])
    err = np.abs(lap / src - 1)
    assert err.max() < 0.04 and np.median(err) < 0.01, (err.max(), np.median(err))  # largest next to corners
    # the nodes sample the duct at unit-spaced cell centres: their mean is the midpoint rule
    # of the area mean U
    assert abs(core.mean() / U - 1) < 0.03


def test_velocity_inlet_and_convective_outlet_carry_the_injected_flux():
    o, p = duct(1, 1)
    w_in = o.array("w_in").copy()
    o.run(1, 3000)
    flux = plane_flux(o)
    q = w_in.sum()
    assert np.abs(flux / q - 1).max() < 0.01, flux / q
    # downstream of the inlet the axial velocity is the inlet profile again
    _, _, _, w = o.macro(zero_solid_phi=False)
    mid = w[1 + N[2] // 2, 1:-1, 1:-1]
    assert np.abs(mid - w_in[1:-1, 1:-1]).max() < 0.01 * w_in.max()
    # the convective outlet's stored ghost phi follows the (uniform) outlet phase
    conv = o.array("convec")
    assert np.allclose(conv[38, 1:-1, 1:-1], 1.0, rtol=0, atol=1e-12)


def test_porous_plate_stops_the_blocked_colour():
    o, _ = duct(2, 2, plate=1, plate_z=N[2] // 2, body_force=1e-4)
    o.run(1, 2000)
    flux = plane_flux(o)
    free = np.delete(flux, N[2] // 2 - 1)  # the plate plane itself holds reflected populations
    ref, _ = duct(2, 2, body_force=1e-4)
    ref.run(1, 2000)
    q = plane_flux(ref).mean()
    assert q > 1e-3
    assert np.abs(free).max() < 1e-8 * q, np.abs(free).max()


def test_porous_plate_passes_the_other_colour():
    zp = N[2] // 2
    plate, _ = duct(2, 2, plate=2, plate_z=zp, body_force=1e-4)
    plain, _ = duct(2, 2, body_force=1e-4)
    plate.run(1, 2000)
    plain.run(1, 2000)
    ratio = np.delete(plane_flux(plate), zp - 1).mean() / plane_flux(plain).mean()
    assert abs(ratio / (N[2] / (N[2] - 1)) - 1) < 0.005, ratio
