"""Named workloads (geometry + control values) used by the tests and the benchmark.

Axis convention: the reference drives flow along its z axis (inlet k=1, outlet k=nz,
src/main_iteration_GPU.cu:1930) and requires walls on both x faces
(src/IO_multiphase.cpp:210-213).  A workload stated as "flow along x, periodic y/z" is
therefore laid out with its flow axis on the reference z axis, one transverse axis on the
periodic reference y axis and the other on the walled reference x axis.
"""
from __future__ import annotations

import math

import numpy as np

from . import config, geometry


def nu_from_tau(tau: float) -> float:
    return (tau - 0.5) / 3.0


def drainage_control(n=(64, 64, 64), theta=45.0, nu1=None, nu2=None, gamma=0.01, beta=0.95,
                     body_force=5e-4, interface_z0=4.0, jper=1, option=1, sa_inject=1.0):
    """Colour-gradient drainage: fluid 1 injected through a Zou-He pressure inlet at k=1."""
    nx, ny, nz = n
    c = config.default_control()
    tau_nu = nu_from_tau(0.7)
    c.update(dict(
        nxGlobal=nx, nyGlobal=ny, nzGlobal=nz, initial_fluid_distribution_option=option,
        domain_wall_status_x_min=1, domain_wall_status_x_max=1,
        domain_wall_status_y_min=0 if jper else 1, domain_wall_status_y_max=0 if jper else 1,
        domain_wall_status_z_min=0, domain_wall_status_z_max=0, iper=0, jper=jper, kper=0,
        inlet_BC=2, outlet_BC=2, fluid1_viscosity=tau_nu if nu1 is None else nu1,
        fluid2_viscosity=tau_nu if nu2 is None else nu2, surface_tension=gamma, theta=theta, RK_beta=beta,
        saturation_injection=sa_inject, initial_interface_position=interface_z0, body_force_0=body_force,
        n_exclude_inlet=min(10, nz // 8), n_exclude_outlet=min(10, nz // 8),
    ))
    return c


def periodic_control(n=(64, 64, 64), body_force=1e-6, theta=45.0, nu=None, gamma=0.01, option=5,
                     interface_z0=10.0):
    """Periodic in y and z, walls in x, body force along z (single phase when phi = 1)."""
    nx, ny, nz = n
    c = config.default_control()
    nu = nu_from_tau(0.7) if nu is None else nu
    c.update(dict(
        nxGlobal=nx, nyGlobal=ny, nzGlobal=nz, initial_fluid_distribution_option=option,
        domain_wall_status_x_min=1, domain_wall_status_x_max=1, domain_wall_status_y_min=0,
        domain_wall_status_y_max=0, domain_wall_status_z_min=0, domain_wall_status_z_max=0,
        iper=0, jper=1, kper=1, fluid1_viscosity=nu, fluid2_viscosity=nu, surface_tension=gamma,
        theta=theta, body_force_0=body_force, initial_interface_position=interface_z0,
        n_exclude_inlet=0, n_exclude_outlet=0,
    ))
    return c


def sphere_geometry(n, radius, porosity, seed, jper=1, buffer=0):
    """Sphere pack in the reference layout; ``buffer`` empty layers at the inlet and outlet
    ends of the flow axis (the reference's own sample geometries carry such reservoirs)."""
    nx, ny, nz = n
    geo = geometry.sphere_pack((nz - 2 * buffer, ny, nx), radius, porosity, seed, periodic=(False, bool(jper), False))
    if buffer:
        geo = np.concatenate([np.zeros((buffer, ny, nx), np.int8), geo, np.zeros((buffer, ny, nx), np.int8)])
    return geo


def drainage64(seed=1):
    """64^3 FP64 drainage: spheres r=6, porosity 0.40, theta 45, tau 0.7, gamma 0.01."""
    n = (64, 64, 64)
    ctrl = drainage_control(n=n, theta=45.0, gamma=0.01, body_force=5e-4, interface_z0=4.0)
    geo = sphere_geometry(n, 6.0, 0.40, seed, buffer=4)
    return geo, ctrl


def drainage256(seed=2, n=(256, 256, 256), geometry=True):
    """256^3 drainage: spheres r=8, porosity 0.35, theta 30, viscosity ratio 0.1.
    The invading fluid 1 has tau 0.7 and the defending fluid 2 ten times its viscosity
    (tau 2.5); scaling fluid 1 down to a tenth instead (tau 0.52) is unstable and the run
    turns to NaN within a few hundred steps, in the oracle as on the device.
    ``geometry=False`` returns (None, control) without building the sphere pack."""
    nu1 = nu_from_tau(0.7)
    nu2 = 10.0 * nu1  # viscosity ratio nu1 / nu2 = 0.1, fluid 1 invading
    ctrl = drainage_control(n=n, theta=30.0, nu1=nu1, nu2=nu2, gamma=0.01, body_force=2e-4,
                            interface_z0=8.0)
    geo = sphere_geometry(n, 8.0, 0.35, seed, buffer=8) if geometry else None
    return geo, ctrl


def grf_drainage512(seed=3, n=(256, 256, 512), geometry=True):
    """Precision-study drainage, 512 nodes along the flow axis and 256 x 256 across: seeded
    Gaussian-random-field medium (correlation length 10, porosity 0.30) with 8 empty layers
    at the inlet and outlet; the physics of the 256^3 workload (theta 30, viscosity ratio
    0.1 with the invading fluid at tau 0.7, gamma 0.01)."""
    nx, ny, nz = n
    nu1 = nu_from_tau(0.7)
    ctrl = drainage_control(n=n, theta=30.0, nu1=nu1, nu2=10.0 * nu1, gamma=0.01, body_force=2e-4,
                            interface_z0=8.0)
    geo = None
    if geometry:
        geo = geometry_grf((nz, ny, nx), 10.0, 0.30, seed)
        geo[:8] = 0
        geo[nz - 8:] = 0
    return geo, ctrl


def geometry_grf(shape, corr_len, porosity, seed):
    return geometry.gaussian_random_field(shape, corr_len, porosity, seed)


def params_from_control(ctrl: dict) -> dict:
    return config.derive(ctrl)
