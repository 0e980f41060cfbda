"""Shared helpers of the test suite: workload construction and array comparison."""
import os

import numpy as np

from mflbm import cases, config


def small_drainage(n=(24, 24, 32), seed=7, radius=4.0, porosity=0.45, jper=1, theta=45.0, nu_ratio=1.0,
                   body_force=5e-4, buffer=3, inlet=2, outlet=2, ca=0.05, plate=0, plate_z=0):
    """Sphere-pack drainage case with empty inlet/outlet buffers (reference axis layout).
    inlet / outlet: 1 velocity inlet / convective outlet (mean inlet velocity Ca * gamma / nu1),
    2 Zou-He pressure; plate: porous plate blocking fluid 1 or 2 at z = plate_z."""
    nx, ny, nz = n
    nu2 = cases.nu_from_tau(0.7)
    ctrl = cases.drainage_control(n=n, theta=theta, nu1=nu_ratio * nu2, nu2=nu2, gamma=0.01,
                                  body_force=body_force, interface_z0=buffer + 1.0, jper=jper)
    ctrl.update(inlet_BC=inlet, outlet_BC=outlet, capillary_number=ca, porous_plate_cmd=plate,
                Z_porous_plate=plate_z)
    geo = cases.sphere_geometry(n, radius, porosity, seed, jper=jper)
    geo[:buffer] = 0
    geo[nz - buffer:] = 0
    return geo, config.derive(ctrl)


def periodic_drop(n=(24, 24, 24), radius=6.0, theta=45.0, geo_seed=None, body_force=0.0):
    """Droplet (option 5) in a y/z periodic box with x walls; optional sphere obstacles."""
    ctrl = cases.periodic_control(n=n, body_force=body_force, theta=theta, option=5, interface_z0=radius)
    nx, ny, nz = n
    if geo_seed is None:
        geo = np.zeros((nz, ny, nx), np.int8)
    else:
        geo = cases.geometry.sphere_pack((nz, ny, nx), 3.0, 0.85, geo_seed, periodic=(True, True, False))
    return geo, config.derive(ctrl)


def rel_linf(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(np.abs(b).max(), 1e-300)
    return float(np.abs(a - b).max() / scale)


def assert_same(a, b, tol=0.0, what=""):
    a = np.asarray(a)
    b = np.asarray(b)
    assert a.shape == b.shape, f"{what}: shape {a.shape} != {b.shape}"
    if tol == 0.0:
        bad = ~((a == b) | (np.isnan(a) & np.isnan(b)))
        if bad.any():
            idx = np.argwhere(bad)[:5]
            raise AssertionError(f"{what}: {bad.sum()} of {a.size} entries differ; first {idx.tolist()}; "
                                 f"rel Linf {rel_linf(a, b):.3e}")
    else:
        r = rel_linf(a, b)
        assert r <= tol, f"{what}: relative Linf {r:.3e} > {tol:.1e}"


DRIVER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mf-lbm-cuda_amd", "mflbm_run")


def write_case(root, ctrl, geo, status="new_simulation"):
    """Lay out a run directory the way the reference expects it: input/job_status.txt,
    input/simulation_control.txt, input/Geometry_File_Path.txt and the .dat geometry."""
    from mflbm import geometry
    os.makedirs(os.path.join(root, "input", "geometry"), exist_ok=True)
    config.write_control(os.path.join(root, "input", "simulation_control.txt"), ctrl)
    geometry.write_dat(os.path.join(root, "input", "geometry", "sample1.dat"), geo)
    with open(os.path.join(root, "input", "Geometry_File_Path.txt"), "w") as fh:
        fh.write("geo_file_path_\tinput/geometry/sample1\n\ngeo_boundary_file_path_\tinput/geometry/sample1\n")
    with open(os.path.join(root, "input", "job_status.txt"), "w") as fh:
        fh.write(status)  # no newline: the reference compares the whole file
