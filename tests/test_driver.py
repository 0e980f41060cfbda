"""The C++ host driver (mf-lbm-cuda_amd/host/driver.cpp): input-file handling that runs
before any device call (`--check`), on the CPU."""
import os
import subprocess

import numpy as np
import pytest

from helpers import DRIVER, write_case
from mflbm import cases

pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER), reason="mflbm_run is built by __graft_entry__.build()")


def small_case():
    ctrl = cases.drainage_control(n=(16, 16, 24), interface_z0=4.0)
    ctrl.update(max_time_step=10, monitor_timer=10, ntime_visual=10, ntime_animation=10,
                computation_time_timer=10, display_steps_timer=10, n_exclude_inlet=2, n_exclude_outlet=2)
    geo = np.zeros((24, 16, 16), np.int8)
    geo[8:16, 6:10, 6:10] = 1
    return ctrl, geo


def check(root):
    return subprocess.run([DRIVER, "--check"], cwd=root, capture_output=True, text=True, timeout=60)


def test_valid_inputs_pass(tmp_path):
    ctrl, geo = small_case()
    write_case(str(tmp_path), ctrl, geo)
    r = check(tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Everything looks good!" in r.stdout
    assert "sample size: nx = 16, ny = 16, nz = 24" in r.stdout


@pytest.mark.parametrize("change, message", [
    ({"theta": 100.0}, "contact angle is larger than 90 degrees"),
    ({"iper": 1}, "X direction periodic BC enabled"),
    ({"jper": 0}, "non-slip BC not applied at y"),
    ({"porous_plate_cmd": 3}, "porous_plate_cmd must be 0, 1 or 2"),
    ({"inlet_BC": 2, "outlet_BC": 1}, "Inlet pressure + outlet convective BC is not supported"),
    ({"geometry_preprocess_cmd": 1}, "geometry_preprocessing_load is not implemented yet"),
])
def test_rejected_controls(tmp_path, change, message):
    ctrl, geo = small_case()
    ctrl.update(change)
    write_case(str(tmp_path), ctrl, geo)
    r = check(tmp_path)
    assert r.returncode == 1
    assert message in r.stderr


def test_job_status_is_compared_whole(tmp_path):
    ctrl, geo = small_case()
    write_case(str(tmp_path), ctrl, geo, status="new_simulation\n")
    r = check(tmp_path)
    assert r.returncode == 1 and "Wrong simlation status" in r.stderr


def test_missing_key_is_an_error(tmp_path):
    ctrl, geo = small_case()
    write_case(str(tmp_path), ctrl, geo)
    path = os.path.join(tmp_path, "input", "simulation_control.txt")
    lines = [ln for ln in open(path) if not ln.startswith("surface_tension")]
    open(path, "w").writelines(lines)
    r = check(tmp_path)
    assert r.returncode == 1 and "(surface_tension)" in r.stderr


def test_geometry_larger_than_domain_is_an_error(tmp_path):
    ctrl, geo = small_case()
    write_case(str(tmp_path), ctrl, np.zeros((30, 16, 16), np.int8))
    r = check(tmp_path)
    assert r.returncode == 1 and "Domain size is smaller" in r.stderr


@pytest.mark.parametrize("model, rc", [("1", 0), ("3", 0), ("4", 0), ("0", 2), ("5", 2), ("x", 2)])
def test_collision_model_option(tmp_path, model, rc):
    """--mrt takes the reference's compile-time collision models 1-4; anything else is a
    usage error before any input is read."""
    ctrl, geo = small_case()
    write_case(str(tmp_path), ctrl, geo)
    r = subprocess.run([DRIVER, "--mrt", model, "--check"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == rc, r.stdout + r.stderr
    if rc:
        assert "usage: mflbm_run" in r.stdout
