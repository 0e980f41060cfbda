"""The reference's own input files (input/simulation_control.txt, input/job_status.txt,
input/geometry/tube_sphere.dat; there are no other fixtures in the reference) through this
repository's readers and the C++ driver's input checks, on the CPU.

The geometry and the parsed control values are committed under tests/golden/ (written by
tests/golden/make_reference_inputs.py); the tests that read the reference tree itself skip
where it is absent."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from helpers import DRIVER
from mflbm import config, geometry

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF = "/root/reference"
REF_INPUT = os.path.join(REF, "input")
needs_ref = pytest.mark.skipif(not os.path.isdir(REF_INPUT), reason="reference tree not present")


def reference_control():
    with open(os.path.join(GOLDEN, "reference_control.json")) as fh:
        return json.load(fh)


def test_tube_sphere_fixture_reads_as_a_60x60x80_sample():
    geo = geometry.read_dat(os.path.join(GOLDEN, "tube_sphere.dat"))
    assert geo.shape == (80, 60, 60)
    assert set(np.unique(geo)) <= {0, 1}
    assert 0.1 < geo.mean() < 0.4


@needs_ref
def test_fixtures_match_the_reference_files():
    with open(os.path.join(REF_INPUT, "geometry", "tube_sphere.dat"), "rb") as a, \
            open(os.path.join(GOLDEN, "tube_sphere.dat"), "rb") as b:
        assert a.read() == b.read()
    assert config.read_control(os.path.join(REF_INPUT, "simulation_control.txt")) == reference_control()


def test_reference_control_derives_the_reference_constants():
    """src/IO_multiphase.cpp:200-206 and src/Init_multiphase.cpp:173-214 on the example
    control: pressure inlet/outlet on an open z axis, rho_out = 1 and
    rho_in = rho_out + body_force_0 / 3 * nz, force 0, phi_inlet = 2 S_inj - 1 = -1,
    cos(180 - 65 degrees)."""
    c = reference_control()
    p = config.derive(c)
    assert p["force_z"] == 0.0 and p["rho_out"] == 1.0
    assert p["rho_in"] == 1.0 - (-1e-4 / 3.0) * 260.0
    assert p["phi_inlet"] == -1.0
    assert abs(p["cos_theta"] - np.cos(np.radians(115.0))) < 1e-15
    assert p["la_nui1"] == 1.0 / 0.04 and p["la_nui2"] == 1.0 / 0.4
    assert (p["nx"], p["ny"], p["nz"]) == (240, 240, 260)
    assert p["porous_plate"] == 2 and p["z_porous_plate"] == 252


@needs_ref
@pytest.mark.skipif(not os.path.exists(DRIVER), reason="mflbm_run is built by __graft_entry__.build()")
def test_driver_accepts_the_reference_input_files(tmp_path):
    """mflbm_run --check on the reference's control and job-status files, unchanged, with
    the shipped geometry.  The geometry path rule (src/Misc.cpp:31-33) strips trailing
    non-digits and appends .dat, so the file must be named with a trailing digit, as the
    reference's own example path (bentheimer_in10_240_240_240_out10) is."""
    inp = tmp_path / "input"
    (inp / "geometry").mkdir(parents=True)
    shutil.copyfile(os.path.join(REF_INPUT, "simulation_control.txt"), inp / "simulation_control.txt")
    shutil.copyfile(os.path.join(REF_INPUT, "job_status.txt"), inp / "job_status.txt")
    shutil.copyfile(os.path.join(GOLDEN, "tube_sphere.dat"), inp / "geometry" / "tube_sphere1.dat")
    (inp / "Geometry_File_Path.txt").write_text("geo_file_path_\tinput/geometry/tube_sphere1\n\n"
                                                "geo_boundary_file_path_\tinput/geometry/tube_sphere1\n")
    r = subprocess.run([DRIVER, "--check"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Everything looks good!" in r.stdout
    assert "Porous media sample size: nx = 60, ny = 60, nz = 80" in r.stdout
