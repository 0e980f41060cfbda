"""The C++ host driver on the GPU: a drainage run from the reference's input files, its
monitor series, VTK files and checkpoint, and a restart from that checkpoint that
continues bit for bit where an uninterrupted run would be."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import DRIVER, write_case
from mflbm import cases, config, gpu

pytestmark = pytest.mark.gpu

N = (24, 20, 32)


def say(msg):
    print(f"[gpu-driver] {msg}", flush=True)
    sys.stdout.flush()


def case(max_step):
    nx, ny, nz = N
    ctrl = cases.drainage_control(n=N, interface_z0=4.0)
    ctrl.update(max_time_step=max_step, monitor_timer=50, monitor_profile_timer_ratio=2, ntime_visual=100,
                ntime_animation=100, computation_time_timer=50, display_steps_timer=100, n_exclude_inlet=3,
                n_exclude_outlet=3, steady_state_option=0, checkpoint_save_timer=1e3, checkpoint_2rd_save_timer=1e3,
                simulation_duration_timer=1e3)
    geo = cases.sphere_geometry(N, 3.0, 0.6, 5, buffer=3)
    return ctrl, geo


def run_driver(root, *args):
    r = subprocess.run([DRIVER, *args], cwd=root, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def read_checkpoint(path, es=8):
    raw = open(path, "rb").read()
    nt = int(np.frombuffer(raw[:4], np.int32)[0])
    body = np.frombuffer(raw[4:], np.float64 if es == 8 else np.float32)
    return nt, body


def reference_run(ctrl, geo, steps):
    p = config.derive(ctrl)
    s = gpu.Solver(p, geo, gpu.FP64)
    s.setup()
    s.run(1, steps)
    return s


def test_driver_run_outputs_and_restart(tmp_path):
    if not os.path.exists(DRIVER):
        pytest.fail("mflbm_run is missing: build with __graft_entry__.build()")
    root = str(tmp_path)
    ctrl, geo = case(199)  # steps 1..200
    write_case(root, ctrl, geo)
    say("driver: new simulation, 200 steps")
    out = run_driver(root)
    assert "Simulation ended after\t200 iterations" in out
    assert open(os.path.join(root, "job_status.txt")).read() == "simulation_reached_max_step\n"
    sat = np.loadtxt(os.path.join(root, "results/out1.output/saturation.dat"))
    assert sat.shape == (4, 6) and list(sat[:, 0]) == [50, 100, 150, 200]
    for name in ("Ca_number.dat", "flowrate_time.dat", "saturation_full_domain.dat", "pre.dat", "time.dat", "info.txt"):
        assert os.path.exists(os.path.join(root, "results/out1.output", name)), name
    assert os.path.exists(os.path.join(root, "results/out1.output/profile/monitor 00000100"))
    nx, ny, nz = N
    n = nx * ny * nz
    full = open(os.path.join(root, "results/out3.field_data/full_flow_field/full_flow_field_0000000200.vtk"), "rb").read()
    assert full.startswith(b"# vtk DataFile Version 3.0\nvtk output\nBINARY\nDATASET STRUCTURED_POINTS\n")
    assert full.count(b"SCALARS") == 5 and len(full) > 5 * 4 * n
    small = open(os.path.join(root, "results/out3.field_data/phase_distribution/small_0000000100.vtk"), "rb").read()
    header_end = small.index(b"LOOKUP_TABLE default\n") + len(b"LOOKUP_TABLE default\n")
    phi_vtk = np.frombuffer(small[header_end:], ">f4")
    assert phi_vtk.size == n
    walls = open(os.path.join(root, "results/out3.field_data/walls_bin.vtk"), "rb").read()
    assert walls.count(b"SCALARS walls int") == 1

    # the checkpoint holds the state of the Python-driven run after the same 200 steps
    say("reference: python API, 200 steps")
    ref = reference_run(ctrl, geo, 200)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    assert nt == 201
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    assert np.array_equal(body[2 + 38 * s1:], ref.download("phi").ravel())
    # saturation of the last monitor = the device monitor of the same state
    m = ref.monitor()
    k0, k1 = 3, nz - 3
    s_ref = m["vol1"][k0:k1].sum() / (m["vol1"][k0:k1].sum() + m["vol2"][k0:k1].sum())
    assert abs(sat[-1, 1] - s_ref) < 1e-5

    # restart: 100 more steps, bitwise equal to an uninterrupted 300-step run
    ctrl2, _ = case(99)
    write_case(root, ctrl2, geo, status="continue_simulation")
    say("driver: continue_simulation, steps 201..300")
    out = run_driver(root)
    assert "Simulation ended after\t300 iterations" in out
    ref.run(201, 100)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 301
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    ref.close()


def test_driver_fp32_run_and_cross_precision_restart(tmp_path):
    root = str(tmp_path)
    ctrl, geo = case(99)
    write_case(root, ctrl, geo)
    say("driver: fp32 new simulation, 100 steps")
    run_driver(root, "--precision", "fp32")
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"), es=4)
    assert nt == 101 and np.isfinite(body).all()
    write_case(root, ctrl, geo, status="continue_simulation")
    say("driver: fp64 continues the fp32 checkpoint")
    out = run_driver(root, "--precision", "fp64")
    assert "Simulation ended after\t200 iterations" in out
    sat = np.loadtxt(os.path.join(root, "results/out1.output/saturation.dat"))
    assert np.isfinite(sat).all() and sat[-1, 0] == 200


@pytest.mark.parametrize("option, series", [(1, "steady_monitor_capillary_pressure_error.dat"),
                                            (2, "steady_monitor_max_phi_change.dat")])
def test_driver_steady_state_monitors(tmp_path, option, series):
    """A droplet at rest in a periodic box relaxes to a steady state that the capillary-
    pressure (1) and phase-field (2) monitors detect."""
    root = str(tmp_path)
    n = (24, 24, 24)
    ctrl = cases.periodic_control(n=n, body_force=0.0, option=5, interface_z0=6.0)
    ctrl.update(steady_state_option=option, convergence_criteria=1e-3, max_time_step=20000, monitor_timer=200,
                monitor_profile_timer_ratio=100, ntime_visual=100000, ntime_animation=100000,
                computation_time_timer=1000, display_steps_timer=1000, n_exclude_inlet=0, n_exclude_outlet=0,
                checkpoint_save_timer=1e3, checkpoint_2rd_save_timer=1e3, simulation_duration_timer=1e3)
    write_case(root, ctrl, np.zeros((24, 24, 24), np.int8))
    say(f"driver: droplet, steady_state_option {option}")
    out = run_driver(root)
    assert "converged" in out, out[-1500:]
    assert open(os.path.join(root, "job_status.txt")).read() == "simulation_done\n"
    rows = np.loadtxt(os.path.join(root, "results/out1.output", series), ndmin=2)
    assert rows[-1, 1] < 1e-3 and rows[-1, 0] < 20000
