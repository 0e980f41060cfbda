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


def test_driver_velocity_inlet_convective_outlet_and_restart(tmp_path):
    """Velocity inlet + convective outlet through the driver: the injected-volume timers, a
    checkpoint that carries the outlet state, and a bitwise restart from it."""
    root = str(tmp_path)
    nx, ny, nz = N
    ctrl, geo = case(99)
    ctrl.update(inlet_BC=1, outlet_BC=1, capillary_number=0.05, target_inject_pore_volume=0.1,
                d_vol_monitor=0.02, d_vol_animation=-1.0, d_vol_detail=-1.0)
    write_case(root, ctrl, geo)
    say("driver: velocity inlet + convective outlet, injected-volume timers")
    out = run_driver(root)
    pore_sum = int(out.split("Total number of pore nodes = ")[1].split()[0])
    uin = 0.05 * ctrl["surface_tension"] / ctrl["fluid1_viscosity"]
    flowrate = uin * ((nx - 2.0) * (ny - 2.0))

    def even(t):
        return t + 1 if t % 2 == 1 else t

    ntime_max, t_mon = even(int(0.1 * pore_sum / flowrate)), even(int(0.02 * pore_sum / flowrate))
    assert f"ntime_max = {ntime_max}" in out and f"Monitor timer is modified based on injected volume: {t_mon}" in out
    last = 1 + ntime_max
    assert f"Simulation ended after\t{last} iterations" in out
    sat = np.loadtxt(os.path.join(root, "results/out1.output/saturation.dat"), ndmin=2)
    assert list(sat[:, 0]) == list(range(t_mon, last + 1, t_mon))

    say(f"reference: python API, {last} steps")
    p = config.derive(ctrl)
    ref = gpu.Solver(p, geo, gpu.FP64)
    ref.setup()
    ref.run(1, last)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    s1, s4, tail = (nx + 2) * (ny + 2) * (nz + 2), (nx + 8) * (ny + 8) * (nz + 8), 39 * (nx + 2) * (ny + 2)
    assert nt == last + 1 and body.size == 2 + 38 * s1 + s4 + tail
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    assert np.array_equal(body[2 + 38 * s1 + s4:], ref.download("convec").ravel())

    ctrl2 = dict(ctrl, target_inject_pore_volume=-1.0, max_time_step=99)
    write_case(root, ctrl2, geo, status="continue_simulation")
    say("driver: continue_simulation, 100 more steps")
    out = run_driver(root)
    assert f"Simulation ended after\t{last + 100} iterations" in out
    ref.run(last + 1, 100)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == last + 101
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    assert np.array_equal(body[2 + 38 * s1 + s4:], ref.download("convec").ravel())
    ref.close()


def test_driver_porous_plate_run(tmp_path):
    """porous_plate_cmd / Z_porous_plate from the control file reach the step: the run's
    checkpoint equals the Python API run of the same plate."""
    root = str(tmp_path)
    nx, ny, nz = N
    ctrl, geo = case(59)
    ctrl.update(porous_plate_cmd=1, Z_porous_plate=nz - 2)
    write_case(root, ctrl, geo)
    say("driver: porous plate blocking fluid 1 near the outlet, 60 steps")
    run_driver(root)
    ref = gpu.Solver(config.derive(ctrl), geo, gpu.FP64)
    ref.setup()
    ref.run(1, 60)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    assert nt == 61 and np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    ref.close()


def test_driver_modified_geometry_and_inlet_phase_switch(tmp_path):
    """modify_geometry_cmd = 1 on a duct (no geometry file) and change_inlet_fluid_phase_cmd = 2
    on the resumed run, against the same edits made through the Python API."""
    root = str(tmp_path)
    n = (24, 24, 40)
    nx, ny, nz = n
    ctrl = cases.drainage_control(n=n, interface_z0=6.0, jper=0)
    ctrl.update(max_time_step=59, monitor_timer=20, monitor_profile_timer_ratio=100, ntime_visual=1000,
                ntime_animation=1000, computation_time_timer=20, display_steps_timer=20, n_exclude_inlet=2,
                n_exclude_outlet=2, steady_state_option=0, checkpoint_save_timer=1e3, checkpoint_2rd_save_timer=1e3,
                simulation_duration_timer=1e3, modify_geometry_cmd=1, external_geometry_read_cmd=0)
    write_case(root, ctrl, np.zeros(n[::-1], np.int8))
    say("driver: modified duct geometry, 60 steps")
    out = run_driver(root)
    assert "Internal geometry modified!" in out
    # the sample edit of src/Misc.cpp:105-136: sphere r = ny/4 at the centre, tube r = ny/2
    # outside 10-layer reservoirs
    k, j, i = np.meshgrid(np.arange(1, nz + 1), np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    d2 = (i - 0.5 * (nx + 1)) ** 2 + (j - 0.5 * (ny + 1)) ** 2 + (k - 0.5 * (nz + 1)) ** 2
    geo = ((d2 < (0.25 * ny) ** 2) | ((d2 > (0.5 * ny) ** 2) & (k > 10) & (k < nz - 10 + 1))).astype(np.int8)
    p = config.derive(ctrl)
    ref = gpu.Solver(p, geo, gpu.FP64)
    ref.setup()
    ref.run(1, 60)
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 61 and np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())

    write_case(root, dict(ctrl, change_inlet_fluid_phase_cmd=2, max_time_step=39), np.zeros(n[::-1], np.int8),
               status="continue_simulation")
    say("driver: resume with change_inlet_fluid_phase_cmd = 2, 40 steps")
    out = run_driver(root)
    assert "Inlet fluid phase was changed (1 to fluid1; 2 to fluid2): 2" in out
    pdf = ref.download("pdf")  # (2, 19, nz + 2, ny + 2, nx + 2)
    below = np.arange(nz + 2) <= 6.0
    sl = (slice(None), below, slice(None), slice(0, nx))  # i = 0 .. nx - 1 (src/Misc.cpp:283-285)
    pdf[1][sl] += pdf[0][sl]
    pdf[0][sl] = 0.0
    ref.upload("pdf", pdf)
    ref.color_gradient()
    ref.run(61, 40)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 101 and np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    ref.close()


def test_driver_breakthrough_ends_the_run(tmp_path):
    """breakthrough_check 1: the run stops at the first monitor step at which fluid 1
    (phi > 0) holds a fluid node of the plane nz - 1 (src/Monitor.cpp:447-468), saves the
    state and reports simulation_done (src/main.cpp:296-312).  The Python API run of the
    same case has no fluid 1 on that plane one monitor interval earlier."""
    root = str(tmp_path)
    n = (16, 16, 32)
    nx, ny, nz = n
    ctrl = cases.drainage_control(n=n, interface_z0=float(nz - 8), body_force=5e-4)
    ctrl.update(breakthrough_check=1, steady_state_option=0, max_time_step=6000, monitor_timer=20,
                monitor_profile_timer_ratio=1000, ntime_visual=100000, ntime_animation=100000,
                computation_time_timer=1000, display_steps_timer=1000, n_exclude_inlet=3, n_exclude_outlet=3,
                checkpoint_save_timer=1e3, checkpoint_2rd_save_timer=1e3, simulation_duration_timer=1e3)
    geo = cases.sphere_geometry(n, 3.0, 0.8, 5, buffer=3)
    write_case(root, ctrl, geo)
    say("driver: drainage to breakthrough")
    out = run_driver(root)
    assert "Breakthrough point reached! Exiting program!" in out
    t = int(out.split("Simulation ended successfully after\t")[1].split()[0])
    say(f"breakthrough at step {t}")
    assert t % 20 == 0 and 20 < t < 6000
    assert open(os.path.join(root, "job_status.txt")).read() == "simulation_done\n"
    assert os.path.exists(os.path.join(root, f"results/out3.field_data/phase_distribution/small_{t:010d}.vtk"))

    ref = reference_run(ctrl, geo, t - 20)
    walls = ref.download("walls")[nz, 2:ny + 2, 2:nx + 2] == 0  # plane nz - 1 (two ghost layers)

    def fluid1_on_plane():
        return int((walls & (ref.download("phi")[nz + 2, 4:ny + 4, 4:nx + 4] > 0.0)).sum())

    assert fluid1_on_plane() == 0
    ref.run(t - 19, 20)
    assert fluid1_on_plane() >= 1
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == t + 1
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    ref.close()


def test_driver_random_initial_distribution(tmp_path):
    """initial_fluid_distribution_option 6 from the control file with this build's
    random_seed key: the run equals the Python API run from the same seeded field, and the
    initial saturation is the target's."""
    root = str(tmp_path)
    n = (16, 16, 20)
    nx, ny, nz = n
    ctrl = cases.periodic_control(n=n, body_force=1e-5, option=6, interface_z0=0.0)
    ctrl.update(target_fluid1_saturation=0.3, random_seed=77, max_time_step=99, monitor_timer=50,
                monitor_profile_timer_ratio=100, ntime_visual=100000, ntime_animation=100000,
                computation_time_timer=50, display_steps_timer=100, n_exclude_inlet=0, n_exclude_outlet=0,
                steady_state_option=0, checkpoint_save_timer=1e3, checkpoint_2rd_save_timer=1e3,
                simulation_duration_timer=1e3)
    geo = np.zeros((nz, ny, nx), np.int8)
    write_case(root, ctrl, geo)
    say("driver: option 6, seed 77")
    out = run_driver(root)
    s0 = float(out.split("Initial saturation: ")[1].split()[0])
    assert abs(s0 - 0.3) < 0.03, s0
    ref = reference_run(ctrl, geo, 100)
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 101
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    assert np.array_equal(body[2 + 38 * s1:], ref.download("phi").ravel())
    ref.close()


def test_driver_checkpoint_and_restart_at_256(tmp_path):
    """Config 2's 256^3 drainage through the driver at its real size: 40 steps with the full
    flow-field VTK, then a restart for 20 more; both checkpoints (5.2 GB of populations) equal
    the Python-driven state bit for bit."""
    root = str(tmp_path)
    geo, ctrl = cases.drainage256()
    nx, ny, nz = int(ctrl["nxGlobal"]), int(ctrl["nyGlobal"]), int(ctrl["nzGlobal"])
    timers = dict(monitor_timer=20, monitor_profile_timer_ratio=2, ntime_visual=40, ntime_animation=40,
                  computation_time_timer=20, display_steps_timer=20, n_exclude_inlet=8, n_exclude_outlet=8,
                  steady_state_option=0, checkpoint_save_timer=1e3, checkpoint_2rd_save_timer=1e3,
                  simulation_duration_timer=1e3)
    ctrl = dict(ctrl, max_time_step=39, **timers)
    write_case(root, ctrl, geo)
    say("driver: 256^3 new simulation, 40 steps")
    run_driver(root)
    vtk = os.path.join(root, "results/out3.field_data/full_flow_field/full_flow_field_0000000040.vtk")
    assert os.path.getsize(vtk) > 5 * 4 * nx * ny * nz
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    say("reference: python API, 40 steps")
    ref = reference_run(dict(ctrl), geo, 40)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 41
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    assert np.array_equal(body[2 + 38 * s1:], ref.download("phi").ravel())
    del body
    write_case(root, dict(ctrl, max_time_step=19), geo, status="continue_simulation")
    say("driver: 256^3 continue_simulation, steps 41..60")
    out = run_driver(root)
    assert "Simulation ended after\t60 iterations" in out
    ref.run(41, 20)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 61
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    ref.close()


def test_driver_collision_model_option(tmp_path):
    """--mrt 3 (the reference's SRT build) runs the SRT collision: the checkpoint equals the
    API-driven state with mrt = 3 bit for bit, and differs from the default model's."""
    root = str(tmp_path)
    ctrl, geo = case(49)  # steps 1..50
    write_case(root, ctrl, geo)
    say("driver: --mrt 3, 50 steps")
    run_driver(root, "--mrt", "3")
    nx, ny, nz = N
    s1 = (nx + 2) * (ny + 2) * (nz + 2)
    nt, body = read_checkpoint(os.path.join(root, "results/out2.checkpoint/id0000"))
    assert nt == 51
    p = config.derive(ctrl)
    ref = gpu.Solver(dict(p, mrt=3), geo, gpu.FP64)
    ref.setup()
    ref.run(1, 50)
    assert np.array_equal(body[2:2 + 38 * s1], ref.download("pdf").ravel())
    ref.close()
    dflt = reference_run(ctrl, geo, 50)
    assert not np.array_equal(body[2:2 + 38 * s1], dflt.download("pdf").ravel())
    dflt.close()
