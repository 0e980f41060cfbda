"""Control-file reader and derived step parameters.

Mirrors the reference's parameter handling:

* ``read_control`` follows ``read_int``/``read_double``/``read_string`` in
  src/utils.cpp:17-154: one ``key value`` pair per line, lines starting with ``#`` are
  comments, the first matching key wins, a missing key is an error.
* ``derive`` follows src/IO_multiphase.cpp:196-233 (contact-angle conversion and the
  consistency checks) and src/Init_multiphase.cpp:173-214 (relaxation parameters,
  inlet order parameter, pressure-boundary densities).

The result is a plain dict of the constants the time step needs; the C ABI
(``include/mflbm.h``) and the oracle take the same values.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field

# Keys read by the reference (src/IO_multiphase.cpp:47-187) with their types.
INT_KEYS = [
    "initial_fluid_distribution_option", "breakthrough_check", "steady_state_option",
    "benchmark_cmd", "output_fieldData_precision_cmd", "extreme_large_sim_cmd",
    "modify_geometry_cmd", "external_geometry_read_cmd", "geometry_dims_type_size",
    "geometry_preprocess_cmd", "porous_plate_cmd", "change_inlet_fluid_phase_cmd",
    "n_exclude_inlet", "n_exclude_outlet",
    "domain_wall_status_x_min", "domain_wall_status_x_max",
    "domain_wall_status_y_min", "domain_wall_status_y_max",
    "domain_wall_status_z_min", "domain_wall_status_z_max",
    "iper", "jper", "kper", "inlet_BC", "outlet_BC", "Z_porous_plate",
    "max_time_step", "max_time_step_benchmark", "ntime_visual", "ntime_animation",
    "monitor_timer", "monitor_profile_timer_ratio", "computation_time_timer",
    "display_steps_timer", "rho_in_new", "rho_out_BC",
    "block_Threads_X", "block_Threads_Y", "block_Threads_Z",
]
LONG_KEYS = ["nxGlobal", "nyGlobal", "nzGlobal"]
REAL_KEYS = [
    "convergence_criteria", "fluid1_viscosity", "fluid2_viscosity", "surface_tension",
    "theta", "RK_beta", "target_inject_pore_volume", "capillary_number",
    "saturation_injection", "initial_interface_position", "body_force_0",
    "target_fluid1_saturation", "checkpoint_save_timer", "checkpoint_2rd_save_timer",
    "simulation_duration_timer", "d_vol_animation", "d_vol_detail", "d_vol_monitor",
]

PI = 3.14159265358979323846  # includes/Module.h:9


class ControlError(ValueError):
    """Raised where the reference calls ERROR() and exits."""


def _lookup(lines, key):
    for line in lines:
        if not line or line[0] == "#":
            continue
        toks = line.split()
        if toks and toks[0] == key:
            if len(toks) < 2:
                raise ControlError(f"no value for ({key})")
            return toks[1]
    raise ControlError(f"ERROR! Couldn't find the variable named: ({key}) in the file")


def read_control(path: str) -> dict:
    """Parse a simulation_control.txt into a dict of raw control values."""
    with open(path, "r") as fh:
        lines = [ln.rstrip("\n").rstrip("\r") for ln in fh]
    out = {}
    for k in INT_KEYS:
        out[k] = int(_lookup(lines, k))
    for k in LONG_KEYS:
        out[k] = int(_lookup(lines, k))
    for k in REAL_KEYS:
        out[k] = float(_lookup(lines, k))
    return out


def read_geometry_path(path: str) -> tuple[str, str]:
    """Geometry_File_Path.txt rule (src/Misc.cpp:24-43): strip trailing non-digits, add .dat."""
    with open(path, "r") as fh:
        lines = [ln.rstrip("\n").rstrip("\r") for ln in fh]

    def fix(v):
        while v and not v[-1].isdigit():
            v = v[:-1]
        return v + ".dat"

    return fix(_lookup(lines, "geo_file_path_")), fix(_lookup(lines, "geo_boundary_file_path_"))


def read_job_status(path: str) -> str:
    """The whole file is the status; it is compared exactly (src/IO_multiphase.cpp:18-36)."""
    with open(path, "r") as fh:
        status = fh.read()
    if status not in ("new_simulation", "continue_simulation"):
        raise ControlError("Wrong simlation status! Exiting program!")
    return status


def derive(ctrl: dict) -> dict:
    """Constants of the step, derived as the reference derives them."""
    errors = []
    theta = float(ctrl["theta"])
    if theta > 90.0:
        errors.append("contact angle is larger than 90 degrees")
    theta = 180.0 - theta
    theta = theta * PI / 180.0
    cos_theta = math.cos(theta)
    if ctrl["iper"] == 1 or ctrl["domain_wall_status_x_max"] == 0 or ctrl["domain_wall_status_x_min"] == 0:
        errors.append("X direction periodic BC enabled or non-slip BC not applied at x = xmin or x = xmax")
    if ctrl["jper"] == 0 and (ctrl["domain_wall_status_y_max"] == 0 or ctrl["domain_wall_status_y_min"] == 0):
        errors.append("non-slip BC not applied at y walls while y is not periodic")
    if ctrl["jper"] == 1 and (ctrl["domain_wall_status_y_max"] == 1 or ctrl["domain_wall_status_y_min"] == 1):
        errors.append("non-slip BC applied at y walls while y is periodic")
    if ctrl["kper"] == 1 and (ctrl["domain_wall_status_z_max"] == 1 or ctrl["domain_wall_status_z_min"] == 1):
        errors.append("non-slip BC applied at z walls while z is periodic")
    if ctrl["outlet_BC"] == 1 and ctrl["inlet_BC"] == 2:
        errors.append("Inlet pressure + outlet convective BC is not supported")
    if errors:
        raise ControlError("; ".join(errors))

    nz = int(ctrl["nzGlobal"])
    la_nu1 = float(ctrl["fluid1_viscosity"])
    la_nu2 = float(ctrl["fluid2_viscosity"])
    sa_inject = float(ctrl["saturation_injection"])
    force_z0 = float(ctrl["body_force_0"])
    force_z = force_z0
    rho_out = 1.0
    rho_in = 0.0
    open_z = ctrl["kper"] == 0 and ctrl["domain_wall_status_z_min"] == 0 and ctrl["domain_wall_status_z_max"] == 0
    if open_z:
        if ctrl["inlet_BC"] == 1:
            force_z = 0.0
        elif ctrl["inlet_BC"] == 2:
            force_z = 0.0
            p_gradient = -force_z0 / 3.0
            if ctrl["rho_out_BC"]:
                rho_out = 1.0 - p_gradient * float(nz)
            else:
                rho_in = rho_out - p_gradient * float(nz)
    return {
        "nx": int(ctrl["nxGlobal"]), "ny": int(ctrl["nyGlobal"]), "nz": nz,
        "iper": int(ctrl["iper"]), "jper": int(ctrl["jper"]), "kper": int(ctrl["kper"]),
        "wall_x_min": int(ctrl["domain_wall_status_x_min"]), "wall_x_max": int(ctrl["domain_wall_status_x_max"]),
        "wall_y_min": int(ctrl["domain_wall_status_y_min"]), "wall_y_max": int(ctrl["domain_wall_status_y_max"]),
        "wall_z_min": int(ctrl["domain_wall_status_z_min"]), "wall_z_max": int(ctrl["domain_wall_status_z_max"]),
        "inlet_bc": int(ctrl["inlet_BC"]), "outlet_bc": int(ctrl["outlet_BC"]),
        "lbm_gamma": float(ctrl["surface_tension"]),
        "force_z": force_z,
        "la_nui1": 1.0 / la_nu1,
        "la_nui2": 1.0 / la_nu2,
        "lbm_beta": float(ctrl["RK_beta"]),
        "phi_inlet": 2.0 * sa_inject - 1.0,
        "sa_inject": sa_inject,
        "cos_theta": cos_theta,
        "rho_in": rho_in,
        "rho_out": rho_out,
        "relaxation": 1.0,
        "porous_plate": int(ctrl.get("porous_plate_cmd", 0)),
        "z_porous_plate": int(ctrl.get("Z_porous_plate", 0)),
        # mean inlet velocity, set only for a velocity inlet on an open z axis (src/Init_multiphase.cpp:195-199,
        # 262-263); the convective outlet advects with it, so with another inlet it advects with 0
        "uin_avg": (float(ctrl.get("capillary_number", 0.0)) * float(ctrl["surface_tension"]) / la_nu1
                    if open_z and int(ctrl["inlet_BC"]) == 1 else 0.0),
        # not step constants, kept for initialisation / monitors
        "la_nu1": la_nu1, "la_nu2": la_nu2,
        "initial_fluid_distribution_option": int(ctrl["initial_fluid_distribution_option"]),
        "interface_z0": float(ctrl["initial_interface_position"]),
        "sa_target": float(ctrl["target_fluid1_saturation"]),
        "random_seed": int(ctrl.get("random_seed", 1)),
        "n_exclude_inlet": int(ctrl["n_exclude_inlet"]),
        "n_exclude_outlet": int(ctrl["n_exclude_outlet"]),
    }


def default_control() -> dict:
    """A complete control dict with the reference example's values (input/simulation_control.txt)."""
    c = {k: 0 for k in INT_KEYS}
    c.update({k: 0 for k in LONG_KEYS})
    c.update({k: 0.0 for k in REAL_KEYS})
    c.update(dict(
        initial_fluid_distribution_option=1, steady_state_option=3, convergence_criteria=1e-4,
        external_geometry_read_cmd=1, geometry_dims_type_size=4, n_exclude_inlet=10, n_exclude_outlet=10,
        domain_wall_status_x_min=1, domain_wall_status_x_max=1, domain_wall_status_y_min=1,
        domain_wall_status_y_max=1, inlet_BC=2, outlet_BC=2, fluid1_viscosity=0.04, fluid2_viscosity=0.4,
        surface_tension=0.03, theta=65.0, RK_beta=0.95, target_inject_pore_volume=-1.0,
        capillary_number=100e-6, saturation_injection=1.0, initial_interface_position=8.0,
        body_force_0=1e-4, target_fluid1_saturation=0.4, max_time_step=500000, max_time_step_benchmark=100,
        ntime_visual=5000000, ntime_animation=10000, monitor_timer=10000, monitor_profile_timer_ratio=5,
        computation_time_timer=10000, display_steps_timer=10000, checkpoint_save_timer=1.0,
        checkpoint_2rd_save_timer=5.5, simulation_duration_timer=168.0, d_vol_animation=-0.05,
        d_vol_detail=-1.0, d_vol_monitor=-0.01, rho_in_new=1, rho_out_BC=0,
        block_Threads_X=128, block_Threads_Y=1, block_Threads_Z=1,
        nxGlobal=64, nyGlobal=64, nzGlobal=64,
    ))
    return c


def write_control(path: str, ctrl: dict) -> None:
    """Write a control dict in the reference's ``key value`` format."""
    with open(path, "w") as fh:
        fh.write("# simulation control (MF-LBM key/value format)\n")
        for k in INT_KEYS + LONG_KEYS:
            fh.write(f"{k} {int(ctrl[k])}\n")
        for k in REAL_KEYS:
            fh.write(f"{k} {float(ctrl[k])!r}\n")
        if "random_seed" in ctrl:  # option 6's seed (a key of this build only)
            fh.write(f"random_seed {int(ctrl['random_seed'])}\n")
