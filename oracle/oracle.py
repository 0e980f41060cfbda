"""ctypes binding of the CPU FP64 oracle (liboracle.so).

Test infrastructure: imported only by tests/, the smoke entry point and the CPU-baseline
leg of bench.py, always as the checker.  See mflbm_oracle.cpp for what each entry point
restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class OracleParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("iper", C.c_int32), ("jper", C.c_int32), ("kper", C.c_int32),
        ("wall_x_min", C.c_int32), ("wall_x_max", C.c_int32), ("wall_y_min", C.c_int32),
        ("wall_y_max", C.c_int32), ("wall_z_min", C.c_int32), ("wall_z_max", C.c_int32),
        ("inlet_bc", C.c_int32), ("outlet_bc", C.c_int32),
        ("lbm_gamma", C.c_double), ("force_z", C.c_double), ("la_nui1", C.c_double), ("la_nui2", C.c_double),
        ("lbm_beta", C.c_double), ("phi_inlet", C.c_double), ("sa_inject", C.c_double), ("cos_theta", C.c_double),
        ("rho_in", C.c_double), ("rho_out", C.c_double), ("relaxation", C.c_double),
        ("porous_plate", C.c_int32), ("z_porous_plate", C.c_int32), ("uin_avg", C.c_double),
        ("mrt", C.c_int32),
    ]


class OracleMonitor(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("fl1", "fl2", "vol1", "vol2", "mass1", "mass2", "pre")] + [
        ("umax2", C.c_double), ("usq1", C.c_double), ("usq2", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        vp, dp, i64p = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int64)
        L.oracle_create.restype = vp
        L.oracle_create.argtypes = [C.POINTER(OracleParams), C.POINTER(C.c_int8)]
        L.oracle_destroy.argtypes = [vp]
        L.oracle_init_phi.argtypes = [vp, C.c_int, C.c_double]
        L.oracle_init_phi.restype = C.c_int
        L.oracle_init_phi_random.argtypes = [vp, C.c_double, C.c_uint64]
        L.oracle_init_phi_random.restype = C.c_int
        L.oracle_node_uniform.argtypes = [C.c_uint64, C.c_uint64]
        L.oracle_node_uniform.restype = C.c_double
        L.oracle_init_pdf.argtypes = [vp, dp, dp, dp, dp]
        L.oracle_init_pdf_relaxed.argtypes = [vp, dp, dp, dp, dp]
        L.oracle_color_gradient.argtypes = [vp]
        L.oracle_step.argtypes = [vp, C.c_int]
        L.oracle_run.argtypes = [vp, C.c_int, C.c_int]
        L.oracle_collide.argtypes = [vp, C.c_int]
        L.oracle_boundaries.argtypes = [vp, C.c_int]
        L.oracle_collide_node.restype = C.c_double
        L.oracle_collide_node.argtypes = [C.POINTER(OracleParams), dp, dp, C.c_double, C.c_double, C.c_double,
                                          C.c_double, C.c_double, dp, dp]
        L.oracle_array.restype = vp
        L.oracle_array.argtypes = [vp, C.c_int, i64p]
        L.oracle_int_array.restype = C.POINTER(C.c_int32)
        L.oracle_int_array.argtypes = [vp, C.c_int, i64p]
        L.oracle_macro.argtypes = [vp, dp, dp, dp, dp, C.c_int]
        L.oracle_monitor_sums.argtypes = [vp, dp, dp, dp, dp, C.POINTER(OracleMonitor)]
        L.oracle_count.restype = C.c_int64
        L.oracle_count.argtypes = [vp, C.c_int]
        L.oracle_threads.restype = C.c_int
        _lib = L
    return _lib


def make_params(p: dict) -> OracleParams:
    op = OracleParams()
    for name, _ in OracleParams._fields_:
        setattr(op, name, p.get(name, 2) if name == "mrt" else p[name])
    return op


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


ARRAYS = {"pdf": 0, "phi": 1, "cn_x": 2, "cn_y": 3, "cn_z": 4, "c_norm": 5, "curv": 6, "s_nx": 7, "s_ny": 8, "s_nz": 9,
          "w_in": 10, "convec": 11, "smooth": 12}


class Oracle:
    """One oracle simulation; arrays are exposed as numpy views in the reference layout."""

    def __init__(self, params: dict, geometry: np.ndarray):
        self.p = dict(params)
        self.nx, self.ny, self.nz = int(params["nx"]), int(params["ny"]), int(params["nz"])
        geo = np.ascontiguousarray(geometry, dtype=np.int8)
        assert geo.shape == (self.nz, self.ny, self.nx)
        self._op = make_params(self.p)
        self.h = lib().oracle_create(C.byref(self._op), geo.ctypes.data_as(C.POINTER(C.c_int8)))
        if not self.h:
            raise RuntimeError("oracle_create failed (unsupported boundary condition?)")

    def close(self):
        if self.h:
            lib().oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- shapes -------------------------------------------------------------------
    def shape(self, halo):
        return (self.nz + 2 * halo, self.ny + 2 * halo, self.nx + 2 * halo)

    def array(self, name):
        n = C.c_int64()
        ptr = lib().oracle_array(self.h, ARRAYS[name], C.byref(n))
        if n.value == 0:
            return np.zeros(0)
        a = np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(n.value,))
        if name == "w_in":
            return a.reshape(self.ny + 2, self.nx + 2)
        if name == "convec":
            return a.reshape(39, self.ny + 2, self.nx + 2)
        halo = {"pdf": 1, "phi": 4, "curv": 1, "s_nx": 4, "s_ny": 4, "s_nz": 4, "smooth": 4}.get(name, 2)
        if name == "pdf":
            return a.reshape((2, 19) + self.shape(1))
        return a.reshape(self.shape(halo))

    def int_array(self, name):
        which = {"walls": 0, "walls_type": 1, "pore_profile_z": 2}[name]
        n = C.c_int64()
        ptr = lib().oracle_int_array(self.h, which, C.byref(n))
        a = np.ctypeslib.as_array(ptr, shape=(n.value,))
        if name == "walls":
            return a.reshape(self.shape(2))
        if name == "walls_type":
            return a.reshape(self.shape(4))
        return a

    # ---- operations -----------------------------------------------------------------
    def init_phi(self, option, interface_z0):
        if lib().oracle_init_phi(self.h, int(option), float(interface_z0)) != 0:
            raise ValueError(f"unsupported initial_fluid_distribution_option {option}")

    def init_phi_random(self, sa_target, seed):
        """option 6: phi = -1 where the node's seeded deviate exceeds sa_target"""
        lib().oracle_init_phi_random(self.h, float(sa_target), int(seed))

    def init_pdf(self, rho=None, u=None, v=None, w=None, relaxed=False):
        """relaxed: rest weights 4/21, 13/126, 1/63 (the collision's own rest state)"""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w)]
        fn = lib().oracle_init_pdf_relaxed if relaxed else lib().oracle_init_pdf
        fn(self.h, *[_dp(a) for a in arrs])

    def color_gradient(self):
        lib().oracle_color_gradient(self.h)

    def step(self, ntime):
        lib().oracle_step(self.h, int(ntime))

    def run(self, ntime0, nsteps):
        lib().oracle_run(self.h, int(ntime0), int(nsteps))

    def collide(self, odd):
        lib().oracle_collide(self.h, int(bool(odd)))

    def boundaries(self, odd):
        lib().oracle_boundaries(self.h, int(bool(odd)))

    def macro(self, zero_solid_phi=True):
        shp = self.shape(1)
        rho, u, v, w = (np.zeros(shp) for _ in range(4))
        lib().oracle_macro(self.h, _dp(rho), _dp(u), _dp(v), _dp(w), int(zero_solid_phi))
        return rho, u, v, w

    def monitor_sums(self, rho, u, v, w):
        nz = self.nz
        bufs = {k: np.zeros(nz) for k in ("fl1", "fl2", "vol1", "vol2", "mass1", "mass2", "pre")}
        m = OracleMonitor()
        for k, b in bufs.items():
            setattr(m, k, _dp(b))
        lib().oracle_monitor_sums(self.h, _dp(rho), _dp(u), _dp(v), _dp(w), C.byref(m))
        bufs.update(umax2=m.umax2, usq1=m.usq1, usq2=m.usq2)
        return bufs

    def counts(self):
        return lib().oracle_count(self.h, 0), lib().oracle_count(self.h, 1)

    def setup(self, option=None, interface_z0=None):
        """New-simulation initialisation: phi, equilibrium populations, initial gradient."""
        option = self.p["initial_fluid_distribution_option"] if option is None else option
        if option == 6:
            self.init_phi_random(self.p["sa_target"], self.p["random_seed"])
        else:
            self.init_phi(option, self.p["interface_z0"] if interface_z0 is None else interface_z0)
        self.init_pdf()
        self.color_gradient()


def collide_node(params: dict, g1, g2, cn, curv, cnorm):
    op = make_params(params)
    g1 = np.ascontiguousarray(g1, dtype=np.float64)
    g2 = np.ascontiguousarray(g2, dtype=np.float64)
    o1, o2 = np.zeros(19), np.zeros(19)
    phi = lib().oracle_collide_node(C.byref(op), _dp(g1), _dp(g2), float(cn[0]), float(cn[1]), float(cn[2]),
                                    float(curv), float(cnorm), _dp(o1), _dp(o2))
    return phi, o1, o2


def threads():
    return lib().oracle_threads()
