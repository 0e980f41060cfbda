"""ctypes binding of libmflbm.so, the HIP implementation of the colour-gradient step.

There is no CPU fallback: if the library or a GPU is missing, constructing a ``Solver``
raises.  Arrays returned by the download helpers use the reference layout with numpy
shape (z, y, x) (pdf: (2, 19, z, y, x)).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libmflbm.so")
REPO = os.path.dirname(PKG_DIR)
HEADER = os.path.join(REPO, "include", "mflbm.h")

FP64, FP32 = 0, 1
FIELDS = {"pdf": 0, "phi": 1, "cn_x": 2, "cn_y": 3, "cn_z": 4, "c_norm": 5, "curv": 6,
          "s_nx": 7, "s_ny": 8, "s_nz": 9, "walls": 10, "walls_type": 11, "w_in": 12, "convec": 13}
HALO = {"pdf": 1, "phi": 4, "cn_x": 2, "cn_y": 2, "cn_z": 2, "c_norm": 2, "curv": 1,
        "s_nx": 4, "s_ny": 4, "s_nz": 4, "walls": 2, "walls_type": 4}

# Every symbol declared in include/mflbm.h.
EXPORTS = ["mflbm_create", "mflbm_destroy", "mflbm_last_error", "mflbm_init_phi", "mflbm_init_phi_random",
           "mflbm_init_equilibrium",
           "mflbm_init_relaxed",
           "mflbm_color_gradient", "mflbm_step", "mflbm_run", "mflbm_profile", "mflbm_download", "mflbm_upload",
           "mflbm_macro", "mflbm_monitor", "mflbm_velocity_change", "mflbm_sync", "mflbm_count", "mflbm_field_size", "mflbm_bandwidth_probe",
           "mflbm_ensemble_create", "mflbm_ensemble_destroy", "mflbm_ensemble_last_error",
           "mflbm_ensemble_fluid_nodes", "mflbm_ensemble_solve", "mflbm_ensemble_fields",
           "mflbm_ensemble_init_fields"]


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int64), ("ny", C.c_int64), ("nz", C.c_int64),
        ("iper", C.c_int32), ("jper", C.c_int32), ("kper", C.c_int32),
        ("wall_x_min", C.c_int32), ("wall_x_max", C.c_int32), ("wall_y_min", C.c_int32),
        ("wall_y_max", C.c_int32), ("wall_z_min", C.c_int32), ("wall_z_max", C.c_int32),
        ("inlet_bc", C.c_int32), ("outlet_bc", C.c_int32),
        ("lbm_gamma", C.c_double), ("force_z", C.c_double), ("la_nui1", C.c_double), ("la_nui2", C.c_double),
        ("lbm_beta", C.c_double), ("phi_inlet", C.c_double), ("sa_inject", C.c_double), ("cos_theta", C.c_double),
        ("rho_in", C.c_double), ("rho_out", C.c_double), ("relaxation", C.c_double),
        ("precision", C.c_int32), ("flags", C.c_int32),
        ("porous_plate", C.c_int32), ("z_porous_plate", C.c_int32), ("uin_avg", C.c_double),
        ("mrt", C.c_int32),
    ]


class MonitorOut(C.Structure):
    _fields_ = [(n, C.POINTER(C.c_double)) for n in ("fl1", "fl2", "vol1", "vol2", "mass1", "mass2", "pre")] + [
        ("umax2", C.c_double), ("usq1", C.c_double), ("usq2", C.c_double), ("n_fluid", C.c_int64)]


class KernelTimes(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("collide_odd_ms", "collide_even_ms", "boundary_ms", "gradient_ms")] + [
        ("n_odd", C.c_int32), ("n_even", C.c_int32)]


class MflbmError(RuntimeError):
    pass


_lib = None


def load():
    """Load libmflbm.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MflbmError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                         "(make -C mf-lbm-cuda_amd/csrc)")
    L = C.CDLL(LIB_PATH)
    vp, dp = C.c_void_p, C.POINTER(C.c_double)
    L.mflbm_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_int8), C.POINTER(vp)]
    L.mflbm_create.restype = C.c_int
    L.mflbm_destroy.argtypes = [vp]
    L.mflbm_last_error.argtypes = [vp]
    L.mflbm_last_error.restype = C.c_char_p
    L.mflbm_init_phi.argtypes = [vp, C.c_int32, C.c_double]
    L.mflbm_init_phi_random.argtypes = [vp, C.c_double, C.c_uint64]
    L.mflbm_init_equilibrium.argtypes = [vp, dp, dp, dp, dp]
    L.mflbm_init_relaxed.argtypes = [vp, dp, dp, dp, dp]
    L.mflbm_color_gradient.argtypes = [vp]
    L.mflbm_step.argtypes = [vp, C.c_int32]
    L.mflbm_run.argtypes = [vp, C.c_int32, C.c_int32]
    L.mflbm_profile.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(KernelTimes)]
    L.mflbm_download.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.mflbm_upload.argtypes = [vp, C.c_int32, vp, C.c_size_t]
    L.mflbm_macro.argtypes = [vp, dp, dp, dp, dp]
    L.mflbm_monitor.argtypes = [vp, C.POINTER(MonitorOut)]
    L.mflbm_velocity_change.argtypes = [vp, C.POINTER(C.c_double)]
    L.mflbm_sync.argtypes = [vp]
    L.mflbm_count.argtypes = [vp, C.c_int32]
    L.mflbm_count.restype = C.c_int64
    L.mflbm_field_size.argtypes = [vp, C.c_int32]
    L.mflbm_field_size.restype = C.c_int64
    L.mflbm_bandwidth_probe.argtypes = [vp, C.c_int64, C.c_int32, C.POINTER(C.c_double)]
    L.mflbm_bandwidth_probe.restype = C.c_int
    i32p = C.POINTER(C.c_int32)
    L.mflbm_ensemble_create.argtypes = [C.POINTER(Params), C.POINTER(C.c_int8), C.c_int32, dp, C.POINTER(vp)]
    L.mflbm_ensemble_create.restype = C.c_int
    L.mflbm_ensemble_destroy.argtypes = [vp]
    L.mflbm_ensemble_last_error.argtypes = [vp]
    L.mflbm_ensemble_last_error.restype = C.c_char_p
    L.mflbm_ensemble_fluid_nodes.argtypes = [vp]
    L.mflbm_ensemble_fluid_nodes.restype = C.c_int64
    L.mflbm_ensemble_solve.argtypes = [vp, C.c_int32, C.c_double, C.c_int32, i32p, dp, dp, dp]
    L.mflbm_ensemble_solve.restype = C.c_int
    L.mflbm_ensemble_fields.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_float)]
    L.mflbm_ensemble_fields.restype = C.c_int
    L.mflbm_ensemble_init_fields.argtypes = [vp, C.c_int32, C.c_int32, C.POINTER(C.c_float)]
    L.mflbm_ensemble_init_fields.restype = C.c_int
    for name in ("mflbm_init_phi", "mflbm_init_phi_random", "mflbm_init_equilibrium", "mflbm_init_relaxed", "mflbm_color_gradient",
                 "mflbm_step", "mflbm_run", "mflbm_profile", "mflbm_download", "mflbm_upload", "mflbm_macro",
                 "mflbm_monitor", "mflbm_sync"):
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


FLAG_SPLIT_GRADIENT = 1  # MFLBM_FLAG_SPLIT_GRADIENT


def make_params(p: dict, precision=FP64, flags=0) -> Params:
    cp = Params()
    for name, _ in Params._fields_:
        if name == "precision":
            cp.precision = precision
        elif name == "flags":
            cp.flags = flags
        elif name == "mrt":
            cp.mrt = int(p.get("mrt", 2))
        else:
            setattr(cp, name, p[name])
    return cp


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


class Solver:
    """One device-resident simulation (handle of the C ABI)."""

    def __init__(self, params: dict, geometry: np.ndarray, precision=FP64, split_gradient=False):
        L = load()
        self.p = dict(params)
        self.precision = precision
        self.dtype = np.float64 if precision == FP64 else np.float32
        self.nx, self.ny, self.nz = int(params["nx"]), int(params["ny"]), int(params["nz"])
        geo = np.ascontiguousarray(geometry, dtype=np.int8)
        if geo.shape != (self.nz, self.ny, self.nx):
            raise ValueError(f"geometry shape {geo.shape} != {(self.nz, self.ny, self.nx)}")
        self._cp = make_params(self.p, precision, FLAG_SPLIT_GRADIENT if split_gradient else 0)
        h = C.c_void_p()
        rc = L.mflbm_create(C.byref(self._cp), geo.ctypes.data_as(C.POINTER(C.c_int8)), C.byref(h))
        self.h = h
        if rc != 0:
            msg = L.mflbm_last_error(h).decode() if h else "mflbm_create failed"
            self.close()
            raise MflbmError(msg)

    def _check(self, rc):
        if rc != 0:
            raise MflbmError(load().mflbm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            load().mflbm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, halo):
        return (self.nz + 2 * halo, self.ny + 2 * halo, self.nx + 2 * halo)

    # ---- set-up ------------------------------------------------------------------------
    def init_phi(self, option, interface_z0):
        self._check(load().mflbm_init_phi(self.h, int(option), float(interface_z0)))

    def init_phi_random(self, sa_target, seed):
        """option 6: phi = -1 where the node's seeded deviate exceeds sa_target"""
        self._check(load().mflbm_init_phi_random(self.h, float(sa_target), int(seed)))

    def init_equilibrium(self, rho=None, u=None, v=None, w=None, relaxed=False):
        """relaxed: rest weights of the collision (mflbm_init_relaxed) instead of the D3Q19 ones"""
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (rho, u, v, w)]
        fn = load().mflbm_init_relaxed if relaxed else load().mflbm_init_equilibrium
        self._check(fn(self.h, *[_dp(a) for a in arrs]))

    def color_gradient(self):
        self._check(load().mflbm_color_gradient(self.h))

    def setup(self, option=None, interface_z0=None):
        option = self.p["initial_fluid_distribution_option"] if option is None else option
        if option == 6:
            self.init_phi_random(self.p["sa_target"], self.p["random_seed"])
        else:
            self.init_phi(option, self.p["interface_z0"] if interface_z0 is None else interface_z0)
        self.init_equilibrium()
        self.color_gradient()

    # ---- stepping ----------------------------------------------------------------------
    def step(self, ntime):
        self._check(load().mflbm_step(self.h, int(ntime)))

    def run(self, ntime0, nsteps):
        self._check(load().mflbm_run(self.h, int(ntime0), int(nsteps)))

    def profile(self, ntime0, nsteps) -> dict:
        kt = KernelTimes()
        self._check(load().mflbm_profile(self.h, int(ntime0), int(nsteps), C.byref(kt)))
        return {n: getattr(kt, n) for n, _ in KernelTimes._fields_}

    def sync(self):
        self._check(load().mflbm_sync(self.h))

    # ---- data --------------------------------------------------------------------------
    def download(self, name):
        f = FIELDS[name]
        n = load().mflbm_field_size(self.h, f)
        dt = np.int8 if name == "walls_type" else (np.uint8 if name == "walls" else self.dtype)
        out = np.empty(n, dtype=dt)
        self._check(load().mflbm_download(self.h, f, out.ctypes.data_as(C.c_void_p), out.nbytes))
        if name == "pdf":
            return out.reshape((2, 19) + self.shape(1))
        if name == "w_in":
            return out.reshape(self.shape(1)[1:])
        if name == "convec":
            return out.reshape((39,) + self.shape(1)[1:])
        return out.reshape(self.shape(HALO[name]))

    def upload(self, name, arr):
        f = FIELDS[name]
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        self._check(load().mflbm_upload(self.h, f, a.ctypes.data_as(C.c_void_p), a.nbytes))

    def macro(self):
        shp = self.shape(1)
        rho, u, v, w = (np.zeros(shp) for _ in range(4))
        self._check(load().mflbm_macro(self.h, _dp(rho), _dp(u), _dp(v), _dp(w)))
        return rho, u, v, w

    def velocity_change(self) -> float:
        """|u - u_prev| / |u| since the previous call (-1 after an initialisation)"""
        c = C.c_double()
        self._check(load().mflbm_velocity_change(self.h, C.byref(c)))
        return c.value

    def monitor(self) -> dict:
        nz = self.nz
        bufs = {k: np.zeros(nz) for k in ("fl1", "fl2", "vol1", "vol2", "mass1", "mass2", "pre")}
        m = MonitorOut()
        for k, b in bufs.items():
            setattr(m, k, _dp(b))
        self._check(load().mflbm_monitor(self.h, C.byref(m)))
        bufs.update(umax2=m.umax2, usq1=m.usq1, usq2=m.usq2, n_fluid=m.n_fluid)
        return bufs

    def bandwidth_probe(self, nbytes, repeats=10):
        """Mean device time (ms) of one plain copy of nbytes (read nbytes + write nbytes)."""
        ms = C.c_double()
        self._check(load().mflbm_bandwidth_probe(self.h, int(nbytes), int(repeats), C.byref(ms)))
        return ms.value

    def counts(self):
        L = load()
        return {"fluid": L.mflbm_count(self.h, 0), "fluid_ext": L.mflbm_count(self.h, 1),
                "solid_boundary": L.mflbm_count(self.h, 2), "long_link_waves": L.mflbm_count(self.h, 4)}

    def gradient_path(self):
        """'march' (one-pass z-march), 'tiles' (two-pass 3-D tiles) or 'split' (four kernels)"""
        k = load().mflbm_count(self.h, 3)
        return "split" if k == 6 else "tiles" if k == 5 else "march"
