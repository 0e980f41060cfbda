"""The warm-start network on the device: a four-scale 3-D convolutional U-Net trained and
evaluated by the project's own HIP kernels (csrc/nn.hip, C ABI in include/mfnn.h).

Device memory comes from libmflbm as well (mfnn_alloc), so the kernels, the copies and the
synchronisation all go through one HIP runtime; every operation of the forward and backward
passes and of the optimiser is a kernel of libmflbm, launched on the default stream in
order.  Tensors are channels-last (D, H, W, C) with C a multiple of 8.

Architecture (widths c0..c3, e.g. 16, 32, 64, 128):
    e1, e2   conv3 + ReLU at full resolution        (cin -> c0 -> c0)
    e3, e4   2x2x2 average pool, conv3 + ReLU x 2   (c0 -> c1 -> c1)
    e5, e6   pool, conv3 + ReLU x 2                 (c1 -> c2 -> c2)
    e7, e8   pool, conv3 + ReLU x 2                 (c2 -> c3 -> c3)
    d9       nearest upsample of e8 ++ e6, conv3 + ReLU   (c3 + c2 -> c2)
    d10      upsample of d9 ++ e4, conv3 + ReLU            (c2 + c1 -> c1)
    d11      upsample of d10 ++ e2, conv3 + ReLU           (c1 + c0 -> c0)
    head     conv3, linear, fp32 output                    (c0 -> cout)
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import gpu

LAYERS = ("e1", "e2", "e3", "e4", "e5", "e6", "e7", "e8", "d9", "d10", "d11", "head")
_bound = False

_vp, _i32, _i64, _fp = C.c_void_p, C.c_int32, C.c_int64, C.c_float
# argument types of every entry point of include/mfnn.h
SIGNATURES = {
    "mfnn_alloc": [_i64],
    "mfnn_free": [_vp],
    "mfnn_copy": [_vp, _vp, _i64, _i32],
    "mfnn_zero": [_vp, _i64],
    "mfnn_sync": [],
    "mfnn_conv3": [_vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp],
    "mfnn_wgrad_workspace": [_i32, _i32, _i32, _i32, _i32, _i32],
    "mfnn_conv3_wgrad": [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _vp, _i32, _vp],
    "mfnn_conv3_wgrad_static": [_vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _vp, _i32, _vp],
    "mfnn_flip_weights": [_vp, _i32, _i32, _i32, _i32, _i32, _vp, _vp],
    "mfnn_to_bf16": [_vp, _i64, _vp, _vp],
    "mfnn_pool2": [_vp, _i32, _i32, _i32, _i32, _vp, _i32, _vp],
    "mfnn_up2": [_vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp, _i32, _vp],
    "mfnn_channels": [_vp, _i64, _i32, _i32, _i32, _i32, _vp, _vp],
    "mfnn_add_gate": [_vp, _vp, _vp, _i64, _vp, _vp],
    "mfnn_mse_partials": [_i64, _i32],
    "mfnn_mse": [_vp, _vp, _vp, _i64, _i32, _fp, _vp, _vp, _vp],
    "mfnn_adam": [_vp, _vp, _vp, _vp, _i64, _fp, _fp, _fp, _fp, _i32, _vp, _vp],
}


# voxels per weight-gradient chunk: each chunk's fp32 partials are summed in chunk order.
# Grids below 65536 voxels use WGRAD_CHUNK_SMALL, so that they still spread over many waves.
WGRAD_CHUNK = 2048
WGRAD_CHUNK_SMALL = 512


def wgrad_chunk(D, H, W):
    return WGRAD_CHUNK if D * H * W >= 65536 else WGRAD_CHUNK_SMALL


def lib():
    global _bound
    L = gpu.load()
    if not _bound:
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = (C.c_int64 if name in ("mfnn_wgrad_workspace", "mfnn_mse_partials")
                         else (C.c_void_p if name == "mfnn_alloc" else C.c_int))
        _bound = True
    return L


def _chk(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def rup(x, m):
    return (x + m - 1) // m * m


def wrows(c):
    """weight rows the conv kernel reads for c output channels"""
    return rup(c, 16) if c <= 128 else 192


def ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def bf16_bits(a):
    """float array -> uint16 bf16 bit patterns (round to nearest even)"""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def bf16_values(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


class DevArray:
    """A device buffer from libmflbm with a shape and an element type (bf16, f32, f64)."""

    SIZES = {"bf16": 2, "f32": 4, "f64": 8}

    def __init__(self, shape, dtype="f32", zero=False):
        self.shape = tuple(int(v) for v in shape)
        self.dtype = dtype
        self.nbytes = max(1, int(np.prod(self.shape)) * self.SIZES[dtype])
        p = lib().mfnn_alloc(self.nbytes)
        if not p:
            raise MemoryError(f"device allocation of {self.nbytes} bytes failed")
        self.ptr = p
        if zero:
            self.zero()

    @classmethod
    def from_numpy(cls, a, dtype="f32"):
        d = cls(np.shape(a), dtype)
        d.upload(a)
        return d

    def numel(self):
        return int(np.prod(self.shape))

    def data_ptr(self):
        return self.ptr

    def zero(self):
        _chk(lib().mfnn_zero(C.c_void_p(self.ptr), self.nbytes), "zero")

    def upload(self, a):
        if self.dtype == "bf16":
            h = bf16_bits(a)
        else:
            h = np.ascontiguousarray(a, dtype=np.float32 if self.dtype == "f32" else np.float64)
        if h.size != self.numel():
            raise ValueError(f"size {h.size} != {self.numel()}")
        _chk(lib().mfnn_copy(C.c_void_p(self.ptr), h.ctypes.data_as(C.c_void_p), h.nbytes, 0), "upload")

    def numpy(self):
        """host copy as float32 (bf16, f32) or float64 (f64); waits for queued kernels"""
        dt = {"bf16": np.uint16, "f32": np.float32, "f64": np.float64}[self.dtype]
        h = np.empty(self.shape, dt)
        _chk(lib().mfnn_copy(h.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), h.nbytes, 1), "download")
        return bf16_values(h) if self.dtype == "bf16" else h

    def copy_from(self, other):
        _chk(lib().mfnn_copy(C.c_void_p(self.ptr), C.c_void_p(other.ptr), min(self.nbytes, other.nbytes), 2), "copy")

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().mfnn_free(C.c_void_p(self.ptr))
                self.ptr = None
        except Exception:
            pass


def sync():
    _chk(lib().mfnn_sync(), "sync")


class Conv:
    """One 3x3x3 convolution: fp32 master weights W[n][k] (zero padded), bf16 copy, bias,
    gradients, Adam moments, and the flipped bf16 weights of its data gradient."""

    def __init__(self, cin, cout, rng, zero_rows_from=None):
        self.cin, self.cout = cin, cout
        self.kp = rup(27 * cin, 32)
        self.rows = wrows(cout)
        w = np.zeros((self.rows, self.kp), np.float32)
        n_init = cout if zero_rows_from is None else zero_rows_from
        w[:n_init, :27 * cin] = rng.normal(0.0, np.sqrt(2.0 / (27 * cin)), (n_init, 27 * cin))
        self.w = DevArray.from_numpy(w, "f32")
        self.b = DevArray((cout,), "f32", zero=True)
        self.wb = DevArray((self.rows, self.kp), "bf16")
        self.dw = DevArray((self.rows, self.kp), "f32", zero=True)
        self.db = DevArray((cout,), "f32", zero=True)
        self.m = [DevArray((self.rows, self.kp), "f32", zero=True), DevArray((self.rows, self.kp), "f32", zero=True),
                  DevArray((cout,), "f32", zero=True), DevArray((cout,), "f32", zero=True)]
        self.kp2 = rup(27 * cout, 32)
        self.wt = DevArray((wrows(cin), self.kp2), "bf16")
        self.refresh()

    def refresh(self, stream=None):
        _chk(lib().mfnn_to_bf16(ptr(self.w), self.w.numel(), ptr(self.wb), stream), "to_bf16")

    def set_numpy(self, w, b, stream=None):
        full = np.zeros((self.rows, self.kp), np.float32)
        full[:w.shape[0], :w.shape[1]] = w
        self.w.upload(full)
        self.b.upload(np.asarray(b, np.float32))
        self.refresh(stream)

    def get_numpy(self):
        return (self.w.numpy()[:self.cout, :27 * self.cin].astype(np.float64), self.b.numpy().astype(np.float64))

    def forward(self, x, shape, relu, yb=None, yf=None, stream=None):
        D, H, W = shape
        _chk(lib().mfnn_conv3(ptr(x), ptr(self.wb), ptr(self.b), None, D, H, W, self.cin, self.kp, self.cout,
                              int(relu), ptr(yb), ptr(yf), stream), "conv3")

    def backward(self, x, dz, shape, work, accumulate, dx=None, gate=None, stream=None):
        """weight/bias gradients from input x and output gradient dz; data gradient into dx
        (times [gate > 0] when gate is given)."""
        D, H, W = shape
        L = lib()
        _chk(L.mfnn_conv3_wgrad(ptr(x), ptr(dz), D, H, W, self.cin, self.kp, self.cout, wgrad_chunk(D, H, W),
                                ptr(work), ptr(self.dw), ptr(self.db), int(accumulate), stream), "wgrad")
        if dx is not None:
            _chk(L.mfnn_flip_weights(ptr(self.wb), self.cin, self.kp, self.cout, self.kp2, self.wt.shape[0],
                                     ptr(self.wt), stream), "flip")
            _chk(L.mfnn_conv3(ptr(dz), ptr(self.wt), None, ptr(gate), D, H, W, self.cout, self.kp2, self.cin, 0,
                              ptr(dx), None, stream), "dgrad")

    def adam(self, step, lr, stream=None, b1=0.9, b2=0.999, eps=1e-8):
        L = lib()
        _chk(L.mfnn_adam(ptr(self.w), ptr(self.dw), ptr(self.m[0]), ptr(self.m[1]), self.w.numel(), lr, b1, b2, eps,
                         step, ptr(self.wb), stream), "adam w")
        _chk(L.mfnn_adam(ptr(self.b), ptr(self.db), ptr(self.m[2]), ptr(self.m[3]), self.b.numel(), lr, b1, b2, eps,
                         step, None, stream), "adam b")


class UNet:
    """Device U-Net for grids of shape (D, H, W) divisible by 8."""

    def __init__(self, widths=(16, 32, 64, 128), cin=8, cout=8, n_out=None, seed=0):
        self.widths, self.cin, self.cout = tuple(widths), cin, cout
        c0, c1, c2, c3 = widths
        for c in (cin, cout, c0, c1, c2, c3):
            if c % 8:
                raise ValueError("channel counts must be multiples of 8")
        if c3 + c2 > 192:
            raise ValueError("the data gradient of the first decoder layer has c3 + c2 outputs, at most 192")
        rng = np.random.default_rng(seed)
        shapes = {"e1": (cin, c0), "e2": (c0, c0), "e3": (c0, c1), "e4": (c1, c1), "e5": (c1, c2), "e6": (c2, c2),
                  "e7": (c2, c3), "e8": (c3, c3), "d9": (c3 + c2, c2), "d10": (c2 + c1, c1), "d11": (c1 + c0, c0),
                  "head": (c0, cout)}
        self.conv = {k: Conv(a, b, rng, zero_rows_from=(n_out if k == "head" else None))
                     for k, (a, b) in shapes.items()}
        self.step_count = 0
        self._buf = {}
        self._work = None

    def n_params(self):
        return sum(c.cout * 27 * c.cin + c.cout for c in self.conv.values())

    def _t(self, name, shape, dtype="bf16"):
        key = (name, tuple(shape), dtype)
        t = self._buf.get(key)
        if t is None:
            t = DevArray(shape, dtype)
            self._buf[key] = t
        return t

    def _workspace(self, D, H, W):
        L = lib()
        need = 0
        for k, c in self.conv.items():
            s = 1 if k in ("e1", "e2", "d11", "head") else (2 if k in ("e3", "e4", "d10") else (4 if k in ("e5", "e6", "d9") else 8))
            d, h, w = D // s, H // s, W // s
            need = max(need, L.mfnn_wgrad_workspace(d, h, w, c.kp, c.cout, wgrad_chunk(d, h, w)))
        if self._work is None or self._work.numel() < need:
            self._work = DevArray((need,), "f32")
        return self._work

    def forward(self, x, stream=None):
        """x: bf16 (D, H, W, cin) -> activations dict with the fp32 output under 'y'."""
        D, H, W = x.shape[:3]
        c0, c1, c2, c3 = self.widths
        s0, s1, s2, s3 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4), (D // 8, H // 8, W // 8)
        L = lib()
        a = {"x": x}

        def conv(name, inp, shape, relu=True, fp32=False):
            c = self.conv[name]
            if fp32:
                out = self._t(name, shape + (c.cout,), "f32")
                c.forward(inp, shape, relu, yf=out, stream=stream)
            else:
                out = self._t(name, shape + (c.cout,))
                c.forward(inp, shape, relu, yb=out, stream=stream)
            a[name] = out
            return out

        def pool(name, inp, shape, C_):
            out = self._t(name, (shape[0] // 2, shape[1] // 2, shape[2] // 2, C_))
            _chk(L.mfnn_pool2(ptr(inp), *shape, C_, ptr(out), 0, stream), "pool")
            a[name] = out
            return out

        def cat(name, low, lowC, skip, skipC, shape):
            out = self._t(name, shape + (lowC + skipC,))
            _chk(L.mfnn_up2(ptr(low), *shape, lowC, lowC + skipC, 0, ptr(out), 0, stream), "up")
            _chk(L.mfnn_channels(ptr(skip), int(np.prod(shape)), skipC, lowC + skipC, lowC, 1, ptr(out), stream), "cat")
            a[name] = out
            return out

        conv("e1", x, s0)
        conv("e2", a["e1"], s0)
        pool("p1", a["e2"], s0, c0)
        conv("e3", a["p1"], s1)
        conv("e4", a["e3"], s1)
        pool("p2", a["e4"], s1, c1)
        conv("e5", a["p2"], s2)
        conv("e6", a["e5"], s2)
        pool("p3", a["e6"], s2, c2)
        conv("e7", a["p3"], s3)
        conv("e8", a["e7"], s3)
        cat("c1", a["e8"], c3, a["e6"], c2, s2)
        conv("d9", a["c1"], s2)
        cat("c2", a["d9"], c2, a["e4"], c1, s1)
        conv("d10", a["c2"], s1)
        cat("c3", a["d10"], c1, a["e2"], c0, s0)
        conv("d11", a["c3"], s0)
        conv("head", a["d11"], s0, relu=False, fp32=True)
        a["y"] = a["head"]
        return a

    def backward(self, a, gy, accumulate=False, stream=None):
        """Gradients of every layer from the bf16 output gradient gy (D, H, W, cout)."""
        D, H, W = a["x"].shape[:3]
        c0, c1, c2, c3 = self.widths
        s0, s1, s2, s3 = (D, H, W), (D // 2, H // 2, W // 2), (D // 4, H // 4, W // 4), (D // 8, H // 8, W // 8)
        L = lib()
        work = self._workspace(D, H, W)
        acc = int(accumulate)
        cv = self.conv
        g = lambda name, shape, C_: self._t("g_" + name, shape + (C_,))

        def split_up(gcat, lowC, skipC, shape_fine, name_low, name_skip, gate_low):
            # gradient of the upsampled part -> coarse grid (then the ReLU of the coarse
            # activation), gradient of the skip part -> its own buffer
            shape_c = (shape_fine[0] // 2, shape_fine[1] // 2, shape_fine[2] // 2)
            glow = g(name_low + "_u", shape_c, lowC)
            _chk(L.mfnn_up2(ptr(gcat), *shape_fine, lowC, lowC + skipC, 0, ptr(glow), 1, stream), "up bwd")
            glow_g = g(name_low, shape_c, lowC)
            _chk(L.mfnn_add_gate(ptr(glow), None, ptr(gate_low), glow.numel(), ptr(glow_g), stream), "gate")
            gskip = g(name_skip + "_s", shape_fine, skipC)
            _chk(L.mfnn_channels(ptr(gcat), int(np.prod(shape_fine)), skipC, lowC + skipC, lowC, 0, ptr(gskip), stream),
                 "split")
            return glow_g, gskip

        def merge_pool(gskip, gpooled, shape_fine, C_, gate, name):
            gp = g(name + "_p", shape_fine, C_)
            _chk(L.mfnn_pool2(ptr(gpooled), *shape_fine, C_, ptr(gp), 1, stream), "pool bwd")
            out = g(name, shape_fine, C_)
            _chk(L.mfnn_add_gate(ptr(gskip), ptr(gp), ptr(gate), out.numel(), ptr(out), stream), "merge")
            return out

        gd11 = g("d11", s0, c0)
        cv["head"].backward(a["d11"], gy, s0, work, acc, dx=gd11, gate=a["d11"], stream=stream)
        gc3 = g("c3", s0, c1 + c0)
        cv["d11"].backward(a["c3"], gd11, s0, work, acc, dx=gc3, stream=stream)
        gd10, ge2_a = split_up(gc3, c1, c0, s0, "d10", "e2", a["d10"])
        gc2 = g("c2", s1, c2 + c1)
        cv["d10"].backward(a["c2"], gd10, s1, work, acc, dx=gc2, stream=stream)
        gd9, ge4_a = split_up(gc2, c2, c1, s1, "d9", "e4", a["d9"])
        gc1 = g("c1", s2, c3 + c2)
        cv["d9"].backward(a["c1"], gd9, s2, work, acc, dx=gc1, stream=stream)
        ge8, ge6_a = split_up(gc1, c3, c2, s2, "e8", "e6", a["e8"])
        ge7 = g("e7", s3, c3)
        cv["e8"].backward(a["e7"], ge8, s3, work, acc, dx=ge7, gate=a["e7"], stream=stream)
        gp3 = g("p3", s3, c2)
        cv["e7"].backward(a["p3"], ge7, s3, work, acc, dx=gp3, stream=stream)
        ge6 = merge_pool(ge6_a, gp3, s2, c2, a["e6"], "e6")
        ge5 = g("e5", s2, c2)
        cv["e6"].backward(a["e5"], ge6, s2, work, acc, dx=ge5, gate=a["e5"], stream=stream)
        gp2 = g("p2", s2, c1)
        cv["e5"].backward(a["p2"], ge5, s2, work, acc, dx=gp2, stream=stream)
        ge4 = merge_pool(ge4_a, gp2, s1, c1, a["e4"], "e4")
        ge3 = g("e3", s1, c1)
        cv["e4"].backward(a["e3"], ge4, s1, work, acc, dx=ge3, gate=a["e3"], stream=stream)
        gp1 = g("p1", s1, c0)
        cv["e3"].backward(a["p1"], ge3, s1, work, acc, dx=gp1, stream=stream)
        ge2 = merge_pool(ge2_a, gp1, s0, c0, a["e2"], "e2")
        ge1 = g("e1", s0, c0)
        cv["e2"].backward(a["e1"], ge2, s0, work, acc, dx=ge1, gate=a["e1"], stream=stream)
        cv["e1"].backward(a["x"], ge1, s0, work, acc, stream=stream)

    def loss_grad(self, y, target, mask, stream=None):
        """masked squared error of y (fp32) against target (fp32) -> (loss, bf16 gradient)"""
        L = lib()
        M = int(np.prod(y.shape[:3]))
        cout = y.shape[3]
        nb = L.mfnn_mse_partials(M, cout)
        part = self._t("mse_part", (nb,), "f64")
        gy = self._t("gy", tuple(y.shape))
        inv = 1.0 / (M * cout)
        _chk(L.mfnn_mse(ptr(y), ptr(target), ptr(mask), M, cout, inv, ptr(gy), ptr(part), stream), "mse")
        return part, inv, gy

    def adam(self, lr, stream=None):
        self.step_count += 1
        for c in self.conv.values():
            c.adam(self.step_count, lr, stream)

    def params_numpy(self):
        return {k: c.get_numpy() for k, c in self.conv.items()}

    def grads_numpy(self):
        return {k: (c.dw.numpy()[:c.cout, :27 * c.cin].astype(np.float64), c.db.numpy().astype(np.float64))
                for k, c in self.conv.items()}

    def state(self):
        """weights as numpy arrays (for saving)"""
        out = {}
        for k, c in self.conv.items():
            w, b = c.get_numpy()
            out[k + ".w"] = w.astype(np.float32)
            out[k + ".b"] = b.astype(np.float32)
        return out

    def load_state(self, st):
        for k, c in self.conv.items():
            c.set_numpy(st[k + ".w"], st[k + ".b"])
