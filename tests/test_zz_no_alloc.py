"""The weight gradient without a workspace argument (mfnn_conv3_wgrad_static), for callers
that cannot allocate device memory: called through the C ABI, its weight and bias gradients
must equal those of mfnn_conv3_wgrad (same kernels, same chunk order: bit for bit), on two
calls in a row, the second on a smaller grid (the fixed scratch keeps the first call's
partials, which the second must not read), and with accumulation."""
import ctypes as C
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-no-alloc] {msg}", flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def nn():
    from mflbm import nn as mod
    mod.lib()
    return mod


def grads(nn, L, x, dz, shape, cin, cout, kp, static, accumulate=False, dw=None, db=None):
    D, H, W = shape
    chunk = nn.wgrad_chunk(D, H, W)
    coutp = nn.rup(cout, 16)
    if dw is None:
        dw = nn.DevArray((coutp, kp), "f32", zero=True)
        db = nn.DevArray((cout,), "f32", zero=True)
    if static:
        rc = L.mfnn_conv3_wgrad_static(nn.ptr(x), nn.ptr(dz), D, H, W, cin, kp, cout, chunk, nn.ptr(dw), nn.ptr(db),
                                       int(accumulate), None)
    else:
        work = nn.DevArray((L.mfnn_wgrad_workspace(D, H, W, kp, cout, chunk),), "f32")
        rc = L.mfnn_conv3_wgrad(nn.ptr(x), nn.ptr(dz), D, H, W, cin, kp, cout, chunk, nn.ptr(work), nn.ptr(dw),
                                nn.ptr(db), int(accumulate), None)
    assert rc == 0, rc
    assert L.mfnn_sync() == 0
    return dw, db


@pytest.mark.parametrize("first, second", [
    ((16, 32, (32, 32, 64)), (8, 16, (12, 10, 14))),
    ((32, 16, (24, 24, 32)), (16, 8, (8, 8, 8))),
])
def test_static_wgrad_equals_workspace_wgrad(nn, first, second):
    L = nn.lib()
    rng = np.random.default_rng(7)
    for cin, cout, shape in (first, second):
        say(f"wgrad_static cin={cin} cout={cout} shape={shape}")
        kp = nn.rup(27 * cin, 32)
        x = nn.DevArray.from_numpy(rng.normal(size=shape + (cin,)), "bf16")
        dz = nn.DevArray.from_numpy(rng.normal(size=shape + (cout,)), "bf16")
        dw_ref, db_ref = grads(nn, L, x, dz, shape, cin, cout, kp, static=False)
        dw, db = grads(nn, L, x, dz, shape, cin, cout, kp, static=True)
        assert np.array_equal(dw.numpy(), dw_ref.numpy())
        assert np.array_equal(db.numpy(), db_ref.numpy())
        # accumulate onto the previous gradients
        grads(nn, L, x, dz, shape, cin, cout, kp, static=False, accumulate=True, dw=dw_ref, db=db_ref)
        grads(nn, L, x, dz, shape, cin, cout, kp, static=True, accumulate=True, dw=dw, db=db)
        assert np.array_equal(dw.numpy(), dw_ref.numpy())
        assert np.array_equal(db.numpy(), db_ref.numpy())
        assert np.abs(dw_ref.numpy()).max() > 0


def test_static_wgrad_rejects_sizes_beyond_its_scratch(nn):
    L = nn.lib()
    cin, cout, shape = 64, 128, (64, 64, 64)
    kp = nn.rup(27 * cin, 32)
    chunk = nn.wgrad_chunk(*shape)
    assert L.mfnn_wgrad_workspace(*shape, kp, cout, chunk) > 16 * 1024 * 1024
    x = nn.DevArray((16,), "f32")
    assert L.mfnn_conv3_wgrad_static(nn.ptr(x), nn.ptr(x), *shape, cin, kp, cout, chunk, None, None, 0, None) == -3
