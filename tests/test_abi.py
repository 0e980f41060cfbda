"""The C-ABI library loads and exports every symbol include/mflbm.h declares (no GPU calls)."""
import ctypes
import os
import re

from mflbm import gpu


def declared_symbols():
    with open(gpu.HEADER) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"\b(mflbm_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_entry_points():
    assert declared_symbols() == sorted(gpu.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_struct_sizes_match_header_layout():
    # mflbm_params: 3 int64 + 11 int32 (+4 pad) + 11 double + 4 int32 + 1 double + 1 int32 (+4 pad)
    assert ctypes.sizeof(gpu.Params) == 3 * 8 + 11 * 4 + 4 + 11 * 8 + 4 * 4 + 8 + 4 + 4
    assert gpu.Params.uin_avg.offset == 3 * 8 + 11 * 4 + 4 + 11 * 8 + 4 * 4
    assert gpu.Params.mrt.offset == 3 * 8 + 11 * 4 + 4 + 11 * 8 + 4 * 4 + 8
    assert ctypes.sizeof(gpu.MonitorOut) == 7 * 8 + 3 * 8 + 8


def test_network_header_matches_library_and_bindings():
    from mflbm import nn
    with open(os.path.join(os.path.dirname(gpu.HEADER), "mfnn.h")) as fh:
        declared = sorted(set(re.findall(r"\b(mfnn_[a-z_0-9]+)\s*\(", fh.read())))
    lib = ctypes.CDLL(gpu.LIB_PATH)
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, missing
    assert declared == sorted(nn.SIGNATURES)


def test_network_argument_checks_reject_before_any_launch():
    """The convolution entry points validate their sizes and operand alignment on the host and
    return -1 without touching the device, so these calls are safe without a GPU."""
    from mflbm import nn
    L = nn.lib()
    buf = ctypes.create_string_buffer(64)
    aligned = (ctypes.addressof(buf) + 15) // 16 * 16
    # weight gradient: cout not a multiple of 8, cin not a multiple of 8, kp too small,
    # chunk not a multiple of 32, misaligned operands
    assert L.mfnn_conv3_wgrad(aligned, aligned, 4, 4, 4, 8, 224, 12, 32, None, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad(aligned, aligned, 4, 4, 4, 12, 352, 8, 32, None, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad(aligned, aligned, 4, 4, 4, 8, 192, 8, 32, None, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad(aligned, aligned, 4, 4, 4, 8, 224, 8, 48, None, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad(aligned + 2, aligned, 4, 4, 4, 8, 224, 8, 32, None, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad(aligned, aligned + 8, 4, 4, 4, 8, 224, 8, 32, None, None, None, 0, None) == -1
    # weight gradient: an empty grid
    assert L.mfnn_conv3_wgrad(aligned, aligned, 0, 4, 4, 8, 224, 8, 32, None, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad(aligned, aligned, 4, 4, -1, 8, 224, 8, 32, None, None, None, 0, None) == -1
    # the weight gradient without a workspace: the same checks, and -3 for a workspace
    # larger than its fixed scratch (16 Mi floats), all before touching the device
    assert L.mfnn_conv3_wgrad_static(aligned, aligned, 4, 4, 4, 8, 224, 12, 32, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad_static(aligned + 2, aligned, 4, 4, 4, 8, 224, 8, 32, None, None, 0, None) == -1
    assert L.mfnn_conv3_wgrad_static(aligned, aligned, 0, 4, 4, 8, 224, 8, 32, None, None, 0, None) == -1
    assert L.mfnn_wgrad_workspace(64, 64, 64, 1728, 128, 2048) > 16 * 1024 * 1024
    assert L.mfnn_conv3_wgrad_static(aligned, aligned, 64, 64, 64, 64, 1728, 128, 2048, None, None, 0, None) == -3
    # forward: more than 192 output channels, an empty grid, misaligned input or weights
    assert L.mfnn_conv3(aligned, aligned, None, None, 4, 4, 4, 8, 224, 200, 1, aligned, None, None) == -1
    assert L.mfnn_conv3(aligned, aligned, None, None, 0, 4, 4, 8, 224, 16, 1, aligned, None, None) == -1
    assert L.mfnn_conv3(aligned + 2, aligned, None, None, 4, 4, 4, 8, 224, 16, 1, aligned, None, None) == -1
    assert L.mfnn_conv3(aligned, aligned + 8, None, None, 4, 4, 4, 8, 224, 16, 1, aligned, None, None) == -1
