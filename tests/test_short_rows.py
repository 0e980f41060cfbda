"""The 16-bit neighbour rows of the odd collide (encode_short_rows in
mf-lbm-cuda_amd/csrc/layout.h), checked on the host against the format's definition:

* a wave of 64 nodes stores, per direction q and class c (0: cell < nf, 1: extra cell), the
  base nbb[w * 40 + 2 (q - 1) + c]; word 36 flags a wave that keeps the 32-bit rows;
* each node stores (cell - base) | c << 15 in 16 bits, directions 2p + 1 and 2p + 2 in the
  low and high half of nb16[p * nf + t];
* a wave is flagged exactly when the links of one class, over its valid nodes, span 2^15
  cells or more; nodes that are not valid (padding) are left out.

Waves are built with spreads of exactly 32767 and 32768 so that an off-by-one in the 15-bit
capacity fails.  The device decodes the same words (`odd_links` in csrc/kernels.inl); its
results are checked bit for bit against the oracle by the GPU parity tests."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "mf-lbm-cuda_amd", "csrc")
NB_WAVE = 40


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("short_rows") / "short_rows_check")
    subprocess.run(["g++", "-std=c++17", "-O2", "-D__host__=", "-D__device__=", "-D__forceinline__=inline",
                    "-I" + CSRC, "-I" + os.path.join(os.path.dirname(HERE), "include"),
                    os.path.join(HERE, "short_rows_check.cpp"), "-o", out], check=True, timeout=120)
    return out


def encode(harness, tmp_path, nb, valid):
    nf = valid.size
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(src, "wb") as fh:
        fh.write(np.int64(nf).tobytes())
        fh.write(np.ascontiguousarray(nb, np.int32).tobytes())
        fh.write(np.ascontiguousarray(valid, np.uint8).tobytes())
    subprocess.run([harness, str(src), str(dst)], check=True, timeout=60)
    raw = np.fromfile(dst, np.uint8)
    nw = (nf + 63) // 64
    r = int(raw[:4].view(np.int32)[0])
    nbb = raw[4:4 + 4 * nw * NB_WAVE].view(np.int32).reshape(nw, NB_WAVE)
    nb16 = raw[4 + 4 * nw * NB_WAVE:].view(np.uint32).reshape(9, nf)
    return r, nbb, nb16


def expected_flags(nb, valid):
    nf = valid.size
    flags = []
    for w in range((nf + 63) // 64):
        t = np.arange(w * 64, min(nf, w * 64 + 64))
        t = t[valid[t] != 0]
        esc = False
        for q in range(18):
            v = nb[q * nf + t].astype(np.int64)
            for c in (0, 1):
                s = v[(v >= nf) == bool(c)]
                if s.size and s.max() - s.min() >= 32768:
                    esc = True
        flags.append(esc)
    return np.array(flags)


def decode(nbb, nb16, nf):
    """Cells of every (direction, node) from the short rows, as the device decodes them."""
    q = np.arange(18)[:, None]
    t = np.arange(nf)[None, :]
    pr = nb16[q // 2, t]
    u = np.where(q % 2 == 1, pr >> 16, pr & 0xFFFF).astype(np.int64)
    base = nbb[t // 64, 2 * q + (u >> 15)]
    return base + (u & 0x7FFF)


def test_random_links_and_exact_capacity(harness, tmp_path):
    rng = np.random.default_rng(5)
    nw = 1100
    nf = nw * 64 - 17  # a ragged last wave
    nb = np.empty((18, nf), np.int64)
    # fluid links near the node, extra links (>= nf) in a run per direction
    for q in range(18):
        nb[q] = np.clip(np.arange(nf) + rng.integers(-300, 300, nf), 0, nf - 1)
        ext = rng.random(nf) < 0.3
        nb[q, ext] = nf + np.arange(nf)[ext] // 2 + q

    def lowest(q, w, c):
        v = nb[q, w * 64:w * 64 + 64]
        v = v[(v >= nf) == bool(c)]
        return int(v.min()) if v.size else (nf if c else 0)

    # extra cells: wave 3 spans exactly 32767 cells (fits), wave 4 32768 (flagged)
    m = lowest(4, 3, 1)
    nb[4, 3 * 64 + 2], nb[4, 3 * 64 + 3] = m, m + 32767
    m = lowest(7, 4, 1)
    nb[7, 4 * 64 + 5], nb[7, 4 * 64 + 6] = m, m + 32768
    # wave 6: the wide link sits on a padding node, which is left out
    m = lowest(9, 6, 1)
    nb[9, 6 * 64 + 1], nb[9, 6 * 64 + 2] = m, m + 40000
    valid = np.ones(nf, np.uint8)
    valid[6 * 64 + 2] = 0
    # fluid cells: wave 7 spans exactly 32767 cells (fits), wave 8 32768 (flagged)
    m = lowest(11, 7, 0)
    nb[11, 7 * 64], nb[11, 7 * 64 + 63] = m, m + 32767
    m = lowest(12, 8, 0)
    nb[12, 8 * 64 + 10], nb[12, 8 * 64 + 20] = m, m + 32768
    assert m + 32768 < nf
    flat = nb.reshape(-1)
    assert flat.max() < 2 ** 31
    r, nbb, nb16 = encode(harness, tmp_path, flat.astype(np.int32), valid)
    flags = expected_flags(flat, valid)
    assert flags[4] and flags[8] and not flags[3] and not flags[6] and not flags[7]
    assert r == int(flags.sum())
    np.testing.assert_array_equal(nbb[:, 36] != 0, flags)
    cells = decode(nbb, nb16, nf)
    ok = (~flags[np.arange(nf) // 64]) & (valid != 0)
    np.testing.assert_array_equal(cells[:, ok], nb[:, ok])
    # padding nodes are not encoded
    assert (nb16[:, valid == 0] == 0).all()


def test_whole_domain_like_layout(harness, tmp_path):
    """Links of a small porous lattice laid out like the solver's cells (fluid nodes in x-fastest
    order, extra cells per slot sorted by node): no wave needs the 32-bit rows."""
    rng = np.random.default_rng(2)
    n = 24
    solid = rng.random((n, n, n)) < 0.35
    fluid = np.flatnonzero(~solid.reshape(-1))
    nf = fluid.size
    fidx = -np.ones(n ** 3, np.int64)
    fidx[fluid] = np.arange(nf)
    ex = [0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0]
    ey = [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1]
    ez = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
    x, y, z = fluid % n, (fluid // n) % n, fluid // (n * n)
    nb = np.empty((18, nf), np.int64)
    for q in range(1, 19):
        m = (x + ex[q]) % n + n * ((y + ey[q]) % n) + n * n * ((z + ez[q]) % n)
        f = fidx[m]
        ext = f < 0
        order = np.unique(m[ext], return_inverse=True)[1]
        f[ext] = nf + order
        nb[q - 1] = f
    valid = np.ones(nf, np.uint8)
    r, nbb, nb16 = encode(harness, tmp_path, nb.reshape(-1).astype(np.int32), valid)
    assert r == 0
    np.testing.assert_array_equal(decode(nbb, nb16, nf), nb)
