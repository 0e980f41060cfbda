"""Geometry files and seeded porous-medium generators.

The reference reads a binary ``.dat`` file (src/Misc.cpp:139-192): three integers
(4 or 8 bytes each, ``geometry_dims_type_size``) giving the sample size, followed by
one byte per voxel (1 solid, 0 pore), x fastest.  ``read_dat``/``write_dat`` keep that
format so a generated geometry can be fed to the reference unchanged.

The reference indexes the voxel bytes as ``(i-1) + ny*((j-1) + nx*(k-1))``
(src/Misc.cpp:171), which is the plain x-fastest order only when nx == ny; the files
written here are x-fastest, which is what the reference reads whenever nx == ny.

Arrays use numpy shape (nz, ny, nx): the reference's linear index i + nx*(j + ny*k)
is numpy's C order.
"""
from __future__ import annotations

import numpy as np


def write_dat(path: str, solid: np.ndarray, dims_type_size: int = 4) -> None:
    solid = np.asarray(solid)
    nz, ny, nx = solid.shape
    dt = np.int32 if dims_type_size == 4 else np.int64
    with open(path, "wb") as fh:
        fh.write(np.array([nx, ny, nz], dtype=dt).tobytes())
        fh.write(np.ascontiguousarray(solid, dtype=np.int8).tobytes())


def read_dat(path: str, dims_type_size: int = 4) -> np.ndarray:
    dt = np.int32 if dims_type_size == 4 else np.int64
    with open(path, "rb") as fh:
        hdr = np.frombuffer(fh.read(3 * np.dtype(dt).itemsize), dtype=dt)
        nx, ny, nz = (int(v) for v in hdr)
        data = np.frombuffer(fh.read(nx * ny * nz), dtype=np.int8)
    if data.size != nx * ny * nz:
        raise ValueError(f"{path}: expected {nx * ny * nz} voxels, found {data.size}")
    return data.reshape(nz, ny, nx).copy()


def sphere_pack(shape, radius, porosity, seed, periodic=(False, False, False), max_spheres=10_000_000):
    """Random overlapping solid spheres placed until the pore fraction reaches ``porosity``.

    shape = (nz, ny, nx); ``radius`` is a number or a (lo, hi) range drawn per sphere.
    Centres and radii come from numpy's PCG64 stream seeded with ``seed``.  Spheres are
    added in batches and the pore fraction is checked after each batch; the last batch is
    trimmed sphere by sphere so the result is the first prefix of the stream that reaches
    the target.  ``periodic`` wraps spheres across the (z, y, x) faces.
    """
    nz, ny, nx = shape
    rng = np.random.default_rng(seed)
    solid = np.zeros(shape, dtype=bool)
    total = solid.size
    target_solid = (1.0 - porosity) * total
    placed = 0

    def stamp(mask, cz, cy, cx, r):
        ri = int(np.ceil(r))
        zz = np.arange(int(np.floor(cz)) - ri, int(np.floor(cz)) + ri + 2)
        yy = np.arange(int(np.floor(cy)) - ri, int(np.floor(cy)) + ri + 2)
        xx = np.arange(int(np.floor(cx)) - ri, int(np.floor(cx)) + ri + 2)
        d2 = ((zz[:, None, None] - cz) ** 2 + (yy[None, :, None] - cy) ** 2 + (xx[None, None, :] - cx) ** 2)
        inside = d2 <= r * r
        zi, yi, xi = zz, yy, xx
        if periodic[0]:
            zi = zi % nz
        if periodic[1]:
            yi = yi % ny
        if periodic[2]:
            xi = xi % nx
        okz = (zi >= 0) & (zi < nz)
        oky = (yi >= 0) & (yi < ny)
        okx = (xi >= 0) & (xi < nx)
        sub = inside[np.ix_(okz, oky, okx)]
        idx = np.ix_(zi[okz], yi[oky], xi[okx])
        mask[idx] |= sub

    while placed < max_spheres:
        n_solid = solid.sum()
        if n_solid >= target_solid:
            break
        lo_hi = radius if isinstance(radius, (tuple, list)) else (radius, radius)
        # expected spheres still needed, from the overlap-free volume estimate
        vol = 4.0 / 3.0 * np.pi * (0.5 * (lo_hi[0] + lo_hi[1])) ** 3
        batch = int(max(1, min(2048, 0.5 * (target_solid - n_solid) / vol)))
        c = rng.random((batch, 3)) * np.array([nz, ny, nx]) + 0.5
        r = rng.uniform(lo_hi[0], lo_hi[1], batch) if lo_hi[0] != lo_hi[1] else np.full(batch, float(lo_hi[0]))
        for b in range(batch):
            stamp(solid, c[b, 0], c[b, 1], c[b, 2], r[b])
            placed += 1
            if batch > 1 and b % 16 == 15 and solid.sum() >= target_solid:
                break
    return solid.astype(np.int8)


def gaussian_random_field(shape, corr_len, porosity, seed):
    """Seeded Gaussian random field with Gaussian covariance of length ``corr_len``,
    thresholded so that the pore fraction equals ``porosity`` (lowest values are pore)."""
    rng = np.random.default_rng(seed)
    noise = rng.standard_normal(shape).astype(np.float32)
    spec = np.fft.rfftn(noise)
    kz = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kx = np.fft.rfftfreq(shape[2])[None, None, :]
    k2 = (2 * np.pi) ** 2 * (kz ** 2 + ky ** 2 + kx ** 2)
    spec *= np.exp(-0.25 * k2 * corr_len ** 2).astype(np.complex64)
    field = np.fft.irfftn(spec, s=shape).astype(np.float32)
    thr = np.quantile(field, porosity)
    return (field > thr).astype(np.int8)


def porosity_of(solid: np.ndarray) -> float:
    return float(1.0 - np.asarray(solid, dtype=np.float64).mean())
