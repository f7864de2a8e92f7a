"""Synthetic inputs: modified Shepp-Logan phantoms and random-ellipsoid volumes.

The 10 ellipsoids are the published modified Shepp-Logan table (Toft; Kak & Slaney) that
elsa's phantom::modifiedSheppLogan also draws (elsa/generators/Phantoms.cpp:71-148); this is
input data for the projectors, rasterised by point sampling at voxel centres.
"""
from __future__ import annotations

import numpy as np

# amplitude, semi-axes (a, b, c), centre (x, y, z), rotation phi about z (degrees)
_MSL = [
    (1.0, 0.6900, 0.920, 0.810, 0.0, 0.0, 0.0, 0.0),
    (-0.8, 0.6624, 0.874, 0.780, 0.0, -0.0184, 0.0, 0.0),
    (-0.2, 0.1100, 0.310, 0.220, 0.22, 0.0, 0.0, -18.0),
    (-0.2, 0.1600, 0.410, 0.280, -0.22, 0.0, 0.0, 18.0),
    (0.1, 0.2100, 0.250, 0.410, 0.0, 0.35, -0.15, 0.0),
    (0.1, 0.0460, 0.046, 0.050, 0.0, 0.1, 0.25, 0.0),
    (0.1, 0.0460, 0.046, 0.050, 0.0, -0.1, 0.25, 0.0),
    (0.1, 0.0460, 0.023, 0.050, -0.08, -0.605, 0.0, 0.0),
    (0.1, 0.0230, 0.023, 0.020, 0.0, -0.606, 0.0, 0.0),
    (0.1, 0.0230, 0.046, 0.020, 0.06, -0.605, 0.0, 0.0),
]


def modified_shepp_logan(shape, dtype=np.float32) -> np.ndarray:
    """Volume of C-order shape ([z,] y, x) with values in [0, 1]."""
    dim = len(shape)
    if dim not in (2, 3):
        raise ValueError("only 2D and 3D phantoms")
    shape3 = (1,) + tuple(shape) if dim == 2 else tuple(shape)
    nz, ny, nx = shape3
    zs = (np.arange(nz, dtype=np.float32) + 0.5) / nz * 2 - 1 if dim == 3 else np.zeros(1, np.float32)
    ys = (np.arange(ny, dtype=np.float32) + 0.5) / ny * 2 - 1
    xs = (np.arange(nx, dtype=np.float32) + 0.5) / nx * 2 - 1
    out = np.zeros(shape3, np.float32)
    for amp, a, b, c, x0, y0, z0, phi in _MSL:
        if dim == 2 and abs(z0) > 0:
            continue
        ph = np.deg2rad(phi)
        cp, sp = np.float32(np.cos(ph)), np.float32(np.sin(ph))
        r = max(a, b)
        # index ranges of the bounding box of the ellipsoid
        def rng(coords, lo, hi):
            idx = np.nonzero((coords >= lo) & (coords <= hi))[0]
            return (int(idx[0]), int(idx[-1]) + 1) if idx.size else (0, 0)
        zr = rng(zs, z0 - c, z0 + c) if dim == 3 else (0, 1)
        yr = rng(ys, y0 - r, y0 + r)
        xr_ = rng(xs, x0 - r, x0 + r)
        if zr[0] == zr[1] or yr[0] == yr[1] or xr_[0] == xr_[1]:
            continue
        z = zs[zr[0]:zr[1], None, None]
        y = ys[None, yr[0]:yr[1], None]
        x = xs[None, None, xr_[0]:xr_[1]]
        xr = (x - x0) * cp + (y - y0) * sp
        yr2 = -(x - x0) * sp + (y - y0) * cp
        q = (xr / a) ** 2 + (yr2 / b) ** 2
        if dim == 3:
            q = q + ((z - z0) / c) ** 2
        out[zr[0]:zr[1], yr[0]:yr[1], xr_[0]:xr_[1]] += np.where(q <= 1.0, np.float32(amp), np.float32(0))
    out = np.clip(out, 0.0, 1.0)
    return (out[0] if dim == 2 else out).astype(dtype)


def random_ellipsoids(shape, n=20, seed=0, dtype=np.float32) -> np.ndarray:
    """Sum of n random ellipsoids with amplitudes U(0, 1): centres U(-0.6, 0.6), semi-axes
    U(0.05, 0.4) of the half-width, uniform random rotations."""
    rng = np.random.default_rng(seed)
    dim = len(shape)
    axes = [(np.arange(s, dtype=np.float32) + 0.5) / s * 2 - 1 for s in shape]
    grid = np.stack(np.meshgrid(*axes, indexing="ij"), -1)  # (..., dim) in C order
    out = np.zeros(shape, np.float32)
    for _ in range(n):
        amp = rng.uniform(0, 1)
        ctr = rng.uniform(-0.6, 0.6, dim)
        semi = rng.uniform(0.05, 0.4, dim)
        q, _ = np.linalg.qr(rng.standard_normal((dim, dim)))
        d = (grid - ctr) @ q
        out[np.sum((d / semi) ** 2, -1) <= 1.0] += amp
    return out.astype(dtype)
