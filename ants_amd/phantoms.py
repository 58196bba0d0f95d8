"""Seeded synthetic registration pairs (numpy/scipy, host side).

Each generator returns float32 arrays in numpy (z, y, x) / (y, x) order,
i.e. x-fastest buffers as the library and ITK expect, plus the grid.

* ``blob_pair_2d``   -- 256x256 Gaussian-blob pair (2-D SyN workload).
* ``phantom_pair_3d`` -- nested-ellipsoid phantom with cubic bias field and
  Rician noise; moving = fixed warped by an exponentiated smooth stationary
  velocity field (192x224x192 at 1 mm by default).
* ``voronoi_labels`` -- integer label map of Voronoi cells of random seeds
  inside the phantom.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional, Sequence, Tuple

import numpy as np
from scipy import ndimage


@dataclass
class Pair:
    fixed: np.ndarray          # float32, numpy order (z,y,x) / (y,x)
    moving: np.ndarray
    size: Tuple[int, ...]      # (nx, ny[, nz])
    spacing: Tuple[float, ...]
    origin: Tuple[float, ...]
    labels_fixed: Optional[np.ndarray] = None   # int32
    labels_moving: Optional[np.ndarray] = None
    true_disp: Optional[np.ndarray] = None      # displacement used to make `moving`, (...,D) in voxels of x,y,z


def _warp_linear(img: np.ndarray, disp_xyz: np.ndarray, order: int = 1) -> np.ndarray:
    """img(x + u(x)) with u given per voxel as (..., D) in (x,y[,z]) voxel units; 0 outside."""
    D = img.ndim
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in img.shape], indexing="ij")
    coords = []
    for ax in range(D):  # numpy axis ax corresponds to xyz component D-1-ax
        coords.append(grids[ax] + disp_xyz[..., D - 1 - ax])
    return ndimage.map_coordinates(img, coords, order=order, mode="constant", cval=0.0)


def _compose_disp(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """(a o b)(x) = b(x) + a(x + b(x)) for displacement fields in voxel units (edge-clamped)."""
    D = a.shape[-1]
    shape = a.shape[:-1]
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    coords = [grids[ax] + b[..., D - 1 - ax] for ax in range(D)]
    out = np.empty_like(b)
    for c in range(D):
        out[..., c] = b[..., c] + ndimage.map_coordinates(a[..., c], coords, order=1, mode="nearest")
    return out


def _exp_svf(v: np.ndarray, steps: int = 6) -> np.ndarray:
    """Scaling and squaring of a stationary velocity field (voxel units)."""
    phi = v / (2.0 ** steps)
    for _ in range(steps):
        phi = _compose_disp(phi, phi)
    return phi


def _smooth_noise_field(rng, shape, D, sigma_vox) -> np.ndarray:
    f = np.empty(tuple(shape) + (D,))
    for c in range(D):
        f[..., c] = ndimage.gaussian_filter(rng.standard_normal(shape), sigma_vox, mode="nearest")
    return f


def blob_pair_2d(seed: int = 0, n: int = 256, n_blobs: int = 20, max_disp: float = 6.0) -> Pair:
    """Fixed: 20 isotropic Gaussian blobs (centres U[0,n)^2, sigma U[8,30] px,
    amplitude U[0.2,1]) + N(0, 0.01^2).  Moving: fixed warped by white noise
    smoothed at sigma 12 px rescaled to `max_disp` px max magnitude, then
    intensity * 1.1 + 0.05."""
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid(np.arange(n, dtype=np.float64), np.arange(n, dtype=np.float64), indexing="ij")
    img = np.zeros((n, n))
    for _ in range(n_blobs):
        cx, cy = rng.uniform(0, n, 2)
        s = rng.uniform(8, 30)
        a = rng.uniform(0.2, 1.0)
        img += a * np.exp(-((xx - cx) ** 2 + (yy - cy) ** 2) / (2 * s * s))
    img += rng.normal(0.0, 0.01, img.shape)
    u = _smooth_noise_field(rng, (n, n), 2, 12.0)
    mag = np.sqrt((u ** 2).sum(-1)).max()
    u *= max_disp / mag
    mov = _warp_linear(img, u) * 1.1 + 0.05
    return Pair(img.astype(np.float32), mov.astype(np.float32), (n, n), (1.0, 1.0), (0.0, 0.0), true_disp=u)


def _ellipsoid_labels(shape_zyx, spacing, semi_axes_mm) -> np.ndarray:
    nz, ny, nx = shape_zyx
    z = (np.arange(nz) - (nz - 1) / 2.0) * spacing[2]
    y = (np.arange(ny) - (ny - 1) / 2.0) * spacing[1]
    x = (np.arange(nx) - (nx - 1) / 2.0) * spacing[0]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    lab = np.zeros(shape_zyx, dtype=np.int32)
    for k, (ax, ay, az) in enumerate(semi_axes_mm):
        inside = (X / ax) ** 2 + (Y / ay) ** 2 + (Z / az) ** 2 <= 1.0
        lab[inside] = k + 1
    return lab


def _cubic_bias(rng, shape_zyx, amplitude=0.2) -> np.ndarray:
    nz, ny, nx = shape_zyx
    z = np.linspace(-1, 1, nz)
    y = np.linspace(-1, 1, ny)
    x = np.linspace(-1, 1, nx)
    p = np.zeros(shape_zyx)
    for i in range(4):
        for j in range(4 - i):
            for k in range(4 - i - j):
                c = rng.standard_normal()
                p += c * (x[None, None, :] ** i) * (y[None, :, None] ** j) * (z[:, None, None] ** k)
    p -= p.mean()
    p *= amplitude / np.abs(p).max()
    return p


def phantom_pair_3d(seed: int = 0, size=(192, 224, 192), spacing=(1.0, 1.0, 1.0),
                    max_disp_mm: float = 8.0, svf_sigma_mm: float = 10.0, noise: float = 0.02) -> Pair:
    """Nested ellipsoids (semi-axes (80,95,80), (70,85,70), (40,50,40) mm for
    192x224x192, scaled with the volume otherwise), labels 1/2/3, intensities
    0.3/0.6/0.9, times (1 + cubic bias in +-20%), Rician noise sigma `noise`.
    Moving = fixed o exp(v), v Gaussian-smoothed noise (sigma 10 mm) scaled to
    `max_disp_mm` max displacement."""
    rng = np.random.default_rng(seed)
    nx, ny, nz = size
    shape = (nz, ny, nx)
    scale = np.array([nx / 192.0, ny / 224.0, nz / 192.0])
    axes = [tuple(np.array(a) * scale) for a in [(80, 95, 80), (70, 85, 70), (40, 50, 40)]]
    lab = _ellipsoid_labels(shape, spacing, axes)
    inten = np.array([0.0, 0.3, 0.6, 0.9])[lab]
    inten = inten * (1.0 + _cubic_bias(rng, shape))
    n1 = rng.normal(0.0, noise, shape)
    n2 = rng.normal(0.0, noise, shape)
    fixed = np.sqrt((inten + n1) ** 2 + n2 ** 2)
    # velocity field on a half-resolution grid, exponentiated there, upsampled
    cs = 2
    cshape = tuple(max(2, s // cs) for s in shape)
    v = _smooth_noise_field(rng, cshape, 3, svf_sigma_mm / (cs * spacing[0]))
    phi = None
    target = max_disp_mm / (cs * spacing[0])
    for _ in range(3):
        phi = _exp_svf(v)
        m = np.sqrt((phi ** 2).sum(-1)).max()
        v *= target / m
    phi = _exp_svf(v)
    zoom = [s / c for s, c in zip(shape, cshape)]
    up = np.stack([ndimage.zoom(phi[..., c], zoom, order=1) for c in range(3)], -1)[:nz, :ny, :nx] * cs
    up = up / np.array(spacing)[None, None, None, :]  # mm -> voxels per axis
    moving = _warp_linear(fixed, up)
    lab_m = _warp_linear(lab.astype(np.float64), up, order=0).astype(np.int32)
    return Pair(fixed.astype(np.float32), moving.astype(np.float32), tuple(size), tuple(spacing), (0.0, 0.0, 0.0),
                labels_fixed=lab, labels_moving=lab_m, true_disp=up)


def voronoi_labels(seed: int, mask: np.ndarray, n_labels: int = 32) -> np.ndarray:
    """Voronoi cells (labels 1..n) of `n_labels` random seeds drawn inside `mask`; 0 outside."""
    rng = np.random.default_rng(seed)
    idx = np.argwhere(mask)
    seeds = idx[rng.choice(len(idx), n_labels, replace=False)].astype(np.float64)
    out = np.zeros(mask.shape, dtype=np.int32)
    pts = np.argwhere(mask).astype(np.float64)
    best = np.full(len(pts), np.inf)
    lab = np.zeros(len(pts), dtype=np.int32)
    for k, s in enumerate(seeds):
        d = ((pts - s) ** 2).sum(-1)
        better = d < best
        best[better] = d[better]
        lab[better] = k + 1
    out[tuple(pts.astype(np.int64).T)] = lab
    return out
