"""Host-side field diagnostics used in reports (not on the GPU path).

``jacobian_determinant``: det(I + grad u) of a displacement field by central
differences (one-sided at the edges), u in mm on a grid of the given
spacing; numpy (z,y,x,D) / (y,x,D) layout with components in x,y,z order.
"""
from __future__ import annotations

import numpy as np


def jacobian_determinant(u: np.ndarray, spacing) -> np.ndarray:
    D = u.shape[-1]
    sp = list(spacing)
    J = np.zeros(u.shape[:-1] + (D, D))
    for c in range(D):  # component x, y, z
        for d in range(D):  # derivative along x, y, z -> numpy axis D-1-d
            J[..., c, d] = np.gradient(u[..., c].astype(np.float64), sp[d], axis=D - 1 - d)
        J[..., c, c] += 1.0
    return np.linalg.det(J)


def field_errors(a: np.ndarray, b: np.ndarray, spacing) -> dict:
    """Max and mean |a - b| in voxels, and Jacobian-determinant differences."""
    sp = np.asarray(spacing, dtype=np.float64)
    d = np.sqrt((((a.astype(np.float64) - b.astype(np.float64)) / sp) ** 2).sum(-1))
    ja, jb = jacobian_determinant(a, spacing), jacobian_determinant(b, spacing)
    return {"max_vox": float(d.max()), "mean_vox": float(d.mean()),
            "jac_max_abs": float(np.abs(ja - jb).max()), "jac_mean_abs": float(np.abs(ja - jb).mean())}
