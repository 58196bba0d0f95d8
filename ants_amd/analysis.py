"""Field diagnostics.

``jacobian_determinant``: host-side det(I + grad u) of a displacement field by
central differences (one-sided at the edges), u in mm on a grid of the given
spacing; numpy (z,y,x,D) / (y,x,D) layout with components in x,y,z order; used
in reports.  ``create_jacobian_determinant_image`` is the GPU form of ANTs'
CreateJacobianDeterminantImage.
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


def create_jacobian_determinant_image(field, geometry, do_log: bool = False, deformation_gradient: bool = False,
                                      geometric: bool = False):
    """CreateJacobianDeterminantImage dim field out [doLogJacobian] [0]
    [deformationGradient] (Examples/CreateJacobianDeterminantImage.cxx:60-160):
    max(det(I + R-rotated five-point gradient), 0), or its log clamped at
    0.001, on the GPU (ants_jacobian_determinant).  ``field``: numpy or
    DeviceArray of shape geometry.shape + (dim,).  With
    ``deformation_gradient`` returns the (..., dim, dim) matrices instead;
    ``geometric`` is useGeometric (GeometricJacobianDeterminantImageFilter:
    moved-simplex volume ratios)."""
    import ctypes as C
    from . import _lib as L
    D = geometry.dim
    f = field if isinstance(field, L.DeviceArray) else L.DeviceArray.from_host(np.asarray(field, np.float32))
    if tuple(f.shape) != tuple(geometry.shape) + (D,):
        raise ValueError("field shape does not match its geometry")
    out = L.DeviceArray(geometry.shape)
    grad = L.DeviceArray(tuple(geometry.shape) + (D, D)) if deformation_gradient else None
    if geometric and deformation_gradient:
        raise ValueError("the geometric form has no gradient output")
    L.check(L.lib().ants_jacobian_determinant(C.byref(geometry.desc()), f.ptr, int(bool(do_log)), int(bool(geometric)),
                                              out.ptr, grad.ptr if grad is not None else None, None))
    return grad.to_host() if deformation_gradient else out.to_host()


def measure_cc_similarity(fixed, fixed_geometry, moving, moving_geometry, radius: int) -> float:
    """MeasureImageSimilarity -m CC[fixed, moving, 1, radius]
    (Examples/MeasureImageSimilarity.cxx:205-220, 312-414): the
    ANTSNeighborhoodCorrelation metric value -mean(local CC) on the fixed
    image's grid with identity transforms, on the GPU (ants_op_cc); a
    moving image on another grid is first resampled onto the fixed one
    (linear, 0 outside, as the metric's interpolator samples it)."""
    import ctypes as C
    from . import _lib as L
    from .apply_transforms import apply_transforms
    if radius < 1:
        raise ValueError("CC radius must be >= 1")
    f = np.asarray(fixed, np.float32)
    m = np.asarray(moving, np.float32)
    def parts(gm):
        D = gm.dim
        sp = np.ones(D) if gm.spacing is None else np.asarray(gm.spacing, float)
        org = np.zeros(D) if gm.origin is None else np.asarray(gm.origin, float)
        dr = np.eye(D) if gm.direction is None else np.asarray(gm.direction, float).reshape(D, D)
        if not dr.any():
            dr = np.eye(D)  # an all-zero direction means identity
        return tuple(int(v) for v in gm.size), sp, org, dr

    (sf, pf, of, rf), (sm, pm, om, rm) = parts(fixed_geometry), parts(moving_geometry)
    same = sf == sm and np.array_equal(pf, pm) and np.array_equal(of, om) and np.array_equal(rf, rm)
    if not same:
        m = apply_transforms(fixed_geometry, m, moving_geometry, [], "Linear", 0.0)
    D = fixed_geometry.dim
    d_f, d_m = L.DeviceArray.from_host(f), L.DeviceArray.from_host(m)
    g1, g2 = L.DeviceArray(f.shape + (D,)), L.DeviceArray(f.shape + (D,))
    value = C.c_double()
    L.check(L.lib().ants_op_cc(C.byref(fixed_geometry.desc()), d_f.ptr, d_m.ptr, int(radius), g1.ptr, g2.ptr,
                               C.byref(value), None))
    return float(value.value)


def smooth_displacement_field(field, geometry, variance=None, mesh=None, order: int = 3) -> np.ndarray:
    """SmoothDisplacementField (Examples/SmoothDisplacementField.cxx:46-180) on
    the GPU: with ``variance``, per-axis GaussianOperator smoothing (maximum
    error 0.001), the boundary zeroed and a blend with the input below
    variance 0.5 (ants_op_smooth_field); with ``mesh`` (spans per axis), the
    one-level DisplacementFieldToBSplineImageFilter fit of the given order
    with a stationary boundary, evaluated on the field's grid
    (ants_op_bspline_field)."""
    import ctypes as C
    from . import _lib as L
    if (variance is None) == (mesh is None):
        raise ValueError("give either a variance or a mesh")
    D = geometry.dim
    f = np.asarray(field, np.float32)
    if f.shape != tuple(geometry.shape) + (D,):
        raise ValueError("field shape does not match its geometry")
    d_in, d_out = L.DeviceArray.from_host(f), L.DeviceArray(f.shape)
    desc = geometry.desc()
    if variance is not None:
        L.check(L.lib().ants_op_smooth_field(C.byref(desc), d_in.ptr, d_out.ptr, float(variance), None))
    else:
        if len(list(mesh)) != D:
            raise ValueError(f"mesh needs {D} entries")
        m = (C.c_int * 3)(*[int(v) for v in mesh], *([1] * (3 - D)))
        mx = C.c_double()
        L.check(L.lib().ants_op_bspline_field(C.byref(desc), d_in.ptr, d_out.ptr, m, int(order), C.byref(mx), None))
    return d_out.to_host()
