"""antsApplyTransforms subset on the GPU: resample an image through a
composite of affine and displacement-field transforms
(``Examples/antsApplyTransforms.cxx:546-640``).

``apply_transforms(reference, image, image_geometry, transforms, ...)``
mirrors ``antsApplyTransforms -r reference -i image -t T1 -t T2 ... -n I
--default-value v``: the transforms are listed as on the command line, and
each reference point is mapped through T1 first, then T2, ...  (the
command-line parser stores repeated options newest-first and ITK's
CompositeTransform applies the last added transform first,
``Examples/itkantsRegistrationHelper.h:1262-1455``).  Interpolators: Linear,
NearestNeighbor, GenericLabel[Linear].
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple, Union

import numpy as np

from . import _lib as L

INTERPOLATORS = {"linear": 0, "nearestneighbor": 1, "genericlabel": 2,
                 # WindowedSincInterpolateImageFunction, radius 3, zero boundary
                 "cosinewindowedsinc": 32, "hammingwindowedsinc": 33, "lanczoswindowedsinc": 34,
                 "blackmanwindowedsinc": 35, "welchwindowedsinc": 36}


@dataclass
class Geometry:
    """Image grid: size (x, y[, z]), spacing, origin, direction (row-major)."""
    size: Tuple[int, ...]
    spacing: Optional[Sequence[float]] = None
    origin: Optional[Sequence[float]] = None
    direction: Optional[np.ndarray] = None

    @property
    def dim(self) -> int:
        return len(self.size)

    @property
    def shape(self) -> Tuple[int, ...]:
        """numpy (z, y, x) order"""
        return tuple(int(v) for v in reversed(self.size))

    def desc(self) -> L.ImageDesc:
        return L.desc(self.size, self.spacing, self.origin, self.direction)


@dataclass
class AffineTransform:
    """ITK MatrixOffsetTransformBase: y = A (x - c) + t + c.  ``invert``
    selects the inverse transform (``-t [affine.mat,1]``)."""
    matrix: np.ndarray
    translation: Sequence[float]
    center: Optional[Sequence[float]] = None
    invert: bool = False

    def parameters(self):
        A = np.asarray(self.matrix, dtype=np.float64)
        D = A.shape[0]
        t = np.asarray(self.translation, dtype=np.float64)
        c = np.zeros(D) if self.center is None else np.asarray(self.center, dtype=np.float64)
        if not self.invert:
            return A, t, c
        # same centre: A' = A^-1, offset' = -A^-1 offset, t' = offset' - c + A' c
        Ai = np.linalg.inv(A)
        off = t + c - A @ c
        offi = -Ai @ off
        return Ai, offi - c + Ai @ c, c


@dataclass
class DisplacementFieldTransform:
    """ITK DisplacementFieldTransform: y = x + u(x), u linear in the field's
    grid, zero outside it.  ``field`` is (z, y, x, D) float32 in physical
    units, on the host or an ``ants_amd._lib.DeviceArray``."""
    field: Union[np.ndarray, L.DeviceArray]
    geometry: Geometry


Transform = Union[AffineTransform, DisplacementFieldTransform, None]


BSPLINE_BASE = 16          # ANTS_INTERP_BSPLINE_ORDER(q) = 16 + q
BSPLINE_DEFAULT_ORDER = 3  # BSplineInterpolateImageFunction's default


def parse_interpolation(name: str) -> int:
    """-n Linear | NearestNeighbor | GenericLabel[Linear] | BSpline[order] |
    {Cosine,Hamming,Lanczos,Blackman,Welch}WindowedSinc
    (Examples/make_interpolator_snip.tmpl:18-110; BSpline's order is 0..5,
    3 when omitted)."""
    m = re.fullmatch(r"\s*([A-Za-z]+)\s*(?:\[\s*([A-Za-z0-9]*)\s*\])?\s*", name)
    if not m:
        raise ValueError(f"bad interpolation {name!r}")
    base = m.group(1).lower()
    if base == "bspline":
        arg = m.group(2)
        if arg and not arg.isdigit():
            raise ValueError(f"bad B-spline order in {name!r}")
        order = int(arg) if arg else BSPLINE_DEFAULT_ORDER
        if not 0 <= order <= 5:
            raise ValueError(f"B-spline order must be 0..5, got {order}")
        return BSPLINE_BASE + order
    if base == "genericlabel" and m.group(2) and m.group(2).lower() != "linear":
        raise ValueError("GenericLabel is supported with its default Linear interpolator only")
    if base not in INTERPOLATORS:
        raise ValueError(f"interpolation {name!r} is not supported (Linear, NearestNeighbor, GenericLabel, BSpline, "
                         "*WindowedSinc)")
    return INTERPOLATORS[base]


def apply_transforms(reference: Geometry, image, image_geometry: Geometry, transforms: Sequence[Transform] = (),
                     interpolation: str = "Linear", default_value: float = 0.0, out=None):
    """Resample ``image`` onto ``reference``.  ``image`` and ``out`` may be
    numpy arrays (host) or DeviceArrays; returns a numpy array when ``out``
    is None, else ``out``."""
    if reference.dim != image_geometry.dim:
        raise ValueError("reference and image dimensions differ")
    D = reference.dim
    if len(transforms) > 8:
        raise ValueError("at most 8 transforms")
    d_img = image if isinstance(image, L.DeviceArray) else L.DeviceArray.from_host(np.asarray(image, np.float32))
    xs, keep = transform_array(transforms, D)
    keep.append(d_img)
    d_out = out if out is not None else L.DeviceArray(reference.shape)
    L.check(L.lib().ants_apply_transforms(C.byref(reference.desc()), C.byref(image_geometry.desc()), d_img.ptr, xs,
                                          len(transforms), parse_interpolation(interpolation), float(default_value),
                                          d_out.ptr, None))
    return d_out if out is not None else d_out.to_host()


def compose_affines(transforms: Sequence[Optional[AffineTransform]]) -> AffineTransform:
    """One AffineTransform (centre 0) equal to the list applied in order: a
    point passes through transforms[0] first.  None entries are identities."""
    A = None
    off = None
    for t in transforms:
        if t is None:
            continue
        B, tr, c = t.parameters()
        ob = tr + c - B @ c
        if A is None:
            A, off = B.copy(), ob.copy()
        else:
            A, off = B @ A, B @ off + ob
    if A is None:
        raise ValueError("no affine transform to compose")
    return AffineTransform(A, off, np.zeros(A.shape[0]))


def apply_transforms_series(reference: Geometry, series, image_geometry: Geometry,
                            transforms: Sequence[Transform] = (), interpolation: str = "Linear",
                            default_value: float = 0.0) -> np.ndarray:
    """antsApplyTransforms -e 3 (Examples/antsApplyTransforms.cxx:516-536,
    903-960): every volume of a (t, z, y, x) series resampled onto
    ``reference`` through the same transforms; returns (t,) + reference.shape.
    The transform list is set up once; the volumes go through the device one
    after another in one output buffer."""
    a = np.asarray(series)
    if a.ndim != image_geometry.dim + 1 or a.shape[1:] != image_geometry.shape:
        raise ValueError(f"series shape {a.shape} does not match (t,) + {image_geometry.shape}")
    D = reference.dim
    xs, keep = transform_array(transforms, D)
    code = parse_interpolation(interpolation)
    d_in = L.DeviceArray(image_geometry.shape)
    d_out = L.DeviceArray(reference.shape)
    out = np.empty((a.shape[0],) + reference.shape, np.float32)
    rdesc, idesc = reference.desc(), image_geometry.desc()
    for k in range(a.shape[0]):
        d_in.copy_from_host(np.ascontiguousarray(a[k], np.float32))
        L.check(L.lib().ants_apply_transforms(C.byref(rdesc), C.byref(idesc), d_in.ptr, xs, len(transforms), code,
                                              float(default_value), d_out.ptr, None))
        out[k] = d_out.to_host()
    return out


def compose_transforms_to_field(reference: Geometry, transforms: Sequence[Transform] = (), out=None):
    """antsApplyTransforms -o [field,1] (Examples/antsApplyTransforms.cxx:706-760):
    the displacement T(p) - p of every reference point through the
    transforms in -t order, shape reference.shape + (dim,), physical
    components.  Returns a numpy array when ``out`` is None, else ``out``."""
    D = reference.dim
    xs, keep = transform_array(transforms, D)
    d_out = out if out is not None else L.DeviceArray(reference.shape + (D,))
    L.check(L.lib().ants_compose_transforms_to_field(C.byref(reference.desc()), xs, len(transforms), d_out.ptr, None))
    return d_out if out is not None else d_out.to_host()


def apply_transforms_to_points(points, transforms: Sequence[Transform] = (), dim: int = 3) -> np.ndarray:
    """antsApplyTransformsToPoints (Examples/antsApplyTransformsToPoints.cxx:145-190):
    the first ``dim`` columns of ``points`` (n rows, physical coordinates)
    through the transforms in -t order; further columns are copied.  Returns
    a float64 array of the same shape."""
    pts = np.asarray(points, dtype=np.float64)
    if pts.ndim != 2 or pts.shape[1] < dim:
        raise ValueError(f"points must be an (n, >= {dim}) array")
    out = pts.copy()
    n = pts.shape[0]
    if n == 0:
        return out
    xs, keep = transform_array(transforms, dim)
    d_pts = L.DeviceArray.from_host(np.ascontiguousarray(pts[:, :dim]))
    L.check(L.lib().ants_apply_transforms_to_points(dim, d_pts.ptr, n, xs, len(transforms), d_pts.ptr, None))
    out[:, :dim] = d_pts.to_host()
    return out


def collapse_linear(transforms: Sequence[Transform], D: int) -> "AffineTransform":
    """antsApplyTransforms -o Linear[file] (Examples/antsApplyTransforms.cxx:647-688):
    the single affine equal to the -t list of affine (and identity)
    transforms; a displacement field among them is refused, as there."""
    if any(t is not None and not isinstance(t, AffineTransform) for t in transforms):
        raise ValueError("the transform or set of transforms is not linear")
    if all(t is None for t in transforms):
        return AffineTransform(np.eye(D), np.zeros(D), np.zeros(D))
    return compose_affines(transforms)


def transform_array(transforms: Sequence[Transform], D: int):
    """ants_transform array for the C-ABI, plus the device buffers it refers
    to (keep them alive while the array is in use)."""
    if len(transforms) > 8:
        raise ValueError("at most 8 transforms")
    keep = []
    xs = (L.TransformC * max(len(transforms), 1))()
    for i, t in enumerate(transforms):
        if t is None:
            xs[i].kind = 0
        elif isinstance(t, AffineTransform):
            A, tr, c = t.parameters()
            if A.shape != (D, D):
                raise ValueError("affine matrix must be dim x dim")
            xs[i].kind = 1
            for k, v in enumerate(A.ravel()):
                xs[i].matrix[k] = float(v)
            for k in range(D):
                xs[i].translation[k] = float(tr[k])
                xs[i].center[k] = float(c[k])
        elif isinstance(t, DisplacementFieldTransform):
            f = t.field if isinstance(t.field, L.DeviceArray) else L.DeviceArray.from_host(np.asarray(t.field, np.float32))
            if tuple(f.shape) != t.geometry.shape + (D,):
                raise ValueError("displacement field shape does not match its geometry")
            keep.append(f)
            xs[i].kind = 2
            xs[i].field = t.geometry.desc()
            xs[i].d_field = f.ptr
        else:
            raise TypeError(f"unsupported transform {type(t).__name__}")
    return xs, keep
