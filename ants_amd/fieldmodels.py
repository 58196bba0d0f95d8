"""Host-side mirror of antsRegistration's dense field-transform stages over the
gfx950 library: ``-t GaussianDisplacementField[step,updateVariance,totalVariance]``,
``-t Exponential[step,updateVariance,velocityVariance,integrationSteps]``,
``-t BSplineDisplacementField[step,updateMesh,totalMesh,order]`` and
``-t BSplineExponential[step,updateMesh,velocityMesh,integrationSteps,order]``
with ``-m CC[fixed,moving,1,radius]``
(Examples/antsRegistrationTemplateHeader.h:692-743, 871-935;
Examples/itkantsRegistrationHelper.hxx:1733-1982, 3112-3461).

``FieldStage.from_args`` reads the stage's options, ``register_field`` runs it
through ``ants_field_register`` (include/ants_syn.h) and returns the
displacement field (for Exponential also its inverse and the velocity field)
on the last level's grid, with the optimizer's per-iteration records.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .apply_transforms import Geometry, transform_array
from .linear import LinearRecordC
from .registration import CCMetric, Convergence, SyNTransform, parse_option_function, parse_sigmas, parse_vector

MODELS = {"gaussiandisplacementfield": 0, "gdf": 0, "exponential": 1, "exp": 1, "bsplinedisplacementfield": 2,
          "bsplineexponential": 3}
GAUSSIAN_DISPLACEMENT, EXPONENTIAL, BSPLINE_DISPLACEMENT, BSPLINE_EXPONENTIAL = 0, 1, 2, 3


class FieldParamsC(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("model", C.c_int),
        ("n_levels", C.c_int),
        ("shrink", C.c_int * L.MAX_LEVELS),
        ("sigma", C.c_double * L.MAX_LEVELS),
        ("sigma_in_mm", C.c_int),
        ("iterations", C.c_int * L.MAX_LEVELS),
        ("convergence_threshold", C.c_double),
        ("convergence_window", C.c_int),
        ("learning_rate", C.c_double),
        ("update_variance", C.c_double),
        ("total_variance", C.c_double),
        ("integration_steps", C.c_int),
        ("cc_radius", C.c_int),
        ("preprocess", C.c_int),
        ("force_iterations", C.c_int),
        ("update_mesh", C.c_int * 3),
        ("total_mesh", C.c_int * 3),
        ("spline_order", C.c_int),
    ]


def _bind():
    lib = L.lib()
    if getattr(lib, "_field_bound", False):
        return lib
    vp, i = C.c_void_p, C.c_int
    pdesc = C.POINTER(L.ImageDesc)
    px = C.POINTER(L.TransformC)
    lib.ants_field_register.argtypes = [C.POINTER(FieldParamsC), pdesc, vp, pdesc, vp, px, i, px, i, vp, vp, vp,
                                        C.POINTER(LinearRecordC), C.c_int64, C.POINTER(C.c_int64),
                                        C.POINTER(C.c_int), vp]
    lib.ants_field_metric.argtypes = [C.POINTER(FieldParamsC), i, pdesc, vp, pdesc, vp, px, i, px, i, vp,
                                      C.POINTER(C.c_double), vp, vp]
    lib.ants_op_exponential.argtypes = [pdesc, vp, vp, i, i, C.POINTER(C.c_int), vp]
    lib._field_bound = True
    return lib


@dataclass
class FieldStage:
    """One GaussianDisplacementField or Exponential stage with the CC metric."""
    model: int = GAUSSIAN_DISPLACEMENT
    gradient_step: float = 0.25
    update_variance: float = 3.0
    total_variance: float = 0.0          # Exponential: the velocity field's variance
    integration_steps: int = 0           # Exponential: 0 = automatic
    radius: int = 4
    iterations: Sequence[int] = (20,)
    threshold: float = 1e-6
    window: int = 10
    shrink: Sequence[int] = (1,)
    sigma: Sequence[float] = (0.0,)
    sigma_in_mm: bool = False
    preprocess: bool = True
    force_iterations: bool = False
    fixed: str = ""
    moving: str = ""
    update_mesh: Sequence[int] = (0, 0, 0)   # B-spline forms: spans per axis at the base level
    total_mesh: Sequence[int] = (0, 0, 0)    # B-spline forms: total / velocity field (0: no fit)
    spline_order: int = 3

    @property
    def expo(self) -> bool:
        return self.model in (EXPONENTIAL, BSPLINE_EXPONENTIAL)

    @staticmethod
    def is_field(transform_arg: str) -> bool:
        return parse_option_function(transform_arg)[0].strip().lower() in MODELS

    @classmethod
    def from_args(cls, transform: str, metric: str, convergence: str, shrink: str, sigmas: str,
                  fixed_size=None, fixed_spacing=None) -> "FieldStage":
        """fixed_size / fixed_spacing (x, y[, z]): the fixed image's geometry,
        needed only for a B-spline mesh given as a knot spacing."""
        tname, tp = parse_option_function(transform)
        key = tname.strip().lower()
        if key not in MODELS:
            raise ValueError("field stages are GaussianDisplacementField, Exponential, BSplineDisplacementField "
                             f"or BSplineExponential, got {tname!r}")
        model = MODELS[key]
        um, tm, order = (0, 0, 0), (0, 0, 0), 3
        if model in (BSPLINE_DISPLACEMENT, BSPLINE_EXPONENTIAL):
            # Examples/antsRegistrationTemplateHeader.h:701-743 (order is
            # parameter 3) and 888-935 (integration steps 3, order 4)
            if len(tp) < 2:
                raise ValueError(f"{tname} needs [step, updateMesh, ...]")
            order = int(tp[3 if model == BSPLINE_DISPLACEMENT else 4]) if len(tp) > (
                3 if model == BSPLINE_DISPLACEMENT else 4) else 3
            fake = "BSplineSyN[" + ",".join([tp[0], tp[1], tp[2] if len(tp) > 2 else "0", str(order)]) + "]"
            t = SyNTransform.parse(fake, fixed_size, fixed_spacing)
            um, tm = tuple(t.update_mesh), tuple(t.total_mesh)
            steps = int(tp[3]) if model == BSPLINE_EXPONENTIAL and len(tp) > 3 else 0
            uvar = tvar = 0.0
        else:
            if len(tp) < 3:
                raise ValueError(f"{tname} needs [step, updateVariance, "
                                 f"{'totalVariance' if model == 0 else 'velocityVariance'}]")
            steps = int(tp[3]) if model == EXPONENTIAL and len(tp) > 3 else 0
            uvar, tvar = float(tp[1]), float(tp[2])
        m = CCMetric.parse(metric)
        if m.weight != 1.0:
            raise ValueError("a single CC metric with weight 1 is supported")
        conv = Convergence.parse(convergence)
        f = parse_vector(shrink)
        s, in_mm = parse_sigmas(sigmas, len(conv.iterations))
        if not (len(conv.iterations) == len(f) == len(s)):
            raise ValueError("-c, -f and -s must give the same number of levels")
        return cls(model, float(tp[0]), uvar, tvar, steps, m.radius, conv.iterations, conv.threshold, conv.window, f,
                   s, in_mm, True, False, m.fixed, m.moving, um, tm, order)

    def params(self, dim: int) -> FieldParamsC:
        p = FieldParamsC()
        p.dim = dim
        p.model = int(self.model)
        n = len(self.iterations)
        if not 1 <= n <= L.MAX_LEVELS:
            raise ValueError("1..8 levels")
        p.n_levels = n
        for k in range(n):
            p.shrink[k] = int(self.shrink[k])
            p.sigma[k] = float(self.sigma[k])
            p.iterations[k] = int(self.iterations[k])
        p.sigma_in_mm = int(self.sigma_in_mm)
        p.convergence_threshold = self.threshold
        p.convergence_window = self.window
        p.learning_rate = self.gradient_step
        p.update_variance = self.update_variance
        p.total_variance = self.total_variance
        p.integration_steps = int(self.integration_steps)
        p.cc_radius = int(self.radius)
        p.preprocess = int(self.preprocess)
        p.force_iterations = int(self.force_iterations)
        for d in range(3):
            p.update_mesh[d] = int(self.update_mesh[d]) if d < len(self.update_mesh) else 0
            p.total_mesh[d] = int(self.total_mesh[d]) if d < len(self.total_mesh) else 0
        p.spline_order = int(self.spline_order)
        return p


@dataclass
class FieldResult:
    field: np.ndarray                        # displacement on the last level's grid, shape + (D,)
    inverse: Optional[np.ndarray] = None     # Exponential: exp(-v)
    velocity: Optional[np.ndarray] = None    # Exponential: v
    iterations: List[int] = field(default_factory=list)
    records: List[tuple] = field(default_factory=list)   # (level, iteration, value, convergence, learning rate)

    @property
    def energy(self) -> np.ndarray:
        return np.array([r[2] for r in self.records])


def _device_image(img):
    return img if isinstance(img, L.DeviceArray) else L.DeviceArray.from_host(np.asarray(img, np.float32))


def level_geometry(stage: FieldStage, fgeom: Geometry, level: int) -> Geometry:
    """A level's grid (the shrink rule of ants_syn_level_geometry, the
    helper's for every stage type): the grid of that level's fields."""
    D = fgeom.dim
    size = tuple(reversed(level_shape(stage, fgeom, level)))
    sp = L.SynParamsC()
    sp.dim = D
    sp.n_levels = len(stage.shrink)
    for k, f in enumerate(stage.shrink):
        sp.shrink[k] = int(f)
        sp.iterations[k] = 1
    sp.convergence_threshold, sp.convergence_window = 1e-6, 10
    sp.learning_rate, sp.update_variance, sp.cc_radius = 0.1, 3.0, int(stage.radius)
    out = L.ImageDesc()
    L.check(L.lib().ants_syn_level_geometry(C.byref(sp), C.byref(fgeom.desc()), level, C.byref(out)))
    return Geometry(size, tuple(out.spacing[k] for k in range(D)), tuple(out.origin[k] for k in range(D)),
                    np.array([out.direction[k] for k in range(D * D)]).reshape(D, D))


def level_shape(stage: FieldStage, fgeom: Geometry, level: int):
    """numpy shape (z, y, x) of a level's grid (ants_syn_level_geometry: the
    helper's shrink rule, the same for every stage type)."""
    D = fgeom.dim
    sp = L.SynParamsC()
    sp.dim = D
    sp.n_levels = len(stage.shrink)
    for k, f in enumerate(stage.shrink):
        sp.shrink[k] = int(f)
        sp.iterations[k] = 1
    sp.convergence_threshold, sp.convergence_window = 1e-6, 10
    sp.learning_rate, sp.update_variance, sp.cc_radius = 0.1, 3.0, int(stage.radius)
    out = L.ImageDesc()
    L.check(L.lib().ants_syn_level_geometry(C.byref(sp), C.byref(fgeom.desc()), level, C.byref(out)))
    return tuple(int(out.size[d]) for d in reversed(range(D)))


def register_field(stage: FieldStage, fixed, fgeom: Geometry, moving, mgeom: Geometry, moving_initial=(),
                   fixed_initial=()) -> FieldResult:
    """Run one field-transform stage; the images may be numpy arrays or DeviceArrays."""
    lib = _bind()
    D = fgeom.dim
    p = stage.params(D)
    dF, dM = _device_image(fixed), _device_image(moving)
    xm, keep_m = transform_array(list(moving_initial), D)
    xf, keep_f = transform_array(list(fixed_initial), D)
    shape = level_shape(stage, fgeom, len(stage.shrink) - 1) + (D,)
    out = L.DeviceArray(shape)
    inv = L.DeviceArray(shape) if stage.expo else None
    vel = L.DeviceArray(shape) if stage.expo else None
    cap = int(sum(int(v) for v in stage.iterations)) + 1
    tr = (LinearRecordC * cap)()
    n = C.c_int64()
    its = (C.c_int * L.MAX_LEVELS)()
    try:
        L.check(lib.ants_field_register(C.byref(p), C.byref(fgeom.desc()), dF.ptr, C.byref(mgeom.desc()), dM.ptr,
                                        xm, len(moving_initial), xf, len(fixed_initial), out.ptr,
                                        inv.ptr if inv else None, vel.ptr if vel else None, tr, cap, C.byref(n),
                                        its, None))
        res = FieldResult(out.to_host(), inv.to_host() if inv else None, vel.to_host() if vel else None,
                          [int(its[k]) for k in range(p.n_levels)],
                          [(tr[k].level, tr[k].iteration, tr[k].value, tr[k].convergence, tr[k].learning_rate)
                           for k in range(min(n.value, cap))])
    finally:
        for b in (out, inv, vel):
            if b is not None:
                b.free()
    return res


def field_metric(stage: FieldStage, level: int, fixed, fgeom: Geometry, moving, mgeom: Geometry, field=None,
                 moving_initial=(), fixed_initial=(), derivative: bool = True):
    """(value, derivative field or None) of one metric evaluation on a level
    at displacement `field` (level grid, shape + (D,); None: zero)."""
    lib = _bind()
    D = fgeom.dim
    p = stage.params(D)
    dF, dM = _device_image(fixed), _device_image(moving)
    xm, keep_m = transform_array(list(moving_initial), D)
    xf, keep_f = transform_array(list(fixed_initial), D)
    shape = level_shape(stage, fgeom, level) + (D,)
    du = L.DeviceArray.from_host(np.asarray(field, np.float32).reshape(shape)) if field is not None else None
    dd = L.DeviceArray(shape) if derivative else None
    v = C.c_double()
    try:
        L.check(lib.ants_field_metric(C.byref(p), level, C.byref(fgeom.desc()), dF.ptr, C.byref(mgeom.desc()), dM.ptr,
                                      xm, len(moving_initial), xf, len(fixed_initial), du.ptr if du else None,
                                      C.byref(v), dd.ptr if dd else None, None))
        return v.value, (dd.to_host() if dd else None)
    finally:
        for b in (du, dd):
            if b is not None:
                b.free()


def exponential(geom: Geometry, v, steps: int = 0, inverse: bool = False):
    """(exp(v) or exp(-v), squarings used) on the GPU (ants_op_exponential)."""
    lib = _bind()
    dv = L.DeviceArray.from_host(np.asarray(v, np.float32))
    out = L.DeviceArray(dv.shape)
    used = C.c_int()
    try:
        L.check(lib.ants_op_exponential(C.byref(geom.desc()), dv.ptr, out.ptr, int(steps), -1 if inverse else 1,
                                        C.byref(used), None))
        return out.to_host(), used.value
    finally:
        dv.free()
        out.free()
