"""Host-side mirror of antsRegistration's linear stages over the gfx950 library.

``LinearStage.from_args`` reads the options antsRegistration takes for one
stage, `-t Rigid[0.1] -m Mattes[fixed,moving,1,32,Regular,0.25]
-c [1000x500x250x100,1e-6,10] -f 8x4x2x1 -s 3x2x1x0vox`
(Examples/antsRegistrationTemplateHeader.h:542-815 transforms and schedules,
985-1086 metrics), and ``register_linear`` runs it through
``ants_linear_register`` (include/ants_syn.h), returning the stage transform
as an ``AffineTransform`` and the optimizer's per-iteration records (the
2DIAGNOSTIC lines of Examples/antsRegistrationOptimizerCommandIterationUpdate.h:158-171).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import _lib as L
from .apply_transforms import AffineTransform, Geometry, transform_array
from .registration import Convergence, parse_option_function, parse_sigmas, parse_vector

TRANSFORMS = {"translation": 0, "rigid": 1, "affine": 2}
METRICS = {"meansquares": 0, "gc": 1, "mattes": 2}
# metric names antsRegistration accepts for these metrics
# (StringToMetricType, Examples/itkantsRegistrationHelper.hxx:139-190): "mi"
# is Mattes there; "mi2" is the joint-histogram metric, not provided here
METRIC_ALIASES = {"meansquares": "MeanSquares", "msq": "MeanSquares", "ssd": "MeanSquares", "gc": "GC",
                  "mattes": "Mattes", "mi": "Mattes"}
SAMPLING = {"none": 0, "": 0, "regular": 1, "random": 2}


class LinearParamsC(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("transform", C.c_int),
        ("gradient_step", C.c_double),
        ("metric", C.c_int),
        ("bins", C.c_int),
        ("sampling", C.c_int),
        ("sampling_percentage", C.c_double),
        ("seed", C.c_uint32),
        ("n_levels", C.c_int),
        ("shrink", C.c_int * L.MAX_LEVELS),
        ("sigma", C.c_double * L.MAX_LEVELS),
        ("sigma_in_mm", C.c_int),
        ("iterations", C.c_int * L.MAX_LEVELS),
        ("convergence_threshold", C.c_double),
        ("convergence_window", C.c_int),
        ("preprocess", C.c_int),
    ]


class LinearRecordC(C.Structure):
    _fields_ = [("level", C.c_int), ("iteration", C.c_int), ("value", C.c_double), ("convergence", C.c_double),
                ("learning_rate", C.c_double)]


class LinearResultC(C.Structure):
    _fields_ = [("n_params", C.c_int), ("params", C.c_double * 12), ("transform", L.TransformC),
                ("iterations", C.c_int * L.MAX_LEVELS), ("metric_evaluations", C.c_int64)]


def _bind():
    lib = L.lib()
    if getattr(lib, "_linear_bound", False):
        return lib
    vp, i, d = C.c_void_p, C.c_int, C.c_double
    pdesc = C.POINTER(L.ImageDesc)
    px = C.POINTER(L.TransformC)
    pd = C.POINTER(C.c_double)
    lib.ants_linear_register.argtypes = [C.POINTER(LinearParamsC), pdesc, vp, pdesc, vp, px, i, px, i, pd, pd,
                                         C.POINTER(LinearResultC), C.POINTER(LinearRecordC), C.c_int64,
                                         C.POINTER(C.c_int64), vp]
    lib.ants_linear_metric.argtypes = [C.POINTER(LinearParamsC), i, pdesc, vp, pdesc, vp, px, i, px, i, pd, pd, i,
                                       pd, pd, C.POINTER(C.c_int64), vp]
    lib.ants_linear_sample_points.argtypes = [C.POINTER(LinearParamsC), i, pdesc, pd, C.c_int64,
                                              C.POINTER(C.c_int64)]
    lib.ants_linear_estimate_scales.argtypes = [C.POINTER(LinearParamsC), i, pdesc, px, i, pd, pd, pd]
    lib.ants_linear_release_points.argtypes = []
    lib._linear_bound = True
    return lib


@dataclass
class LinearStage:
    """One linear stage: transform, metric, sampling and schedule."""
    transform: str = "Rigid"
    gradient_step: float = 0.1
    metric: str = "MeanSquares"
    bins: int = 32
    sampling: str = "None"
    percentage: float = 1.0
    iterations: Sequence[int] = (20,)
    threshold: float = 1e-6
    window: int = 10
    shrink: Sequence[int] = (1,)
    sigma: Sequence[float] = (0.0,)
    sigma_in_mm: bool = False
    seed: int = 0
    preprocess: bool = True
    fixed: str = ""
    moving: str = ""

    @staticmethod
    def is_linear(transform_arg: str) -> bool:
        return parse_option_function(transform_arg)[0].strip().lower() in TRANSFORMS

    @classmethod
    def from_args(cls, transform: str, metric: str, convergence: str, shrink: str, sigmas: str,
                  seed: int = 0) -> "LinearStage":
        tname, tp = parse_option_function(transform)
        if tname.strip().lower() not in TRANSFORMS:
            raise ValueError(f"linear stages are Translation, Rigid or Affine, got {tname!r}")
        if not tp:
            raise ValueError(f"{tname} needs a gradient step: {tname}[step]")
        mname, mp = parse_option_function(metric)
        key = mname.strip().lower()
        if key == "mi2":
            raise ValueError("the joint-histogram MI metric (MI2) is not provided; MI / Mattes is the "
                             "mutual-information metric of the linear stages here")
        if key not in METRIC_ALIASES:
            raise ValueError(f"linear-stage metrics are MeanSquares (MSQ, SSD), GC and Mattes (MI), got {mname!r}")
        canon = METRIC_ALIASES[key]
        key = canon.lower()
        if len(mp) < 2:
            raise ValueError(f"{mname} needs fixed and moving images")
        bins = int(mp[3]) if key == "mattes" and len(mp) > 3 else 32
        samp = mp[4].strip().lower() if len(mp) > 4 else "none"
        if samp not in SAMPLING:
            raise ValueError(f"invalid sampling strategy {mp[4]!r}")
        pct = float(mp[5]) if len(mp) > 5 else 1.0
        conv = Convergence.parse(convergence)
        f = parse_vector(shrink)
        s, in_mm = parse_sigmas(sigmas, len(conv.iterations))
        if not (len(conv.iterations) == len(f) == len(s)):
            raise ValueError("-c, -f and -s must give the same number of levels")
        return cls(tname.strip(), float(tp[0]), canon, bins, samp, pct, conv.iterations, conv.threshold,
                   conv.window, f, s, in_mm, int(seed), True, mp[0], mp[1])

    def params(self, dim: int) -> LinearParamsC:
        p = LinearParamsC()
        p.dim = dim
        p.transform = TRANSFORMS[self.transform.lower()]
        p.gradient_step = self.gradient_step
        p.metric = METRICS[self.metric.lower()]
        p.bins = self.bins
        p.sampling = SAMPLING[self.sampling.lower()]
        p.sampling_percentage = self.percentage
        p.seed = int(self.seed) & 0xFFFFFFFF
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
        p.preprocess = int(self.preprocess)
        return p


@dataclass
class LinearResult:
    transform: AffineTransform
    params: np.ndarray
    iterations: List[int]
    records: List[tuple] = field(default_factory=list)   # (level, iteration, value, convergence, learning rate)
    evaluations: int = 0

    def diagnostic_lines(self) -> List[str]:
        out = []
        for (lev, it, v, cv, lr) in self.records:
            if it == 1:
                out.append(f"METADATA,Level={lev}")
                out.append("DIAGNOSTIC,Iteration,metricValue,convergenceValue,ITERATION_TIME_INDEX,SINCE_LAST")
            out.append(f" 2DIAGNOSTIC, {it:5d}, {v:.12e}, {cv:.12e}, ")
        return out


def _vec(v, n):
    a = (C.c_double * n)()
    for k in range(n):
        a[k] = float(v[k])
    return a


def _device_image(img):
    return img if isinstance(img, L.DeviceArray) else L.DeviceArray.from_host(np.asarray(img, np.float32))


def register_linear(stage: LinearStage, fixed, fgeom: Geometry, moving, mgeom: Geometry, moving_initial=(),
                    fixed_initial=(), center=None, initial_params=None) -> LinearResult:
    """Run one linear stage; the images may be numpy arrays or DeviceArrays."""
    lib = _bind()
    D = fgeom.dim
    p = stage.params(D)
    dF, dM = _device_image(fixed), _device_image(moving)
    xm, keep_m = transform_array(list(moving_initial), D)
    xf, keep_f = transform_array(list(fixed_initial), D)
    cen = _vec(center, D) if center is not None else None
    ip = _vec(initial_params, len(initial_params)) if initial_params is not None else None
    res = LinearResultC()
    cap = int(sum(int(v) for v in stage.iterations)) + 1
    tr = (LinearRecordC * cap)()
    n = C.c_int64()
    L.check(lib.ants_linear_register(C.byref(p), C.byref(fgeom.desc()), dF.ptr, C.byref(mgeom.desc()), dM.ptr, xm,
                                     len(moving_initial), xf, len(fixed_initial), cen, ip, C.byref(res), tr, cap,
                                     C.byref(n), None))
    T = res.transform
    A = np.array([T.matrix[k] for k in range(D * D)]).reshape(D, D)
    t = np.array([T.translation[k] for k in range(D)])
    c = np.array([T.center[k] for k in range(D)])
    recs = [(tr[k].level, tr[k].iteration, tr[k].value, tr[k].convergence, tr[k].learning_rate)
            for k in range(min(n.value, cap))]
    return LinearResult(AffineTransform(A, t, c), np.array([res.params[k] for k in range(res.n_params)]),
                        [int(res.iterations[k]) for k in range(p.n_levels)], recs, int(res.metric_evaluations))


def release_points() -> None:
    """Free the sample-point sets kept on the device between stages."""
    L.check(_bind().ants_linear_release_points())


def metric_value(stage: LinearStage, level: int, fixed, fgeom: Geometry, moving, mgeom: Geometry, params,
                 moving_initial=(), fixed_initial=(), center=None, derivative: bool = True):
    """(value, -dvalue/dparams or None, valid points) of one metric evaluation."""
    lib = _bind()
    D = fgeom.dim
    p = stage.params(D)
    dF, dM = _device_image(fixed), _device_image(moving)
    xm, keep_m = transform_array(list(moving_initial), D)
    xf, keep_f = transform_array(list(fixed_initial), D)
    cen = _vec(center, D) if center is not None else None
    prm = _vec(params, len(params))
    v = C.c_double()
    der = (C.c_double * 12)()
    n = C.c_int64()
    L.check(lib.ants_linear_metric(C.byref(p), level, C.byref(fgeom.desc()), dF.ptr, C.byref(mgeom.desc()), dM.ptr,
                                   xm, len(moving_initial), xf, len(fixed_initial), cen, prm, int(derivative),
                                   C.byref(v), der, C.byref(n), None))
    return v.value, (np.array([der[k] for k in range(len(params))]) if derivative else None), n.value


def sample_points(stage: LinearStage, level: int, fgeom: Geometry) -> np.ndarray:
    lib = _bind()
    p = stage.params(fgeom.dim)
    n = C.c_int64()
    L.check(lib.ants_linear_sample_points(C.byref(p), level, C.byref(fgeom.desc()), None, 0, C.byref(n)))
    out = np.zeros(max(n.value, 1) * fgeom.dim)
    L.check(lib.ants_linear_sample_points(C.byref(p), level, C.byref(fgeom.desc()),
                                          out.ctypes.data_as(C.POINTER(C.c_double)), n.value, C.byref(n)))
    return out[: n.value * fgeom.dim].reshape(n.value, fgeom.dim)


def estimate_scales(stage: LinearStage, level: int, fgeom: Geometry, params, moving_initial=(), center=None):
    lib = _bind()
    D = fgeom.dim
    p = stage.params(D)
    xm, keep = transform_array(list(moving_initial), D)
    cen = _vec(center, D) if center is not None else None
    out = (C.c_double * 12)()
    L.check(lib.ants_linear_estimate_scales(C.byref(p), level, C.byref(fgeom.desc()), xm, len(moving_initial), cen,
                                            _vec(params, len(params)), out))
    return np.array([out[k] for k in range(len(params))])
