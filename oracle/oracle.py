"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, the smoke check and bench.py's cpu_baseline leg import this
module.  It loads ``oracle/libsyn_oracle.so`` built from ``syn_oracle.cpp``
(``make -C oracle``).  The FP64 entry points are the parity oracle; the FP32
driver is the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsyn_oracle.so")


class OgGrid(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int64 * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3),
                ("direction", C.c_double * 9)]


class OgParams(C.Structure):
    _fields_ = [
        ("n_levels", C.c_int),
        ("shrink", C.c_int * 8),
        ("sigma", C.c_double * 8),
        ("sigma_in_mm", C.c_int),
        ("iterations", C.c_int * 8),
        ("learning_rate", C.c_double),
        ("update_variance", C.c_double),
        ("total_variance", C.c_double),
        ("radius", C.c_int),
        ("conv_threshold", C.c_double),
        ("conv_window", C.c_int),
        ("preprocess", C.c_int),
        ("force_iterations", C.c_int),
        ("threads", C.c_int),
        ("inv_tie_rule", C.c_int),
        ("model", C.c_int),
        ("update_mesh", C.c_int * 3),
        ("total_mesh", C.c_int * 3),
        ("spline_order", C.c_int),
    ]


class OgTrace(C.Structure):
    _fields_ = [
        ("max_records", C.c_int),
        ("n_records", C.c_int),
        ("level", C.POINTER(C.c_int)),
        ("iteration", C.POINTER(C.c_int)),
        ("energy", C.POINTER(C.c_double)),
        ("convergence", C.POINTER(C.c_double)),
        ("inv_iters", C.POINTER(C.c_int)),
        ("seconds", C.POINTER(C.c_double)),
    ]


FLT_EPSILON = float(np.finfo(np.float32).eps)

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (run `make -C oracle`)")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        d = C.c_double
        pd = P(C.c_double)
        pg = P(OgGrid)
        L.og_gaussian_kernel.argtypes = [d, d, C.c_int64, pd, C.c_int]
        L.og_gaussian_kernel.restype = C.c_int
        L.og_bessel_in.argtypes = [C.c_int, d]
        L.og_bessel_in.restype = d
        L.og_shrink_grid.argtypes = [pg, C.c_int, pg]
        L.og_smooth_image_f64.argtypes = [pg, pd, pd, d, C.c_int]
        L.og_smooth_field_f64.argtypes = [pg, pd, pd, d]
        L.og_warp_image_f64.argtypes = [pg, pd, pg, pd, pd]
        L.og_compose_f64.argtypes = [pg, pd, pd, pd]
        L.og_resample_field_f64.argtypes = [pg, pd, pg, pd]
        L.og_cc_f64.argtypes = [pg, pd, pd, C.c_int, d, pd, pd]
        L.og_cc_f64.restype = d
        L.og_cc_ext_f64.argtypes = [pg, pd, pd, C.c_int, d, pd, pd]
        L.og_cc_ext_f64.restype = d
        L.og_scale_update_f64.argtypes = [pg, pd, d]
        L.og_scale_update_f64.restype = d
        L.og_bspline_fit_field_f64.argtypes = [pg, pd, pd, P(C.c_int), C.c_int]
        L.og_bspline_point.argtypes = [C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, P(C.c_int), pd]
        L.og_invert_f64.argtypes = [pg, pd, pd, pd, pd, pd]
        L.og_invert_f64.restype = C.c_int
        L.og_preprocess_f64.argtypes = [C.c_int64, pd, pd]
        L.og_conv_value.argtypes = [pd, C.c_int, C.c_int]
        L.og_conv_value.restype = d
        L.og_warp_labels_nn.argtypes = [pg, P(C.c_float), pg, P(C.c_int32), P(C.c_int32)]
        L.og_syn_register_f64.argtypes = [P(OgParams), pg, pd, pg, pd, pd, pd, P(OgTrace), pd, d]
        L.og_syn_register_f64.restype = C.c_int
        L.og_syn_register_f32.argtypes = [P(OgParams), pg, P(C.c_float), pg, P(C.c_float), P(C.c_float),
                                          P(C.c_float), P(OgTrace), P(C.c_float), d]
        L.og_syn_register_f32.restype = C.c_int
        L.og_max_threads.restype = C.c_int
        _lib = L
    return _lib


@dataclass
class Grid:
    """Image grid: size (x fastest), spacing, origin, direction (row-major
    dim x dim orthonormal; identity by default)."""
    size: Sequence[int]
    spacing: Sequence[float] = None
    origin: Sequence[float] = None
    direction: Optional[np.ndarray] = None

    def __post_init__(self):
        D = len(self.size)
        if self.spacing is None:
            self.spacing = [1.0] * D
        if self.origin is None:
            self.origin = [0.0] * D
        if self.direction is None:
            self.direction = np.eye(D)

    @property
    def dim(self):
        return len(self.size)

    @property
    def count(self):
        return int(np.prod(self.size))

    @property
    def shape(self):
        """numpy shape (z, y, x) / (y, x) of an x-fastest array."""
        return tuple(int(s) for s in reversed(self.size))

    def c(self) -> OgGrid:
        g = OgGrid()
        g.dim = self.dim
        for d in range(3):
            g.n[d] = int(self.size[d]) if d < self.dim else 1
            g.spacing[d] = float(self.spacing[d]) if d < self.dim else 1.0
            g.origin[d] = float(self.origin[d]) if d < self.dim else 0.0
        for k, v in enumerate(np.asarray(self.direction, dtype=np.float64).reshape(self.dim, self.dim).ravel()):
            g.direction[k] = float(v)
        return g


def _pd(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def gaussian_kernel(variance: float, max_error: float, max_width: int = 32) -> np.ndarray:
    out = np.zeros(128)
    r = lib().og_gaussian_kernel(variance, max_error, max_width, _pd(out), 128)
    return out[: r + 1].copy()


def bessel_in(n: int, x: float) -> float:
    return lib().og_bessel_in(n, x)


def shrink_grid(g: Grid, factor: int) -> Grid:
    o = OgGrid()
    lib().og_shrink_grid(C.byref(g.c()), factor, C.byref(o))
    D = g.dim
    return Grid([int(o.n[d]) for d in range(D)], [o.spacing[d] for d in range(D)], [o.origin[d] for d in range(D)],
                np.array([o.direction[k] for k in range(D * D)]).reshape(D, D))


def smooth_image(g: Grid, img, sigma: float, in_mm: bool = False):
    a = _f64(img)
    out = np.empty_like(a)
    lib().og_smooth_image_f64(C.byref(g.c()), _pd(a), _pd(out), sigma, int(in_mm))
    return out


def smooth_field(g: Grid, fld, variance: float):
    a = _f64(fld)
    out = np.empty_like(a)
    lib().og_smooth_field_f64(C.byref(g.c()), _pd(a), _pd(out), variance)
    return out


def warp_image(gi: Grid, img, go: Grid, u=None, xforms=()):
    """xforms: initial transforms after the displacement (see _xform_array)."""
    a = _f64(img)
    out = np.empty(go.shape, dtype=np.float64)
    keep = _f64(u) if u is not None else None
    up = _pd(keep) if keep is not None else None
    if xforms:
        xs, hold = _xform_array(xforms, go.dim)
        fn = lib().og_warp_image_xf_f64
        fn.argtypes = [C.POINTER(OgGrid), C.POINTER(C.c_double), C.POINTER(OgGrid), C.POINTER(C.c_double),
                       C.POINTER(OgXform), C.c_int, C.POINTER(C.c_double)]
        fn(C.byref(gi.c()), _pd(a), C.byref(go.c()), up, xs, len(xforms), _pd(out))
        return out
    lib().og_warp_image_f64(C.byref(gi.c()), _pd(a), C.byref(go.c()), up, _pd(out))
    return out


def compose(g: Grid, disp, warp):
    d = _f64(disp)
    w = _f64(warp)
    out = np.empty_like(w)
    lib().og_compose_f64(C.byref(g.c()), _pd(d), _pd(w), _pd(out))
    return out


def resample_field(gi: Grid, fld, go: Grid):
    a = _f64(fld)
    out = np.empty(go.shape + (go.dim,), dtype=np.float64)
    lib().og_resample_field_f64(C.byref(gi.c()), _pd(a), C.byref(go.c()), _pd(out))
    return out


def cc(g: Grid, F, M, radius: int, eps: float = FLT_EPSILON):
    """Returns (value, grad_f2m, grad_m2m)."""
    f = _f64(F)
    m = _f64(M)
    g1 = np.empty(g.shape + (g.dim,))
    g2 = np.empty(g.shape + (g.dim,))
    v = lib().og_cc_f64(C.byref(g.c()), _pd(f), _pd(m), radius, eps, _pd(g1), _pd(g2))
    return v, g1, g2


def cc_ext(g: Grid, F, M, radius: int, gext, eps: float = FLT_EPSILON):
    """CC value and c1 * gext per voxel: the v4 metric's derivative with the
    moving gradient taken at each voxel's mapped point (field transforms)."""
    f = _f64(F)
    m = _f64(M)
    ge = _f64(gext)
    out = np.empty(g.shape + (g.dim,))
    v = lib().og_cc_ext_f64(C.byref(g.c()), _pd(f), _pd(m), radius, eps, _pd(ge), _pd(out))
    return v, out


def scale_update(g: Grid, u, lr: float):
    a = _f64(u).copy()
    mx = lib().og_scale_update_f64(C.byref(g.c()), _pd(a), lr)
    return a, mx


def bspline_fit_field(g: Grid, fld, mesh, order: int = 3):
    """BSplineSyN's B-spline fit of a displacement field (mesh: spans per axis)."""
    a = _f64(fld)
    out = np.empty_like(a)
    m = (C.c_int * 3)(*[int(v) for v in list(mesh) + [1] * (3 - len(mesh))])
    lib().og_bspline_fit_field_f64(C.byref(g.c()), _pd(a), _pd(out), m, int(order))
    return out


def bspline_point(i: int, n: int, spans: int, order: int = 3, max_spans: int = None):
    """(span, basis values b_0..b_3) of voxel i of an axis of n voxels."""
    s = C.c_int()
    b = np.zeros(4)
    lib().og_bspline_point(int(i), int(n), int(spans), int(order), int(max_spans or spans), C.byref(s), _pd(b))
    return s.value, b


def invert(g: Grid, fld, init=None):
    """Returns (inverse, inner_iterations, last_max, last_mean)."""
    f = _f64(fld)
    i0 = _f64(init) if init is not None else None
    out = np.empty_like(f)
    mx = C.c_double()
    mn = C.c_double()
    k = lib().og_invert_f64(C.byref(g.c()), _pd(f), _pd(i0) if i0 is not None else None, _pd(out), C.byref(mx), C.byref(mn))
    return out, k, mx.value, mn.value


def preprocess(img):
    a = _f64(img)
    out = np.empty_like(a)
    lib().og_preprocess_f64(a.size, _pd(a), _pd(out))
    return out


def conv_value(energies, window: int) -> float:
    e = _f64(energies)
    return lib().og_conv_value(_pd(e), len(e), window)


def warp_labels_nn(gout: Grid, u, glab: Grid, labels):
    uu = np.ascontiguousarray(u, dtype=np.float32) if u is not None else None
    lab = np.ascontiguousarray(labels, dtype=np.int32)
    out = np.empty(gout.shape, dtype=np.int32)
    lib().og_warp_labels_nn(C.byref(gout.c()), uu.ctypes.data_as(C.POINTER(C.c_float)) if uu is not None else None,
                            C.byref(glab.c()), lab.ctypes.data_as(C.POINTER(C.c_int32)),
                            out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out


@dataclass
class SynParams:
    """SyN[learning_rate, update_variance, total_variance] with CC[., ., 1, radius]."""
    shrink: List[int]
    sigma: List[float]
    iterations: List[int]
    learning_rate: float = 0.1
    update_variance: float = 3.0
    total_variance: float = 0.0
    radius: int = 4
    conv_threshold: float = 1e-6
    conv_window: int = 10
    sigma_in_mm: bool = False
    preprocess: bool = True
    force_iterations: bool = False
    threads: int = 0
    # inversion max test: 0 guarded (max > 0.1 (1 + 1e-5)), 1 strict (ITK's max > 0.1)
    inv_tie_rule: int = 0
    # 0 SyN; 1 BSplineSyN[learning_rate, update_mesh, total_mesh, spline_order]
    model: int = 0
    update_mesh: Sequence[int] = (0, 0, 0)
    total_mesh: Sequence[int] = (0, 0, 0)
    spline_order: int = 3

    def c(self) -> OgParams:
        p = OgParams()
        p.n_levels = len(self.shrink)
        for i, (s, sg, it) in enumerate(zip(self.shrink, self.sigma, self.iterations)):
            p.shrink[i] = int(s)
            p.sigma[i] = float(sg)
            p.iterations[i] = int(it)
        p.sigma_in_mm = int(self.sigma_in_mm)
        p.learning_rate = self.learning_rate
        p.update_variance = self.update_variance
        p.total_variance = self.total_variance
        p.radius = self.radius
        p.conv_threshold = self.conv_threshold
        p.conv_window = self.conv_window
        p.preprocess = int(self.preprocess)
        p.force_iterations = int(self.force_iterations)
        p.threads = self.threads
        p.inv_tie_rule = int(self.inv_tie_rule)
        p.model = int(self.model)
        for d in range(3):
            p.update_mesh[d] = int(self.update_mesh[d]) if d < len(self.update_mesh) else 0
            p.total_mesh[d] = int(self.total_mesh[d]) if d < len(self.total_mesh) else 0
        p.spline_order = int(self.spline_order)
        return p


@dataclass
class SynResult:
    fwd: np.ndarray
    inv: np.ndarray
    level: np.ndarray
    iteration: np.ndarray
    energy: np.ndarray
    convergence: np.ndarray
    inv_iters: np.ndarray
    state: Optional[np.ndarray] = None
    # (n_levels, ...) results of the schedule stopped after each level
    level_fwd: Optional[np.ndarray] = None
    level_inv: Optional[np.ndarray] = None
    seconds: Optional[np.ndarray] = None     # wall time of each iteration record


def syn_register(params: SynParams, gF: Grid, fixed, gM: Grid, moving, precision: str = "f64",
                 eps: float = FLT_EPSILON, want_state: bool = False, max_records: int = 4096,
                 fixed_xforms=(), moving_xforms=(), restore=None, want_levels: bool = False) -> SynResult:
    """fixed_xforms / moving_xforms: initial transforms (see _xform_array),
    in the order a point passes through them.  restore: (Grid, state) with
    state (4, z, y, x, D) = F2M fwd, F2M inv, M2M fwd, M2M inv (FP64 only;
    moving_xforms may be combined with it).  want_levels (FP64, no initial
    transforms or state): also return each level's result, i.e. the output of
    the same schedule with no iterations after that level."""
    L = lib()
    dt = np.float64 if precision == "f64" else np.float32
    ct = C.c_double if precision == "f64" else C.c_float
    f = np.ascontiguousarray(fixed, dtype=dt)
    m = np.ascontiguousarray(moving, dtype=dt)
    last = shrink_grid(gF, params.shrink[-1])
    fwd = np.empty(last.shape + (gF.dim,), dtype=dt)
    inv = np.empty_like(fwd)
    state = np.empty((4,) + fwd.shape, dtype=dt) if want_state else None
    lv = np.zeros(max_records, np.int32)
    it = np.zeros(max_records, np.int32)
    en = np.zeros(max_records, np.float64)
    cv = np.zeros(max_records, np.float64)
    ki = np.zeros(max_records * 4, np.int32)
    secs = np.zeros(max_records, np.float64)
    tr = OgTrace()
    tr.max_records = max_records
    tr.n_records = 0
    tr.level = lv.ctypes.data_as(C.POINTER(C.c_int))
    tr.iteration = it.ctypes.data_as(C.POINTER(C.c_int))
    tr.energy = en.ctypes.data_as(C.POINTER(C.c_double))
    tr.convergence = cv.ctypes.data_as(C.POINTER(C.c_double))
    tr.inv_iters = ki.ctypes.data_as(C.POINTER(C.c_int))
    tr.seconds = secs.ctypes.data_as(C.POINTER(C.c_double))
    p = params.c()
    P = C.POINTER(ct)
    lvf = lvi = None
    if want_levels:
        if precision != "f64" or restore is not None or fixed_xforms or moving_xforms:
            raise ValueError("want_levels: FP64 without initial transforms or state")
        lvf = np.empty((len(params.shrink),) + fwd.shape, dtype=dt)
        lvi = np.empty_like(lvf)
        fn = L.og_syn_register_levels_f64
        fn.argtypes = [C.POINTER(OgParams), C.POINTER(OgGrid), P, C.POINTER(OgGrid), P, P, P, C.POINTER(OgTrace), P, P,
                       C.c_double]
        fn.restype = C.c_int
        rc = fn(C.byref(p), C.byref(gF.c()), f.ctypes.data_as(P), C.byref(gM.c()), m.ctypes.data_as(P),
                fwd.ctypes.data_as(P), inv.ctypes.data_as(P), C.byref(tr), lvf.ctypes.data_as(P),
                lvi.ctypes.data_as(P), eps)
    elif restore is not None:
        if precision != "f64" or fixed_xforms:
            raise ValueError("restore: FP64 and moving initial transforms only")
        gS, st = restore
        st = np.ascontiguousarray(st, dtype=np.float64)
        if st.shape != (4,) + gS.shape + (gS.dim,):
            raise ValueError("state shape does not match its grid")
        mx, hold_m = _xform_array(moving_xforms, gF.dim)
        fn = L.og_syn_register_restore_f64
        fn.argtypes = [C.POINTER(OgParams), C.POINTER(OgGrid), P, C.POINTER(OgGrid), P, C.POINTER(OgGrid), P,
                       C.POINTER(OgXform), C.c_int, P, P, C.POINTER(OgTrace), P, C.c_double]
        fn.restype = C.c_int
        rc = fn(C.byref(p), C.byref(gF.c()), f.ctypes.data_as(P), C.byref(gM.c()), m.ctypes.data_as(P),
                C.byref(gS.c()), st.ctypes.data_as(P), mx, len(moving_xforms), fwd.ctypes.data_as(P),
                inv.ctypes.data_as(P), C.byref(tr), state.ctypes.data_as(P) if state is not None else None, eps)
    elif fixed_xforms or moving_xforms:
        fx, hold_f = _xform_array(fixed_xforms, gF.dim)
        mx, hold_m = _xform_array(moving_xforms, gF.dim)
        fn = L.og_syn_register_init_f64 if precision == "f64" else L.og_syn_register_init_f32
        fn.argtypes = [C.POINTER(OgParams), C.POINTER(OgGrid), P, C.POINTER(OgGrid), P, C.POINTER(OgXform), C.c_int,
                       C.POINTER(OgXform), C.c_int, P, P, C.POINTER(OgTrace), P, C.c_double]
        fn.restype = C.c_int
        rc = fn(C.byref(p), C.byref(gF.c()), f.ctypes.data_as(P), C.byref(gM.c()), m.ctypes.data_as(P), fx,
                len(fixed_xforms), mx, len(moving_xforms), fwd.ctypes.data_as(P), inv.ctypes.data_as(P), C.byref(tr),
                state.ctypes.data_as(P) if state is not None else None, eps)
    else:
        fn = L.og_syn_register_f64 if precision == "f64" else L.og_syn_register_f32
        rc = fn(C.byref(p), C.byref(gF.c()), f.ctypes.data_as(P), C.byref(gM.c()), m.ctypes.data_as(P),
                fwd.ctypes.data_as(P), inv.ctypes.data_as(P), C.byref(tr),
                state.ctypes.data_as(P) if state is not None else None, eps)
    if rc != 0:
        raise RuntimeError(f"oracle syn_register failed: {rc}")
    n = tr.n_records
    return SynResult(fwd, inv, lv[:n].copy(), it[:n].copy(), en[:n].copy(), cv[:n].copy(),
                     ki[: 4 * n].reshape(n, 4).copy(), state, lvf, lvi, secs[:n].copy())


def max_threads() -> int:
    return lib().og_max_threads()


# --- antsApplyTransforms subset -------------------------------------------

class OgGeom(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int64 * 3), ("spacing", C.c_double * 3), ("origin", C.c_double * 3),
                ("direction", C.c_double * 9)]


class OgXform(C.Structure):
    _fields_ = [("kind", C.c_int), ("matrix", C.c_double * 9), ("translation", C.c_double * 3),
                ("center", C.c_double * 3), ("field", OgGeom), ("f", C.POINTER(C.c_float))]


INTERP = {"linear": 0, "nearestneighbor": 1, "genericlabel": 2}
for _q in range(6):
    INTERP[f"bspline[{_q}]"] = 16 + _q
INTERP["bspline"] = 16 + 3
for _w, _n in enumerate(("cosine", "hamming", "lanczos", "blackman", "welch")):
    INTERP[_n + "windowedsinc"] = 32 + _w


def og_geom(size, spacing=None, origin=None, direction=None) -> OgGeom:
    """size in (x, y[, z]) order; direction row-major dim x dim (identity by default)."""
    D = len(size)
    g = OgGeom()
    g.dim = D
    for k in range(3):
        g.n[k] = int(size[k]) if k < D else 1
        g.spacing[k] = float(spacing[k]) if spacing is not None and k < D else 1.0
        g.origin[k] = float(origin[k]) if origin is not None and k < D else 0.0
    dm = np.eye(D) if direction is None else np.asarray(direction, dtype=np.float64).reshape(D, D)
    for k, v in enumerate(dm.ravel()):
        g.direction[k] = float(v)
    return g


def _xform_array(transforms, D):
    """OgXform array for ("affine", A, t, c) / ("field", OgGeom, field(z,y,x,D)) /
    ("identity",) tuples; returns (array, buffers to keep alive)."""
    xs = (OgXform * max(len(transforms), 1))()
    keep = []
    for i, t in enumerate(transforms):
        if t[0] == "identity":
            xs[i].kind = 0
        elif t[0] == "affine":
            xs[i].kind = 1
            A, tr, ce = np.asarray(t[1], float).reshape(D, D), np.asarray(t[2], float), np.asarray(t[3], float)
            for k, v in enumerate(A.ravel()):
                xs[i].matrix[k] = v
            for k in range(D):
                xs[i].translation[k] = tr[k]
                xs[i].center[k] = ce[k]
        elif t[0] == "field":
            xs[i].kind = 2
            xs[i].field = t[1]
            f = np.ascontiguousarray(t[2], dtype=np.float32)
            keep.append(f)
            xs[i].f = f.ctypes.data_as(C.POINTER(C.c_float))
        else:
            raise ValueError(t[0])
    return xs, keep


def apply_transforms(ref: OgGeom, inp: OgGeom, img, transforms=(), interpolation="linear", default=0.0):
    """transforms: sequence of ("affine", A, t, c) / ("field", OgGeom, field(z,y,x,D)) / ("identity",),
    applied to each reference point in order.  Returns float32 on the reference grid (numpy order)."""
    D = ref.dim
    a = np.ascontiguousarray(img, dtype=np.float32)
    xs, keep = _xform_array(transforms, D)
    shape = tuple(int(ref.n[k]) for k in range(D))[::-1]
    out = np.empty(shape, np.float32)
    fn = lib().og_apply_transforms_f64
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(OgGeom), C.POINTER(OgGeom), C.POINTER(C.c_float), C.POINTER(OgXform), C.c_int, C.c_int,
                   C.c_double, C.POINTER(C.c_float)]
    rc = fn(C.byref(ref), C.byref(inp), a.ctypes.data_as(C.POINTER(C.c_float)), xs, len(transforms),
            INTERP[interpolation.lower()], float(default), out.ctypes.data_as(C.POINTER(C.c_float)))
    if rc != 0:
        raise ValueError("dimension mismatch" if rc == -1 else "unknown interpolation")
    return out


def jacobian(g: OgGeom, field, do_log=False, gradient=False, geometric=False):
    """CreateJacobianDeterminantImage of a (z, y, x, D) field on grid g:
    the clamped determinant (or its log), or with `gradient` the matrices;
    `geometric`: the moved-simplex volume ratio."""
    D = g.dim
    f = np.ascontiguousarray(field, dtype=np.float32)
    shape = f.shape[:-1]
    if geometric:
        out = np.empty(shape, np.float64)
        fn = lib().og_jacobian_geometric_f64
        fn.restype = None
        fn.argtypes = [C.POINTER(OgGeom), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_double)]
        fn(C.byref(g), f.ctypes.data_as(C.POINTER(C.c_float)), int(bool(do_log)),
           out.ctypes.data_as(C.POINTER(C.c_double)))
        return out
    out = np.empty(shape, np.float64)
    grad = np.empty(shape + (D, D), np.float64) if gradient else None
    fn = lib().og_jacobian_f64
    fn.restype = None
    fn.argtypes = [C.POINTER(OgGeom), C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    fn(C.byref(g), f.ctypes.data_as(C.POINTER(C.c_float)), int(bool(do_log)), out.ctypes.data_as(C.POINTER(C.c_double)),
       grad.ctypes.data_as(C.POINTER(C.c_double)) if gradient else None)
    return grad if gradient else out


def transform_points(points, transforms=(), dim=3):
    """The first `dim` columns of each row through the transform list."""
    pts = np.ascontiguousarray(np.asarray(points, np.float64)[:, :dim])
    xs, keep = _xform_array(transforms, dim)
    out = np.empty_like(pts)
    fn = lib().og_transform_points_f64
    fn.restype = None
    fn.argtypes = [C.c_int, C.POINTER(C.c_double), C.c_int64, C.POINTER(OgXform), C.c_int, C.POINTER(C.c_double)]
    fn(dim, pts.ctypes.data_as(C.POINTER(C.c_double)), pts.shape[0], xs, len(transforms),
       out.ctypes.data_as(C.POINTER(C.c_double)))
    return out


def compose_field(ref: OgGeom, transforms=()):
    """T(p) - p at every point of `ref` (transforms as in apply_transforms);
    float32, shape (z, y, x, D) in numpy order."""
    D = ref.dim
    xs, keep = _xform_array(transforms, D)
    shape = tuple(int(ref.n[k]) for k in range(D))[::-1] + (D,)
    out = np.empty(shape, np.float32)
    fn = lib().og_compose_field_f64
    fn.restype = C.c_int
    fn.argtypes = [C.POINTER(OgGeom), C.POINTER(OgXform), C.c_int, C.POINTER(C.c_float)]
    fn(C.byref(ref), xs, len(transforms), out.ctypes.data_as(C.POINTER(C.c_float)))
    return out
