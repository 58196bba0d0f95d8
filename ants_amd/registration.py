"""Host-side mirror of the antsRegistration SyN stage over the gfx950 library.

``SyNRegistration`` is configured the way antsRegistration configures
RegistrationHelper for `-m CC[...] -t SyN[...] -c [...] -f ... -s ...`
(Examples/antsRegistrationTemplateHeader.h:542-815, 985-1086) and runs
through the C-ABI (``ants_syn_register`` / ``ants_syn_register_device``).
Results: forward and inverse displacement fields (numpy, (z,y,x,D) /
(y,x,D), mm) and the per-iteration observer trace
(Examples/antsDisplacementAndVelocityFieldRegistrationCommandIterationUpdate.h:117-195).
"""
from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

import numpy as np

from . import _lib as L


# ---------------------------------------------------------------------------
# antsCommandLineParser option-function grammar: name[p0,p1,...]
# (Utilities/antsCommandLineParser.cxx:87-198) and MxNxO vectors (ConvertVector).
def parse_option_function(text: str) -> Tuple[str, List[str]]:
    text = text.strip()
    m = re.fullmatch(r"([^\[\]]*)\[(.*)\]", text)
    if not m:
        return text, []
    name, body = m.group(1), m.group(2)
    return name, [p.strip() for p in body.split(",")] if body.strip() else []


def parse_vector(text: str, typ=int) -> List:
    return [typ(v) for v in text.strip().split("x") if v != ""]


def mesh_for_knot_spacing(size, spacing, knot_spacing: float) -> List[int]:
    """RegistrationHelper::CalculateMeshSizeForSpecifiedKnotSpacing
    (Examples/itkantsRegistrationHelper.hxx:3737-3767): per axis
    ceil((size - 1) * spacing / knot_spacing); a knot spacing of 0 gives 0."""
    if abs(float(knot_spacing)) <= 4 * np.finfo(np.float32).eps:
        return [0] * len(size)
    return [int(np.ceil((int(n) - 1) * float(sp) / float(knot_spacing))) for n, sp in zip(size, spacing)]


@dataclass
class SyNTransform:
    """`-t SyN[gradientStep, updateFieldVariance=3, totalFieldVariance=0]`
    (Examples/antsRegistrationTemplateHeader.h:801-815) or
    `-t BSplineSyN[gradientStep, updateMesh, totalMesh=0, splineOrder=3]`
    (Examples/antsRegistrationTemplateHeader.h:701-743): a mesh is `AxBxC`
    (spans per axis at the first level) or one value, a knot spacing in mm,
    converted with the fixed image's geometry."""
    gradient_step: float
    update_variance: float = 3.0
    total_variance: float = 0.0
    model: str = "SyN"
    update_mesh: Optional[List[int]] = None
    total_mesh: Optional[List[int]] = None
    spline_order: int = 3

    @classmethod
    def parse(cls, text: str, fixed_size=None, fixed_spacing=None) -> "SyNTransform":
        name, p = parse_option_function(text)
        if name.lower() == "bsplinesyn":
            if len(p) < 2:
                raise ValueError("BSplineSyN needs a gradient step and an update mesh: BSplineSyN[step,mesh,...]")
            order = int(p[3]) if len(p) > 3 else 3

            def mesh(spec, dim=None):
                v = parse_vector(spec, float)
                if len(v) != 1:
                    return [int(x) for x in v]
                if int(v[0]) == 0 and dim is not None:
                    return [0] * dim     # a zero knot spacing is a zero mesh whatever the geometry
                if fixed_size is None:
                    raise ValueError("a knot spacing needs the fixed image geometry (or give the mesh as AxBxC)")
                sp = fixed_spacing if fixed_spacing is not None else [1.0] * len(fixed_size)
                return mesh_for_knot_spacing(fixed_size, sp, int(v[0]))
            um = mesh(p[1])
            tm = mesh(p[2], len(um)) if len(p) > 2 else [0] * len(um)
            return cls(float(p[0]), model="BSplineSyN", update_mesh=um, total_mesh=tm, spline_order=order)
        if name.lower() != "syn":
            raise ValueError(f"only the SyN and BSplineSyN transforms are provided here, got {name!r}")
        if not p:
            raise ValueError("SyN needs a gradient step: SyN[step,...]")
        return cls(float(p[0]), float(p[1]) if len(p) > 1 else 3.0, float(p[2]) if len(p) > 2 else 0.0)


@dataclass
class CCMetric:
    """`-m CC[fixed, moving, weight=1, radius=4]`
    (Examples/antsRegistrationTemplateHeader.h:1005-1086)."""
    fixed: str = ""
    moving: str = ""
    weight: float = 1.0
    radius: int = 4

    @classmethod
    def parse(cls, text: str) -> "CCMetric":
        name, p = parse_option_function(text)
        if name.upper() != "CC":
            raise ValueError(f"only the CC metric is provided here, got {name!r}")
        if len(p) < 2:
            raise ValueError("CC needs fixed and moving images: CC[fixed,moving,...]")
        return cls(p[0], p[1], float(p[2]) if len(p) > 2 else 1.0, int(p[3]) if len(p) > 3 else 4)


@dataclass
class Convergence:
    """`-c [100x70x20, 1e-6, 10]` or `-c 100x70x20` (defaults 1e-6 and 10,
    Examples/antsRegistrationTemplateHeader.h:546-611)."""
    iterations: List[int]
    threshold: float = 1e-6
    window: int = 10

    @classmethod
    def parse(cls, text: str) -> "Convergence":
        name, p = parse_option_function(text)
        if not p:
            return cls(parse_vector(name))
        return cls(parse_vector(p[0]), float(p[1]) if len(p) > 1 else 1e-6, int(p[2]) if len(p) > 2 else 10)


def parse_sigmas(text: str, n_levels: int) -> Tuple[List[float], bool]:
    """`-s 2x1x0vox|mm` (no suffix = voxels; one value is repeated over the
    levels), Examples/antsRegistrationTemplateHeader.h:627-651."""
    t = text.strip()
    in_mm = False
    if "mm" in t:
        t = t.replace("mm", "", 1)
        in_mm = True
    elif "vox" in t:
        t = t.replace("vox", "", 1)
    s = parse_vector(t, float)
    if len(s) == 1:
        s = s * n_levels
    return s, in_mm


@dataclass
class SyNStage:
    transform: SyNTransform
    metric: CCMetric
    convergence: Convergence
    shrink: List[int]
    sigmas: List[float]
    sigmas_in_mm: bool = False

    @classmethod
    def from_args(cls, transform: str, metric: str, convergence: str, shrink: str, smoothing: str,
                  fixed_size=None, fixed_spacing=None) -> "SyNStage":
        """fixed_size / fixed_spacing (x, y[, z]): the fixed image's geometry,
        needed only for a BSplineSyN mesh given as a knot spacing."""
        t = SyNTransform.parse(transform, fixed_size, fixed_spacing)
        m = CCMetric.parse(metric)
        c = Convergence.parse(convergence)
        f = parse_vector(shrink)
        s, mm = parse_sigmas(smoothing, len(c.iterations))
        if not (len(f) == len(s) == len(c.iterations)):
            # RegistrationHelper refuses mismatched schedules (itkantsRegistrationHelper.hxx:1257-1283)
            raise ValueError("the number of shrink factors, smoothing sigmas and iteration counts must match")
        if m.weight != 1.0:
            raise ValueError("a single CC metric with weight 1 is supported")
        return cls(t, m, c, f, s, mm)


@dataclass
class IterationRecord:
    level: int
    iteration: int
    energy: float
    convergence: float
    inv_iters: Tuple[int, int, int, int]
    full_scale_cc: float = 0.0     # --print-similarity-measure-interval value (0: not computed)


@dataclass
class SyNResult:
    fwd: np.ndarray
    inv: np.ndarray
    trace: List[IterationRecord]
    stats: dict

    @property
    def energy(self) -> np.ndarray:
        return np.array([r.energy for r in self.trace])

    @property
    def levels(self) -> np.ndarray:
        return np.array([r.level for r in self.trace])

    def diagnostic_lines(self) -> List[str]:
        """The observer's `1DIAGNOSTIC, iter, E, conv, ...` lines (with the
        full-scale CC appended where it was computed and |value| > 1e-7, as
        the observer prints it)."""
        out = []
        for r in self.trace:
            ln = f"1DIAGNOSTIC, {r.iteration:5d}, {r.energy:.12e}, {r.convergence:.12e}"
            if abs(r.full_scale_cc) > 1e-7:
                ln += f", {r.full_scale_cc:.12e}"
            out.append(ln)
        return out


def stage_params(stage: "SyNStage", dim: int, precision: str = "fp32", preprocess: bool = True,
                 force_iterations: bool = False) -> L.SynParamsC:
    """The C parameter block of a stage (ants_syn_params)."""
    p = L.SynParamsC()
    p.dim = dim
    n = len(stage.shrink)
    if n > L.MAX_LEVELS:
        raise ValueError(f"at most {L.MAX_LEVELS} levels")
    p.n_levels = n
    for i in range(n):
        p.shrink[i] = int(stage.shrink[i])
        p.sigma[i] = float(stage.sigmas[i])
        p.iterations[i] = int(stage.convergence.iterations[i])
    p.sigma_in_mm = int(stage.sigmas_in_mm)
    p.convergence_threshold = stage.convergence.threshold
    p.convergence_window = stage.convergence.window
    p.learning_rate = stage.transform.gradient_step
    p.update_variance = stage.transform.update_variance
    p.total_variance = stage.transform.total_variance
    p.cc_radius = stage.metric.radius
    if stage.transform.model == "BSplineSyN":
        t = stage.transform
        if len(t.update_mesh) != dim or len(t.total_mesh) != dim:
            # itkantsRegistrationHelper.hxx:2239-2243
            raise ValueError("the mesh size(s) don't match the image dimension")
        p.model = L.MODEL_BSPLINE_SYN
        for d in range(dim):
            p.update_mesh[d] = int(stage.transform.update_mesh[d])
            p.total_mesh[d] = int(stage.transform.total_mesh[d])
        p.spline_order = int(stage.transform.spline_order)
    p.precision = {"fp32": 0, "bf16": 1}[precision]
    p.preprocess = int(preprocess)
    p.force_iterations = int(force_iterations)
    return p


class SyNRegistration:
    """One SyN + CC stage (RegistrationHelper::AddSyNTransform + AddMetric +
    schedule setters), run on the GPU through the C-ABI."""

    def __init__(self, stage: SyNStage, dim: int, precision: str = "fp32", preprocess: bool = True,
                 force_iterations: bool = False):
        if dim not in (2, 3):
            raise ValueError("dim must be 2 or 3")
        self.stage = stage
        self.dim = dim
        p = stage_params(stage, dim, precision, preprocess, force_iterations)
        self.params = p
        h = C.c_void_p()
        L.check(L.lib().ants_syn_create(C.byref(p), C.byref(h)))
        self.handle = h

    def close(self):
        if getattr(self, "handle", None):
            L.lib().ants_syn_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def level_geometry(self, size, spacing=None, origin=None, level: int = -1, direction=None) -> L.ImageDesc:
        lvl = level if level >= 0 else self.params.n_levels - 1
        out = L.ImageDesc()
        L.check(L.lib().ants_syn_level_geometry(C.byref(self.params),
                                                C.byref(L.desc(size, spacing, origin, direction)), lvl,
                                                C.byref(out)))
        return out

    def set_initial_transforms(self, moving=(), fixed=()):
        """SetMovingInitialTransform / SetFixedInitialTransform
        (Examples/itkantsRegistrationHelper.hxx:2119-2137): lists of
        ``apply_transforms.AffineTransform`` / ``DisplacementFieldTransform``
        (or None for identity) in the
        order a point passes through them, as ``-t`` options are listed.
        Kept for every later run of this object; empty lists clear them."""
        from .apply_transforms import transform_array
        mx, keep_m = transform_array(list(moving), self.dim)
        fx, keep_f = transform_array(list(fixed), self.dim)
        L.check(L.lib().ants_syn_set_initial_transforms(self.handle, mx, len(moving), fx, len(fixed)))
        # the handle refers to the displacement fields' device buffers for its later runs
        self._initial_keep = keep_m + keep_f

    def set_initial_state(self, state=None, size=None, spacing=None, origin=None, direction=None):
        """Start later runs from a saved state (SetFixedToMiddleTransform /
        SetMovingToMiddleTransform, restored by --restore-state): `state` is
        the four fields F2M fwd, F2M inv, M2M fwd, M2M inv as `state()`
        returns them, on the grid (size (x, y[, z]), spacing, origin,
        direction); None clears it."""
        if state is None:
            L.check(L.lib().ants_syn_set_initial_state_device(self.handle, None, None, None, None, None))
            return
        if len(state) != 4:
            raise ValueError("a state is four fields")
        bufs = [f if isinstance(f, L.DeviceArray) else L.DeviceArray.from_host(np.asarray(f, np.float32))
                for f in state]
        g = L.desc(size, spacing, origin, direction)
        L.check(L.lib().ants_syn_set_initial_state_device(self.handle, C.byref(g), *[b.ptr for b in bufs]))

    def set_full_scale_cc_interval(self, interval: int):
        """--print-similarity-measure-interval: the observer's CC r=4 between
        the original images at iteration 1, every `interval`-th and each
        level's last (ants_syn_set_full_scale_cc_interval); 0 turns it off."""
        L.check(L.lib().ants_syn_set_full_scale_cc_interval(self.handle, int(interval)))

    def set_profiling(self, mask: int, every: int = 1):
        """Time the kernel classes in `mask` on every `every`-th iteration of
        each level (ants_syn_set_profiling / _stride)."""
        L.check(L.lib().ants_syn_set_profiling(self.handle, mask))
        L.check(L.lib().ants_syn_set_profiling_stride(self.handle, every))

    def kernel_stats(self, level: int = None) -> dict:
        """Per kernel class: event-timed ms, launches and algorithmic bytes of
        the last run (of one pyramid level when `level` is given)."""
        k = L.KernelStatsC()
        if level is None:
            L.check(L.lib().ants_syn_get_kernel_stats(self.handle, C.byref(k)))
        else:
            L.check(L.lib().ants_syn_get_kernel_stats_level(self.handle, level, C.byref(k)))
        return {n: {"ms": k.ms[i], "launches": int(k.launches[i]), "bytes": k.bytes[i]}
                for i, n in enumerate(L.KC_NAMES)}

    def stats(self) -> dict:
        s = L.SynStatsC()
        L.check(L.lib().ants_syn_get_stats(self.handle, C.byref(s)))
        n = s.levels_run
        return {"levels_run": n, "iterations": list(s.iterations[:n]), "level_voxels": list(s.level_voxels[:n]),
                "inv_steps": list(s.inv_steps[:n]), "voxel_iterations": s.voxel_iterations,
                "algorithmic_bytes": s.algorithmic_bytes, "graph_launches": int(s.graph_launches),
                "eager_iterations": int(s.eager_iterations), "inv_tails": int(s.inv_tails)}

    @staticmethod
    def _callback(observer, handle=None):
        """ctypes observer for a live per-iteration callback, or a null
        callback.  The engine keeps the trace itself (`trace()`), so no
        callback is passed unless the caller asks for one: a Python call
        between iterations holds the next iteration back.  With `handle` the
        record carries the iteration's full-scale CC."""
        if observer is None:
            return L.ITER_CB()

        def cb(user, level, it, e, conv, kin):
            v = C.c_double(0.0)
            if handle:
                L.check(L.lib().ants_syn_last_full_scale_cc(handle, C.byref(v)))
            rec = IterationRecord(level, it, e, conv, tuple(kin[i] for i in range(4)), v.value)
            return 1 if observer(rec) else 0
        return L.ITER_CB(cb)

    def trace(self) -> List[IterationRecord]:
        """Iteration records of the last run (ants_syn_get_trace)."""
        n = C.c_int64(0)
        L.check(L.lib().ants_syn_get_trace(self.handle, None, 0, C.byref(n)))
        buf = (L.IterRecordC * max(n.value, 1))()
        L.check(L.lib().ants_syn_get_trace(self.handle, buf, n.value, C.byref(n)))
        return [IterationRecord(r.level, r.iteration, r.energy, r.convergence, tuple(r.inv_iters), r.full_scale_cc)
                for r in buf[:n.value]]

    def register(self, fixed: np.ndarray, moving: np.ndarray, fixed_spacing=None, fixed_origin=None,
                 moving_spacing=None, moving_origin=None, fixed_direction=None, moving_direction=None,
                 observer=None) -> SyNResult:
        """Host arrays in (numpy order, x fastest), host fields out (physical
        vectors on the fixed image's last level grid).  Directions are
        row-major dim x dim orthonormal matrices (identity by default).
        `observer(record) -> bool` is called after every iteration (True
        cancels the run); without it the trace is read after the run."""
        fixed = np.ascontiguousarray(fixed, dtype=np.float32)
        moving = np.ascontiguousarray(moving, dtype=np.float32)
        if fixed.ndim != self.dim or moving.ndim != self.dim:
            raise ValueError("image dimension does not match the stage")
        fsize = tuple(reversed(fixed.shape))
        msize = tuple(reversed(moving.shape))
        gf = L.desc(fsize, fixed_spacing, fixed_origin, fixed_direction)
        gm = L.desc(msize, moving_spacing, moving_origin, moving_direction)
        out = self.level_geometry(fsize, fixed_spacing, fixed_origin, direction=fixed_direction)
        oshape = tuple(int(out.size[k]) for k in reversed(range(self.dim))) + (self.dim,)
        fwd = np.empty(oshape, np.float32)
        inv = np.empty(oshape, np.float32)
        cb = self._callback(observer, self.handle)
        L.check(L.lib().ants_syn_register(self.handle, C.byref(gf), fixed.ctypes.data_as(C.c_void_p), C.byref(gm),
                                          moving.ctypes.data_as(C.c_void_p), fwd.ctypes.data_as(C.c_void_p),
                                          inv.ctypes.data_as(C.c_void_p), cb, None))
        return SyNResult(fwd, inv, self.trace(), self.stats())

    def register_device(self, gf: L.ImageDesc, d_fixed: L.DeviceArray, gm: L.ImageDesc, d_moving: L.DeviceArray,
                        d_fwd: Optional[L.DeviceArray], d_inv: Optional[L.DeviceArray], trace: Optional[list] = None,
                        observer=None):
        """Device buffers in and out (inputs already resident in HBM).
        Returns the iteration records (appended to `trace` when given)."""
        tr = trace if trace is not None else []
        cb = self._callback(observer, self.handle)
        L.check(L.lib().ants_syn_register_device(self.handle, C.byref(gf), d_fixed.ptr, C.byref(gm), d_moving.ptr,
                                                 d_fwd.ptr if d_fwd else None, d_inv.ptr if d_inv else None, cb,
                                                 None))
        tr.extend(self.trace())
        return tr

    def state(self, out_shape) -> List[np.ndarray]:
        """F2M fwd, F2M inv, M2M fwd, M2M inv after the last run (host copies),
        the registration state --save-state keeps."""
        bufs = [L.DeviceArray(out_shape) for _ in range(4)]
        L.check(L.lib().ants_syn_get_state_device(self.handle, *[b.ptr for b in bufs]))
        return [b.to_host() for b in bufs]


def center_of_mass(image: np.ndarray, size, spacing=None, origin=None, direction=None) -> np.ndarray:
    """ImageMomentsCalculator's centre of gravity of an image (numpy (z, y,
    x) order), on the GPU (ants_image_center_of_mass)."""
    d = L.DeviceArray.from_host(np.ascontiguousarray(image, np.float32))
    g = L.desc(size, spacing, origin, direction)
    D = len(size)
    c = (C.c_double * 3)()
    L.check(L.lib().ants_image_center_of_mass(C.byref(g), d.ptr, c, None, None))
    return np.array([c[k] for k in range(D)])


def initial_transform_from_images(fixed: np.ndarray, fixed_geom, moving: np.ndarray, moving_geom, feature: int):
    """antsRegistration -r [fixed,moving,feature] (Examples/itkantsRegistrationHelper.h:1296-1360):
    a rigid transform with no rotation, centred at the fixed image's feature
    point and translated to the moving image's: 0 image centres (the
    continuous index (size - 1) / 2, CenteredTransformInitializer::GeometryOn),
    1 centres of mass (MomentsOn), 2 image origins.  Geometries are
    apply_transforms.Geometry objects; returns an AffineTransform."""
    from .apply_transforms import AffineTransform

    def geom_point(geo, idx):
        D = geo.dim
        sp = np.ones(D) if geo.spacing is None else np.asarray(geo.spacing, float)
        org = np.zeros(D) if geo.origin is None else np.asarray(geo.origin, float)
        dm = np.eye(D) if geo.direction is None else np.asarray(geo.direction, float).reshape(D, D)
        return org + dm @ (sp * np.asarray(idx, float))

    D = fixed_geom.dim
    if feature == 0:
        cf = geom_point(fixed_geom, (np.asarray(fixed_geom.size, float) - 1) / 2)
        cm = geom_point(moving_geom, (np.asarray(moving_geom.size, float) - 1) / 2)
    elif feature == 1:
        cf = center_of_mass(fixed, fixed_geom.size, fixed_geom.spacing, fixed_geom.origin, fixed_geom.direction)
        cm = center_of_mass(moving, moving_geom.size, moving_geom.spacing, moving_geom.origin, moving_geom.direction)
    elif feature == 2:
        cf = geom_point(fixed_geom, np.zeros(D))
        cm = geom_point(moving_geom, np.zeros(D))
    else:
        raise ValueError("initialization feature must be 0 (centres), 1 (centres of mass) or 2 (origins)")
    return AffineTransform(np.eye(D), cm - cf, cf)


def registration_from_cli(args: Sequence[str]) -> Tuple[SyNRegistration, SyNStage]:
    """Build a SyNRegistration from antsRegistration-style arguments
    (`-d 3 -m CC[f,m,1,4] -t SyN[0.1,3,0] -c [...] -f 4x2x1 -s 2x1x0vox`)."""
    opts = {}
    it = iter(args)
    for a in it:
        if a in ("-d", "--dimensionality", "-m", "--metric", "-t", "--transform", "-c", "--convergence", "-f",
                 "--shrink-factors", "-s", "--smoothing-sigmas", "--float"):
            opts[a.lstrip("-")[0] if len(a) == 2 else a] = next(it)
    dim = int(opts.get("d", opts.get("--dimensionality", 3)))
    stage = SyNStage.from_args(opts.get("t", opts.get("--transform")), opts.get("m", opts.get("--metric")),
                               opts.get("c", opts.get("--convergence")), opts.get("f", opts.get("--shrink-factors")),
                               opts.get("s", opts.get("--smoothing-sigmas")))
    return SyNRegistration(stage, dim), stage
