"""CPU restatement of antsRegistration's dense field-transform stages with the
CC metric (test infrastructure only: tests/ use it as the checker of the GPU
path; nothing in ants_amd imports it).  Float64 numpy over the FP64 C
primitives of oracle.py (smoothing, composition, resampling, CC sums).

  -t GaussianDisplacementField[step, updateVariance, totalVariance]
     (Examples/itkantsRegistrationHelper.hxx:1733-1847)
  -t Exponential[step, updateVariance, velocityVariance, integrationSteps=0]
     (Examples/itkantsRegistrationHelper.hxx:3112-3277,
      Examples/antsRegistrationTemplateHeader.h:692-699, 871-886)
  -t BSplineDisplacementField[step, updateMesh, totalMesh=0, order=3] and
  -t BSplineExponential[step, updateMesh, velocityMesh=0, steps, order=3]
     (hxx:1848-1982, 3278-3461; h:701-743, 888-935): the same stages with
     the update and total / velocity fields fitted by B-splines (the
     BSplineSyN stage's fit) on meshes doubled per level, mesh << (level + 1)

Both run ImageRegistrationMethodv4: per level the virtual domain is the
shrunk fixed grid, the images are smoothed at full resolution, and the
transform's field is resampled onto the new grid (linear, zero outside).
ITK is not in the reference tree; each piece below restates the ITK class
the helper instantiates:

* Metric (ANTSNeighborhoodCorrelationImageToImageMetricv4, dense): the fixed
  image at the virtual points through the fixed initial transforms, the
  moving image at y = movingInit(x + u(x)); CC box sums over the virtual
  grid (the SyN stage's sums, out-of-buffer samples 0); value -sum localCC/N;
  derivative per voxel c1 * A^T grad M(y), with grad M(y) the
  CentralDifferenceImageFunction of the moving image at y (half a spacing
  either side along each physical axis, 0 where a neighbour leaves the
  buffer) and A the moving initial chain's linear part.  A field transform
  has local support, so the derivative is not divided by the point count.
* GaussianDisplacementField: GaussianSmoothingOnUpdateDisplacementField-
  Transform updated by ConjugateGradientLineSearchOptimizerv4 (the helper's
  optimizer for every non-SyN stage, hxx:1290-1372: learning rate and
  maximum physical step both `step`, line search over [0, 2] x the rate,
  epsilon 0.2, at most 20 steps).  The rate is estimated at each level's
  first iteration as step / max_x |g(x)| (the physical-shift estimator of a
  local-support transform moves each voxel by its own displacement, scales
  1); Polak-Ribiere conjugation with the gamma in [0, 5] restart.  An update
  w becomes u <- u + smooth(w, updateVariance), then u <- smooth(u,
  totalVariance) when that is > 0: additive, as DisplacementFieldTransform
  adds parameter updates.
* Exponential: GaussianExponentialDiffeomorphicTransform with
  GradientDescentOptimizerv4 and no scales estimator (hxx:1405-1418): the
  update is step * g; v <- v + smooth(update, updateVariance), then
  v <- smooth(v, velocityVariance) when > 0; the displacement is exp(v) and
  its inverse exp(-v) by ExponentialDisplacementFieldImageFilter.
* Field smoothing: GaussianSmoothDisplacementField (the SyN stage's: taps
  from GaussianOperator with maximum error 0.001 and width the axis size,
  boundary voxels zeroed, blended with the input for variances below 0.5).
* Convergence: the window monitor of -c, checked after each metric value;
  records (level, iteration, value, convergence, learning rate) after each
  update, as the linear stages keep them.

Restated from ITK without reference outputs (parity with ITK rests on the
restatement): the additive GaussianDisplacementField update, the
exponential's automatic step count and its cap (10, the transform's default
step count), and that a virtual point mapped outside the moving buffer
counts with a zero sample (ITK drops such points from the value and the
derivative).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List, Sequence

import numpy as np

import oracle as O
import linear_oracle as LO

GAUSSIAN_DISPLACEMENT, EXPONENTIAL, BSPLINE_DISPLACEMENT, BSPLINE_EXPONENTIAL = 0, 1, 2, 3
DBL_EPS = np.finfo(np.float64).eps
EXP_DEFAULT_STEPS = 10


@dataclass
class FieldParams:
    model: int = GAUSSIAN_DISPLACEMENT
    learning_rate: float = 0.25
    update_variance: float = 3.0
    total_variance: float = 0.0          # Exponential: the velocity field's variance
    integration_steps: int = 0           # Exponential: 0 = automatic
    radius: int = 4
    iterations: Sequence[int] = (20,)
    shrink: Sequence[int] = (1,)
    sigma: Sequence[float] = (0.0,)
    sigma_in_mm: bool = False
    conv_threshold: float = 1e-6
    conv_window: int = 10
    preprocess: bool = True
    force_iterations: bool = False
    update_mesh: Sequence[int] = (0, 0, 0)   # B-spline forms: spans at the base level
    total_mesh: Sequence[int] = (0, 0, 0)    # B-spline forms: total / velocity field (0: no fit)
    spline_order: int = 3


@dataclass
class FieldResult:
    field: np.ndarray                    # displacement on the last level grid (x fastest, D per voxel)
    inverse: np.ndarray = None           # Exponential: exp(-v)
    velocity: np.ndarray = None          # Exponential: v
    records: List[tuple] = field(default_factory=list)
    iterations: List[int] = field(default_factory=list)
    evaluations: int = 0


def exp_steps(g: O.Grid, v, steps: int) -> int:
    """ExponentialDisplacementFieldImageFilter's squaring count: given, or
    automatic -- the smallest n with max |v| / 2^n below half the smallest
    spacing, as floor(2 + log2(max |v| / min spacing) + 1), capped at
    EXP_DEFAULT_STEPS; 0 for a zero field."""
    if steps > 0:
        return int(steps)
    D = g.dim
    vv = np.asarray(v, dtype=np.float64).reshape(-1, D)
    m2 = float(np.max(np.sum(vv * vv, axis=1))) if vv.size else 0.0
    m2 /= min(float(s) for s in g.spacing) ** 2
    if m2 <= 0.0:
        return 0
    nf = 2.0 + 0.5 * math.log(m2) / math.log(2.0)
    if nf < 0.0:
        return 0
    return min(int(nf + 1.0), EXP_DEFAULT_STEPS)


def exponentiate(g: O.Grid, v, steps: int = 0, inverse: bool = False):
    """exp(v) (or exp(-v)) by scaling and squaring: phi = +-v / 2^n, then n
    times phi <- phi + phi(x + phi) (linear interpolation, zero outside)."""
    n = exp_steps(g, v, steps)
    phi = np.asarray(v, dtype=np.float64) * (-1.0 if inverse else 1.0)
    if n == 0:
        return phi.copy(), 0
    phi = phi / float(1 << n)
    for _ in range(n):
        phi = O.compose(g, phi, phi)
    return phi, n


def smooth(g: O.Grid, fld, variance: float):
    return O.smooth_field(g, fld, variance) if variance > 0.0 else np.asarray(fld, dtype=np.float64).copy()


def _smoother(P: FieldParams, gl: O.Grid, lev: int, total: bool):
    """The update or total / velocity field's smoothing on level lev: Gaussian
    with the stage's variance, or for the B-spline forms the B-spline fit of
    mesh << (lev + 1) (BSplineSmoothingOnUpdateDisplacementFieldTransform /
    BSplineExponentialDiffeomorphicTransform with the helper's per-level
    meshes, Examples/itkantsRegistrationHelper.hxx:1934-1941, 3371-3377),
    none for a zero mesh (control points not above the order)."""
    if P.model in (GAUSSIAN_DISPLACEMENT, EXPONENTIAL):
        var = P.total_variance if total else P.update_variance
        return lambda f: smooth(gl, f, var)
    base = P.total_mesh if total else P.update_mesh
    mesh = [int(m) << (lev + 1) for m in base[: gl.dim]]
    if any(m == 0 for m in mesh):
        return lambda f: np.asarray(f, dtype=np.float64).copy()
    return lambda f: O.bspline_fit_field(gl, f, mesh, P.spline_order)


class _Metric:
    """CC value and derivative of the virtual grid gl for a displacement u."""

    def __init__(self, gl, gF, Fs, gM, Ms, radius, chainM, chainF):
        self.gl, self.gM, self.Ms, self.radius = gl, gM, Ms, radius
        D = gl.dim
        self.pts = LO.grid_points(gl)
        A, b = chainF
        fp = self.pts @ np.asarray(A).T + np.asarray(b)
        self.Fw = LO.interpolate(gF, Fs, fp)[0].reshape(gl.shape)
        self.A, self.b = np.asarray(chainM[0], dtype=np.float64), np.asarray(chainM[1], dtype=np.float64)
        self.D = D

    def __call__(self, u, derivative=True):
        D = self.D
        p = self.pts + np.asarray(u, dtype=np.float64).reshape(-1, D)
        y = p @ self.A.T + self.b
        Mw = LO.interpolate(self.gM, self.Ms, y)[0].reshape(self.gl.shape)
        if not derivative:
            v, _, _ = O.cc(self.gl, self.Fw, Mw, self.radius)
            return v, None
        gr = LO.gradient(self.gM, self.Ms, y) @ self.A      # A^T grad M(y), per voxel
        v, g = O.cc_ext(self.gl, self.Fw, Mw, self.radius, gr.reshape(self.gl.shape + (D,)))
        return v, g


def field_register(P: FieldParams, gF: O.Grid, fixed, gM: O.Grid, moving, moving_init=(), fixed_init=()):
    """One GaussianDisplacementField or Exponential stage, level by level.
    moving_init / fixed_init: (matrix, translation, center) affines in the
    order a point passes them."""
    D = gF.dim
    F = np.asarray(fixed, dtype=np.float64)
    M = np.asarray(moving, dtype=np.float64)
    if P.preprocess:
        F = O.preprocess(F)
        M = O.preprocess(M)
    chainM = LO.affine_chain(moving_init, D)
    chainF = LO.affine_chain(fixed_init, D)
    res = FieldResult(None)
    state = None          # u (GaussianDisplacementField) or v (Exponential) on the level grid
    gprev = None
    for lev in range(len(P.iterations)):
        gl = O.shrink_grid(gF, int(P.shrink[lev]))
        Fs = O.smooth_image(gF, F, float(P.sigma[lev]), P.sigma_in_mm) if P.sigma[lev] > 0 else F
        Ms = O.smooth_image(gM, M, float(P.sigma[lev]), P.sigma_in_mm) if P.sigma[lev] > 0 else M
        if state is None:
            state = np.zeros(gl.shape + (D,))
        else:
            state = O.resample_field(gprev, state, gl)
        gprev = gl
        metric = _Metric(gl, gF, Fs, gM, Ms, P.radius, chainM, chainF)
        expo = P.model in (EXPONENTIAL, BSPLINE_EXPONENTIAL)
        s_upd, s_tot = _smoother(P, gl, lev, False), _smoother(P, gl, lev, True)
        u = exponentiate(gl, state, P.integration_steps)[0] if expo else state
        energies: List[float] = []
        conj = np.zeros_like(state)
        last = np.zeros_like(state)
        lr = P.learning_rate
        it = 0
        while it < P.iterations[lev]:
            val, g = metric(u)
            res.evaluations += 1
            energies.append(val)
            conv = O.conv_value(energies, P.conv_window)
            if conv <= P.conv_threshold and not P.force_iterations:
                break
            if not expo:
                if it == 0:
                    conj[:] = 0.0
                    last[:] = 0.0
                    ss = float(np.sqrt(np.max(np.sum(g.reshape(-1, D) ** 2, axis=1))))
                    lr = 1.0 if ss <= DBL_EPS else P.learning_rate / ss
                gamma = 0.0
                den = float(np.sum(last * last))
                if den > DBL_EPS:
                    gamma = float(np.sum((g - last) * g)) / den
                if gamma < 0.0 or gamma > 5.0:
                    gamma = 0.0
                last = g.copy()
                conj = g + gamma * conj
                direction = conj.copy()

                def update(step, base=state, direction=direction):
                    return s_tot(base + s_upd(step * direction))

                def f(step, update=update):
                    res.evaluations += 1
                    return metric(update(step), False)[0]

                lr = LO._LineSearch(f).search(lr * 0.0, lr, lr * 2.0)
                state = update(lr)
                u = state
            else:
                lr = P.learning_rate
                state = s_tot(state + s_upd(lr * g))
                u = exponentiate(gl, state, P.integration_steps)[0]
            res.records.append((lev, it + 1, val, conv, lr))
            it += 1
        res.iterations.append(it)
    if P.model in (EXPONENTIAL, BSPLINE_EXPONENTIAL):
        res.velocity = state
        res.field = exponentiate(gprev, state, P.integration_steps)[0]
        res.inverse = exponentiate(gprev, state, P.integration_steps, inverse=True)[0]
    else:
        res.field = state
    return res
