"""The field-transform stage oracle (oracle/field_oracle.py) pinned on the
CPU: its metric derivative against the SyN stage's pinned CC derivative
where the two coincide, its mapping direction by an integer translation,
the exponential's step rule by hand values and its result against an
independent Runge-Kutta integration of the flow, and whole stages by
recovering a known translation."""
import math

import numpy as np
import pytest
from scipy import ndimage

import oracle as O
import field_oracle as FO


def blob_pair(shape, shift, seed=0):
    rng = np.random.default_rng(seed)
    img = np.zeros(shape)
    inner = tuple(slice(s // 4, s - s // 4) for s in shape)
    img[inner] = rng.standard_normal(img[inner].shape)
    F = ndimage.gaussian_filter(img, 1.5)
    M = ndimage.shift(F, shift, order=1)
    return F, M


def test_metric_at_zero_field_equals_syn_update():
    """u = 0 on the full-resolution grid with no initial transforms: the
    mapped points are the voxels, the half-spacing central difference of the
    linearly interpolated moving image is the SyN stage's gradient rule
    (half difference at index 0, zero at n - 1), so the derivative equals
    the SyN fixed-to-middle update of the same images, and the value its
    energy.  (Spacings and origin are binary fractions so that the lower
    neighbour of an index-0 voxel lands exactly on the buffer's -0.5 edge.)"""
    g = O.Grid((19, 17, 15), (1.0, 0.5, 2.0), (3.0, -2.0, 1.0))
    F, M = blob_pair(g.shape, (0.7, -1.2, 0.4), seed=1)
    m = FO._Metric(g, g, F, g, M, 2, (np.eye(3), np.zeros(3)), (np.eye(3), np.zeros(3)))
    v, d = m(np.zeros(g.shape + (3,)))
    vs, gs, _ = O.cc(g, F, M, 2)
    assert v == pytest.approx(vs, rel=1e-13)
    np.testing.assert_allclose(d, gs, rtol=0, atol=1e-12 * np.abs(gs).max())


def test_metric_integer_translation_maps_moving_points():
    """u = t (whole voxels): the moving image is sampled at x + t, so the
    value is the CC of F with M shifted by -t (interior unaffected by the
    zero fill), a wrong sign or an index/physical mix-up fails."""
    g = O.Grid((20, 18, 16), (1.0, 1.0, 1.0))
    F, M = blob_pair(g.shape, (0.0, 0.0, 0.0), seed=2)
    t = np.array([2.0, -1.0, 1.0])          # x, y, z in mm (= voxels)
    u = np.broadcast_to(t, g.shape + (3,)).copy()
    m = FO._Metric(g, g, F, g, M, 2, (np.eye(3), np.zeros(3)), (np.eye(3), np.zeros(3)))
    v, _ = m(u)
    Ms = np.zeros_like(M)
    Ms[:-1, 1:, :-2] = M[1:, :-1, 2:]        # Ms(z, y, x) = M(z + 1, y - 1, x + 2), zero outside
    vs, _, _ = O.cc(g, F, Ms, 2)
    assert v == pytest.approx(vs, rel=1e-13)


def test_exp_step_rule_hand_values():
    g = O.Grid((8, 8, 8), (1.0, 0.5, 2.0))
    v = np.zeros(g.shape + (3,))
    assert FO.exp_steps(g, v, 0) == 0
    v[3, 3, 3] = [1.5, 0.0, 0.0]             # max |v| / min spacing = 3 -> 2 + log2(3) = 3.58 -> 4
    assert FO.exp_steps(g, v, 0) == 4
    v[3, 3, 3] = [0.1, 0.0, 0.0]             # 0.2 -> 2 + log2(0.2) = -0.32 < 0 -> 0
    assert FO.exp_steps(g, v, 0) == 0
    v[3, 3, 3] = [0.125, 0.0, 0.0]           # 0.25 -> 2 + log2(0.25) = 0 -> 1
    assert FO.exp_steps(g, v, 0) == 1
    v[3, 3, 3] = [1e6, 0.0, 0.0]             # capped at the default step count
    assert FO.exp_steps(g, v, 0) == FO.EXP_DEFAULT_STEPS
    assert FO.exp_steps(g, v, 7) == 7
    phi, n = FO.exponentiate(g, np.zeros(g.shape + (3,)))
    assert n == 0 and not phi.any()


def _rk4_flow(g, v, steps=64):
    """x(1) - x(0) for dx/dt = v(x), v linearly interpolated (index space,
    unit spacing), fourth-order Runge-Kutta with `steps` steps."""
    D = g.dim
    grids = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in reversed(g.size)], indexing="ij")
    pos = np.stack(grids, axis=0)            # (z, y, x) coordinates
    comps = [v[..., D - 1 - k] for k in range(D)]

    def vel(p):
        return np.stack([ndimage.map_coordinates(c, p, order=1, mode="nearest") for c in comps], axis=0)

    x = pos.copy()
    h = 1.0 / steps
    for _ in range(steps):
        k1 = vel(x)
        k2 = vel(x + 0.5 * h * k1)
        k3 = vel(x + 0.5 * h * k2)
        k4 = vel(x + h * k3)
        x = x + h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
    d = x - pos
    return np.stack([d[D - 1 - k] for k in range(D)], axis=-1)


def test_exponential_matches_integrated_flow():
    """exp(v) of a smooth field equals the flow of v integrated by RK4 in the
    interior (up to the scaling-and-squaring error), exp(-v) inverts it, and
    a shear field shows the second-order term a plain v would miss."""
    g = O.Grid((40, 36, 32))
    z, y, x = np.meshgrid(*[np.arange(n, dtype=np.float64) for n in reversed(g.size)], indexing="ij")
    v = np.zeros(g.shape + (3,))
    v[..., 0] = 2.0 * np.sin(2 * math.pi * y / 36.0)
    v[..., 1] = 1.5 * np.cos(2 * math.pi * x / 40.0)
    v[..., 2] = 0.5 * np.sin(2 * math.pi * (x + y) / 50.0)
    phi, n = FO.exponentiate(g, v)
    assert n == FO.exp_steps(g, v, 0) and n >= 3
    ref = _rk4_flow(g, v)
    core = (slice(6, -6),) * 3
    err = np.abs(phi - ref)[core].max()
    plain = np.abs(v - ref)[core].max()
    print(f"\nexp vs RK4 flow: {err:.4f} voxel (v itself is {plain:.3f} away), {n} squarings")
    assert err < 0.05 and plain > 10 * err
    inv, _ = FO.exponentiate(g, v, inverse=True)
    res = O.compose(g, inv, phi)                # phi + inv(x + phi): the round trip
    assert np.abs(res)[core].max() < 0.05


@pytest.mark.parametrize("model", [FO.GAUSSIAN_DISPLACEMENT, FO.EXPONENTIAL, FO.BSPLINE_DISPLACEMENT,
                                   FO.BSPLINE_EXPONENTIAL])
def test_stage_recovers_translation(model):
    """A blob moved by a known sub-voxel translation: the stage's field over
    the blob's core approaches that translation and the energy falls."""
    g = O.Grid((32, 30, 28))
    shift = np.array([1.2, -0.8, 0.6])        # z, y, x voxels of the moving content
    F, M = blob_pair(g.shape, tuple(shift), seed=3)
    expo = model in (FO.EXPONENTIAL, FO.BSPLINE_EXPONENTIAL)
    P = FO.FieldParams(model=model, learning_rate=20.0 if expo else 0.5, update_variance=2.0, total_variance=0.5,
                       radius=2, iterations=(30, 20), shrink=(2, 1), sigma=(1.0, 0.0), force_iterations=True,
                       update_mesh=(4, 4, 4), total_mesh=(0, 0, 0), spline_order=3)
    res = FO.field_register(P, g, F, g, M)
    core = tuple(slice(s // 2 - 3, s // 2 + 3) for s in g.shape)
    mean = res.field[core].reshape(-1, 3).mean(axis=0)       # x, y, z components
    want = shift[::-1]
    print(f"\nmodel {model}: mean core displacement {mean}, want {want}; energies "
          f"{res.records[0][2]:.4f} -> {res.records[-1][2]:.4f}")
    assert np.abs(mean - want).max() < 0.35
    assert res.records[-1][2] < res.records[0][2]
