"""Pins of the oracle's B-spline field fit (BSplineSyN's update and total
field smoothing: ITK's DisplacementFieldToBSplineImageFilter with one level
of BSplineScatteredDataPointSetToImageFilter, stationary boundary) and of its
BSplineSyN stage.  CPU only.

The fit is one level of multilevel B-spline approximation with confidences
(Lee, Wolberg & Shin, IEEE TVCG 3(3) 1997): for data point c with basis
products B_ck, phi_k = sum_c w_c B_ck^2 (z_c B_ck / sum_l B_cl^2) /
sum_c w_c B_ck^2, evaluated back at every voxel.  mba_dense() below builds
that from scipy's B-spline basis as dense matrices -- the multi-dimensional
basis as a Kronecker product of the per-axis ones, in x-fastest voxel order
-- so a wrong parametric mapping, span index, basis order, axis order or
boundary rule in the oracle's point-by-point scatter fails the comparison.
"""
import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as o

rng = np.random.default_rng(77)
BOUNDARY_WEIGHT = 1.0e10


def end_epsilon(max_spans):
    eps = np.float32(100.0) * np.finfo(np.float32).eps
    S = np.float32(max_spans)
    while np.float32(S - eps) == S:
        eps = np.float32(eps * np.float32(10.0))
    return float(eps)


def axis_basis(n, spans, q, eps):
    """(n x (spans + q)) matrix of the uniform B-spline basis of order q at
    p_i = i spans / (n - 1), the end moved inside by eps."""
    p = np.arange(n) * spans / (n - 1)
    p[np.abs(p - spans) <= eps] = spans - eps
    knots = np.arange(-q, spans + q + 1, dtype=float)
    m = spans + q
    B = np.empty((n, m))
    for k in range(m):
        c = np.zeros(m)
        c[k] = 1.0
        B[:, k] = np.nan_to_num(BSpline(knots, c, q, extrapolate=False)(p))
    return B


def mba_dense(field, size, mesh, q, stationary=True):
    """field: (z, y, x, D) or (y, x, D) array; returns the fitted field."""
    D = len(size)
    eps = end_epsilon(max(mesh))
    Bs = [axis_basis(size[d], mesh[d], q, eps) for d in range(D)]
    B = Bs[0]
    for d in range(1, D):
        B = np.kron(Bs[d], B)          # x fastest: index x + nx (y + ny z)
    z = field.reshape(-1, D).copy()
    idx = np.stack(np.unravel_index(np.arange(z.shape[0]), tuple(reversed(size))), 1)[:, ::-1]
    bnd = np.zeros(z.shape[0], bool)
    if stationary:
        for d in range(D):
            bnd |= (idx[:, d] == 0) | (idx[:, d] == size[d] - 1)
    w = np.where(bnd, BOUNDARY_WEIGHT, 1.0)
    z[bnd] = 0.0
    S = (B ** 2).sum(1)
    delta = ((w[:, None] * B ** 3) / S[:, None]).T @ z
    omega = (w[:, None] * B ** 2).sum(0)
    phi = np.where(omega[:, None] > 0, delta / np.where(omega > 0, omega, 1)[:, None], 0.0)
    return (B @ phi).reshape(field.shape)


@pytest.mark.parametrize("q", [1, 2, 3])
@pytest.mark.parametrize("n,spans", [(24, 8), (17, 5), (9, 12), (40, 3)])
def test_basis_matches_scipy(q, n, spans):
    eps = end_epsilon(spans)
    B = axis_basis(n, spans, q, eps)
    for i in range(n):
        s, b = o.bspline_point(i, n, spans, q)
        assert 0 <= s <= spans - 1
        ref = B[i, s:s + q + 1]
        np.testing.assert_allclose(b[:q + 1], ref, rtol=0, atol=1e-14)
        assert abs(B[i].sum() - 1.0) < 1e-12        # partition of unity
        assert abs(b[:q + 1].sum() - 1.0) < 1e-12   # all of it inside the span's taps


def test_end_epsilon_grows_with_the_span_count():
    assert end_epsilon(8) == pytest.approx(100 * np.finfo(np.float32).eps)
    # at 2^20 spans float32 cannot tell S - 1.2e-5 from S: ten times larger
    assert end_epsilon(1 << 20) > end_epsilon(8)
    s, b = o.bspline_point(10, 11, 1 << 20, 3)
    assert s == (1 << 20) - 1


# Meshes leave control points clear of the boundary's 1e10-weight zeros
# (cubic: spans - 4 per axis), so the fits are not all ~0; (13,), (20,) has
# more spans than voxels; the 2-D/3-D cases have a different size and mesh per axis.
@pytest.mark.parametrize("size,mesh,q", [((26,), (8,), 3), ((13,), (20,), 3), ((19,), (4,), 2), ((11,), (3,), 1),
                                         ((14, 11), (6, 5), 3), ((9, 16), (4, 6), 2),
                                         ((12, 10, 9), (6, 5, 5), 3), ((8, 12, 6), (5, 2, 3), 1)])
def test_fit_matches_dense_mba(size, mesh, q):
    D = len(size)
    shape = tuple(reversed(size)) + (D,)
    fld = rng.standard_normal(shape)
    g = o.Grid(list(size), [1.3, 0.7, 2.0][:D], [4.0, -2.0, 1.0][:D])
    got = o.bspline_fit_field(g, fld, mesh, q)
    ref = mba_dense(fld, size, mesh, q)
    assert np.abs(ref).max() > 1e-2
    np.testing.assert_allclose(got, ref, rtol=1e-9, atol=1e-12)


def test_fit_catches_swapped_meshes():
    """Size and mesh differ per axis: the dense MBA with the two meshes
    given to the wrong axes differs from the oracle by far more than the
    tolerance above."""
    size, mesh = (16, 11), (7, 5)
    fld = rng.standard_normal((11, 16, 2))
    got = o.bspline_fit_field(o.Grid(list(size)), fld, mesh, 3)
    wrong = mba_dense(fld, size, mesh[::-1], 3)
    assert np.abs(got - wrong).max() > 1e-2


def test_stationary_boundary_and_no_boundary_differ():
    size, mesh = (16, 13), (7, 6)
    fld = rng.standard_normal((13, 16, 2))
    got = o.bspline_fit_field(o.Grid(list(size)), fld, mesh, 3)
    free = mba_dense(fld, size, mesh, 3, stationary=False)
    assert np.abs(got - free).max() > 1e-2
    # the boundary voxels' data are zero with weight 1e10: their control
    # points are zero to ~1e-10 and so is the fit on the boundary
    edge = np.concatenate([got[0].ravel(), got[-1].ravel(), got[:, 0].ravel(), got[:, -1].ravel()])
    assert np.abs(edge).max() < 1e-8 * np.abs(fld).max()


def test_fit_is_linear_and_mirror_symmetric():
    size, mesh = (15, 12, 9), (4, 3, 2)
    a = rng.standard_normal((9, 12, 15, 3))
    b = rng.standard_normal((9, 12, 15, 3))
    g = o.Grid(list(size))
    fa, fb = o.bspline_fit_field(g, a, mesh, 3), o.bspline_fit_field(g, b, mesh, 3)
    np.testing.assert_allclose(o.bspline_fit_field(g, 2 * a - 0.5 * b, mesh, 3), 2 * fa - 0.5 * fb, atol=1e-12)
    # mirrored along x: the parametrisation is symmetric up to the end epsilon
    fm = o.bspline_fit_field(g, a[:, :, ::-1], mesh, 3)
    np.testing.assert_allclose(fm[:, :, ::-1], fa, atol=1e-4 * np.abs(fa).max())


def test_zero_mesh_is_a_copy():
    fld = rng.standard_normal((6, 7, 8, 3))
    np.testing.assert_array_equal(o.bspline_fit_field(o.Grid([8, 7, 6]), fld, (0, 0, 0), 3), fld)


def test_bspline_syn_stage_registers():
    """BSplineSyN on a small 3-D pair: the energy falls within each level and
    the mesh sets the smoothness (a coarser mesh gives a smoother field)."""
    from ants_amd import phantoms as ph
    pair = ph.phantom_pair_3d(3, size=(24, 28, 24))
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    res = {}
    for mesh in ((5, 5, 5), (9, 11, 9)):
        P = o.SynParams(shrink=[2, 1], sigma=[1.0, 0.0], iterations=[6, 4], learning_rate=0.25, radius=2,
                        model=1, update_mesh=mesh, total_mesh=(0, 0, 0), spline_order=3, force_iterations=True)
        r = o.syn_register(P, g, pair.fixed, g, pair.moving)
        for lv in (0, 1):
            e = r.energy[r.level == lv]
            assert e[-1] < e[0]
        res[mesh] = r
    def rough(f):
        return np.abs(np.diff(f, 2, axis=2)).mean() / np.abs(f).mean()
    assert rough(res[(5, 5, 5)].fwd) < rough(res[(9, 11, 9)].fwd)
    assert np.abs(res[(5, 5, 5)].fwd).max() > 1e-2
