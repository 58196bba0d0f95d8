"""antsApplyTransforms subset on the GPU against the oracle's restatement:
composite affine + displacement-field transforms with directions and
non-dyadic spacings; Linear within 1e-6, NearestNeighbor and GenericLabel
bit-exact (coordinates are formed in double in the same order on both),
BSpline[0..5] and the windowed sincs within 2e-6 of the image range."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as o
from ants_amd import apply_transforms as AT

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(11)


def rot(D, *angles):
    if D == 2:
        a = angles[0]
        return np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    a, b, c = angles
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


def geoms(D):
    gi = AT.Geometry((41, 37, 29)[:D], (1.1, 0.9, 1.3)[:D], (-3.0, 2.0, 1.0)[:D], rot(D, 0.2, -0.1, 0.3))
    gr = AT.Geometry((33, 35, 27)[:D], (1.2, 1.0, 1.4)[:D], (-1.0, 1.5, 2.0)[:D], rot(D, -0.05, 0.15, -0.2))
    gf = AT.Geometry((20, 22, 18)[:D], (2.1, 1.9, 2.2)[:D], (-4.0, 0.0, -1.0)[:D], rot(D, 0.1, 0.0, -0.1))
    return gi, gr, gf


def og(g):
    return o.og_geom(g.size, g.spacing, g.origin, g.direction)


def case(D, interp):
    gi, gr, gf = geoms(D)
    if interp == "Linear":
        img = ndimage.gaussian_filter(rng.standard_normal(gi.shape), 1.5).astype(np.float32)
    else:
        labels = np.array([0, 2, 5, 9, 14], np.float32)
        img = labels[ndimage.zoom(rng.integers(0, 5, tuple(s // 4 + 1 for s in gi.shape)), 4, order=0)]
        img = img[tuple(slice(0, s) for s in gi.shape)].astype(np.float32)
    fld = ndimage.gaussian_filter(rng.standard_normal(gf.shape + (D,)), (3,) * D + (0,)).astype(np.float32)
    fld *= 3.0 / np.abs(fld).max()
    A = np.eye(D) * 1.03 + 0.04 * rng.standard_normal((D, D))
    t, c = 0.8 * rng.standard_normal(D), np.array([4.0, 6.0, 3.0])[:D]
    xf = [AT.DisplacementFieldTransform(fld, gf), AT.AffineTransform(A, t, c), None]
    ox = [("field", og(gf), fld), ("affine", A, t, c), ("identity",)]
    return gi, gr, img, xf, ox


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("interp", ["Linear", "NearestNeighbor", "GenericLabel[Linear]"])
def test_apply_transforms_matches_oracle(D, interp):
    gi, gr, img, xf, ox = case(D, interp)
    out = AT.apply_transforms(gr, img, gi, xf, interp, default_value=-7.0)
    ref = o.apply_transforms(og(gr), og(gi), img, ox, interp.split("[")[0], -7.0)
    assert out.shape == gr.shape
    inside = ref != -7.0
    assert 0.2 < inside.mean() < 1.0
    if interp == "Linear":
        print(f"\n{D}-D linear max |gpu - oracle| = {np.abs(out - ref).max():.3e}")
        np.testing.assert_allclose(out, ref, rtol=0, atol=1e-6)
    else:
        np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 5])
def test_apply_bspline_matches_oracle(D, q):
    """-n BSpline[q]: coefficients by the recursive decomposition (double),
    evaluation by Cox-de Boor weights on the GPU against the oracle's
    truncated-power weights; float output within 2e-6 of the image range
    (order 0 is a pure selection, compared exactly)."""
    gi, gr, img, xf, ox = case(D, "Linear")
    out = AT.apply_transforms(gr, img, gi, xf, f"BSpline[{q}]", default_value=-7.0)
    ref = o.apply_transforms(og(gr), og(gi), img, ox, f"bspline[{q}]", -7.0)
    inside = ref != -7.0
    assert 0.2 < inside.mean() < 1.0
    np.testing.assert_array_equal(out == -7.0, ~inside)
    err = np.abs(out - ref).max()
    print(f"\n{D}-D BSpline[{q}] max |gpu - oracle| = {err:.3e}")
    if q == 0:
        np.testing.assert_array_equal(out, ref)
    else:
        assert err <= 2e-6 * np.abs(img).max()


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("window", ["Cosine", "Hamming", "Lanczos", "Blackman", "Welch"])
def test_apply_windowed_sinc_matches_oracle(D, window):
    """-n <window>WindowedSinc (radius 3, zero boundary) against the oracle:
    the same sample order and products; device and host sin/cos may differ
    by an ulp, so within 2e-6 of the image range."""
    gi, gr, img, xf, ox = case(D, "Linear")
    name = window + "WindowedSinc"
    out = AT.apply_transforms(gr, img, gi, xf, name, default_value=-7.0)
    ref = o.apply_transforms(og(gr), og(gi), img, ox, name.lower(), -7.0)
    inside = ref != -7.0
    assert 0.2 < inside.mean() < 1.0
    np.testing.assert_array_equal(out == -7.0, ~inside)
    err = np.abs(out - ref).max()
    print(f"\n{D}-D {name} max |gpu - oracle| = {err:.3e}")
    assert err <= 2e-6 * np.abs(img).max()


def test_apply_bspline_default_order_and_thin_axis():
    """BSpline without an order is order 3; an axis of length 1 (a single
    slice) and one of length 2 take the decomposition's short-line paths."""
    gi = AT.Geometry((30, 2, 1), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    img = rng.standard_normal(gi.shape).astype(np.float32)
    xf = [AT.AffineTransform(np.eye(3), [0.3, 0.2, 0.0], [0.0, 0.0, 0.0])]
    ox = [("affine", np.eye(3), np.array([0.3, 0.2, 0.0]), np.zeros(3))]
    a = AT.apply_transforms(gi, img, gi, xf, "BSpline")
    b = AT.apply_transforms(gi, img, gi, xf, "BSpline[3]")
    ref = o.apply_transforms(og(gi), og(gi), img, ox, "bspline[3]")
    np.testing.assert_array_equal(a, b)
    np.testing.assert_allclose(a, ref, rtol=0, atol=2e-6 * np.abs(img).max())


@pytest.mark.parametrize("q", [3, 5])
def test_apply_bspline_long_lines(q):
    """Lines of 1100 samples exceed an LDS tile and take the one-thread-per-
    line decomposition; the short axis stays tiled."""
    gi = AT.Geometry((1100, 5), (1.0, 1.0), (0.0, 0.0))
    img = ndimage.gaussian_filter(rng.standard_normal(gi.shape), 2.0).astype(np.float32)
    A = np.array([[0.97, 0.02], [0.0, 1.0]])
    xf = [AT.AffineTransform(A, [3.3, 0.3], [550.0, 2.0])]
    ox = [("affine", A, np.array([3.3, 0.3]), np.array([550.0, 2.0]))]
    out = AT.apply_transforms(gi, img, gi, xf, f"BSpline[{q}]", default_value=-7.0)
    ref = o.apply_transforms(og(gi), og(gi), img, ox, f"bspline[{q}]", -7.0)
    assert (ref != -7.0).mean() > 0.5
    assert np.abs(out - ref).max() <= 2e-6 * np.abs(img).max()


@pytest.mark.parametrize("D", [2, 3])
def test_compose_transforms_to_field_matches_oracle(D):
    """-o [field,1]: T(p) - p through [field, affine, identity] on an oblique
    reference grid, against the oracle (same double coordinates; float
    output within 2e-5 mm); resampling through that one field equals
    resampling through the list (the field is read at its own grid points)."""
    gi, gr, img, xf, ox = case(D, "Linear")
    fld = AT.compose_transforms_to_field(gr, xf)
    ref = o.compose_field(og(gr), ox)
    assert fld.shape == gr.shape + (D,)
    err = np.abs(fld - ref).max()
    print(f"\n{D}-D composite field max |gpu - oracle| = {err:.3e} mm")
    assert err <= 2e-5
    a = AT.apply_transforms(gr, img, gi, xf, "Linear", default_value=-7.0)
    b = AT.apply_transforms(gr, img, gi, [AT.DisplacementFieldTransform(fld, gr)], "Linear", default_value=-7.0)
    both = (a != -7.0) & (b != -7.0)
    # points on the input buffer's edge may fall either side by an ulp
    assert both.mean() > 0.2 and ((a == -7.0) != (b == -7.0)).sum() <= 0.001 * a.size + 2
    np.testing.assert_allclose(a[both], b[both], rtol=0, atol=1e-4 * np.abs(img).max())


@pytest.mark.parametrize("D", [2, 3])
def test_apply_transforms_to_points_matches_oracle(D):
    """antsApplyTransformsToPoints: points through [field, affine, identity]
    against the oracle (double coordinates, within 1e-9 mm); extra columns
    pass through; at the reference grid's points the result is the point
    plus the composite field there."""
    gi, gr, img, xf, ox = case(D, "Linear")
    pts = np.column_stack([rng.uniform(-10, 40, (500, D)), np.arange(500.0), np.full(500, np.nan)])
    got = AT.apply_transforms_to_points(pts, xf, D)
    ref = o.transform_points(pts, ox, D)
    assert got.shape == pts.shape
    np.testing.assert_allclose(got[:, :D], ref, rtol=0, atol=1e-9)
    np.testing.assert_array_equal(got[:, D], pts[:, D])
    assert np.isnan(got[:, D + 1]).all()
    A = np.asarray(gr.direction, float) * np.asarray(gr.spacing, float)
    idx = np.stack(np.meshgrid(*[np.arange(n, dtype=float) for n in reversed(gr.size)], indexing="ij"), -1)[..., ::-1]
    grid_pts = (np.asarray(gr.origin, float) + idx @ A.T).reshape(-1, D)
    fld = AT.compose_transforms_to_field(gr, xf).reshape(-1, D)
    np.testing.assert_allclose(AT.apply_transforms_to_points(grid_pts, xf, D), grid_pts + fld, rtol=0, atol=3e-5)
    assert AT.apply_transforms_to_points(np.zeros((0, D)), xf, D).shape == (0, D)


@pytest.mark.parametrize("D", [2, 3])
def test_jacobian_determinant_matches_oracle(D):
    """CreateJacobianDeterminantImage on the GPU against the oracle on an
    oblique, anisotropic grid with a folding field (clamps exercised): the
    determinant, its log and the gradient matrices within 2e-6 relative
    (double arithmetic, float outputs)."""
    from ants_amd.analysis import create_jacobian_determinant_image
    gi, gr, img, xf, ox = case(D, "Linear")
    fld = ndimage.gaussian_filter(rng.standard_normal(gr.shape + (D,)), (2,) * D + (0,)).astype(np.float32)
    fld *= 6.0 / np.abs(fld).max()
    ogr = og(gr)
    for do_log in (False, True):
        got = create_jacobian_determinant_image(fld, gr, do_log=do_log)
        ref = o.jacobian(ogr, fld, do_log=do_log)
        np.testing.assert_allclose(got, ref, rtol=2e-6, atol=2e-6)
    J = o.jacobian(ogr, fld)
    assert (J == 0).mean() > 0.01 and (J > 0).mean() > 0.5
    G = create_jacobian_determinant_image(fld, gr, deformation_gradient=True)
    np.testing.assert_allclose(G, o.jacobian(ogr, fld, gradient=True), rtol=2e-6, atol=2e-6)
    # useGeometric: the moved-simplex volume ratio (vertex offsets taken
    # relative to the voxel on the GPU, absolute in the oracle: 1e-12 mm
    # apart in the coordinates)
    for do_log in (False, True):
        got = create_jacobian_determinant_image(fld, gr, do_log=do_log, geometric=True)
        ref = o.jacobian(ogr, fld, do_log=do_log, geometric=True)
        np.testing.assert_allclose(got, ref, rtol=1e-5, atol=1e-5)


def test_inverse_affine_round_trip():
    """-t affine then -t [affine,1] maps every point back onto itself."""
    gi, gr, img, _, _ = case(3, "NearestNeighbor")
    A = np.eye(3) * 0.97 + 0.05 * rng.standard_normal((3, 3))
    fwd = AT.AffineTransform(A, [1.0, -2.0, 0.5], [3.0, 2.0, 1.0])
    inv = AT.AffineTransform(A, [1.0, -2.0, 0.5], [3.0, 2.0, 1.0], invert=True)
    out = AT.apply_transforms(gi, img, gi, [fwd, inv], "NearestNeighbor", default_value=-1.0)
    np.testing.assert_array_equal(out, img)


def test_rejects_bad_arguments():
    gi, gr, img, _, _ = case(2, "Linear")
    for bad in ("BSpline[6]", "BSpline[x]", "Gaussian"):
        with pytest.raises(ValueError):
            AT.parse_interpolation(bad)
    bad = AT.Geometry(gi.size, gi.spacing, gi.origin, np.array([[1.0, 0.3], [0.0, 1.0]]))
    with pytest.raises(Exception):
        AT.apply_transforms(gr, img, bad, [], "Linear")


def test_zero_direction_means_identity():
    gi, gr, img, xf, _ = case(3, "Linear")
    g_id = AT.Geometry(gi.size, gi.spacing, gi.origin, np.eye(3))
    g_zero = AT.Geometry(gi.size, gi.spacing, gi.origin, np.zeros((3, 3)))
    a = AT.apply_transforms(g_id, img, g_id, xf[1:2], "Linear")
    b = AT.apply_transforms(g_zero, img, g_zero, xf[1:2], "Linear")
    np.testing.assert_array_equal(a, b)


def test_center_of_mass_and_image_initializers():
    """ants_image_center_of_mass against numpy's intensity-weighted mean
    physical point (oblique, anisotropic grid); -r [fixed,moving,f]
    initializers: an image shifted by whole voxels gives that shift as the
    translation (f = 1), centres (f = 0) and origins (f = 2) by closed form."""
    from ants_amd import _lib as L
    from ants_amd.registration import center_of_mass, initial_transform_from_images
    Geometry = AT.Geometry
    rng = np.random.default_rng(4)
    size = (30, 26, 22)
    sp = (1.1, 0.9, 1.3)
    org = (5.0, -3.0, 2.0)
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    img = rng.uniform(0, 1, (22, 26, 30)).astype(np.float32)
    img[5:15, 6:20, 8:25] += 3.0
    c = center_of_mass(img, size, sp, org, R)
    z, y, x = np.meshgrid(*[np.arange(n) for n in img.shape], indexing="ij")
    idx = np.stack([x, y, z], -1).reshape(-1, 3).astype(np.float64)
    phys = np.asarray(org) + (idx * np.asarray(sp)) @ R.T
    w = img.reshape(-1).astype(np.float64)
    ref = (phys * w[:, None]).sum(0) / w.sum()
    np.testing.assert_allclose(c, ref, rtol=0, atol=1e-9)
    # a blob moved by (3, -2, 1) voxels inside a zero background
    fixed = np.zeros((22, 26, 30), np.float32)
    fixed[6:14, 8:16, 9:19] = rng.uniform(1, 2, (8, 8, 10))
    moving = np.zeros_like(fixed)
    moving[7:15, 6:14, 12:22] = fixed[6:14, 8:16, 9:19]
    g = Geometry(size, sp, org, R)
    t = initial_transform_from_images(fixed, g, moving, g, 1)
    np.testing.assert_allclose(t.translation, R @ (np.array([3, -2, 1]) * np.asarray(sp)), atol=1e-9)
    np.testing.assert_allclose(t.matrix, np.eye(3))
    np.testing.assert_allclose(t.center, center_of_mass(fixed, size, sp, org, R), atol=1e-12)
    gm = Geometry((20, 26, 40), (1.0, 1.0, 1.0), (1.0, 2.0, 3.0), np.eye(3))
    t0 = initial_transform_from_images(fixed, g, np.zeros((40, 26, 20), np.float32), gm, 0)
    cf = np.asarray(org) + R @ (np.asarray(sp) * (np.array(size) - 1) / 2)
    cm = np.array([1.0, 2.0, 3.0]) + (np.array([20, 26, 40]) - 1) / 2
    np.testing.assert_allclose(t0.center, cf, atol=1e-12)
    np.testing.assert_allclose(t0.translation, cm - cf, atol=1e-12)
    t2 = initial_transform_from_images(fixed, g, moving, gm, 2)
    np.testing.assert_allclose(t2.translation, np.array([1.0, 2.0, 3.0]) - np.asarray(org), atol=1e-12)
    with pytest.raises(L.AntsError):
        center_of_mass(np.zeros((4, 4, 4), np.float32), (4, 4, 4))
