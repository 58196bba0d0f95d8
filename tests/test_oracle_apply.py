"""Pins of the oracle's antsApplyTransforms restatement (CPU only):
identity, integer shifts, scipy linear interpolation under general affines
with directions, field/affine equivalence, composition order, half-up
nearest neighbour, and a brute-force GenericLabel (per-label indicator
images interpolated by scipy, argmax with ties to the smaller label)."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as o

rng = np.random.default_rng(7)


def rot3(a, b, c):
    ca, sa, cb, sb, cc, sc = np.cos(a), np.sin(a), np.cos(b), np.sin(b), np.cos(c), np.sin(c)
    rx = np.array([[1, 0, 0], [0, ca, -sa], [0, sa, ca]])
    ry = np.array([[cb, 0, sb], [0, 1, 0], [-sb, 0, cb]])
    rz = np.array([[cc, -sc, 0], [sc, cc, 0], [0, 0, 1]])
    return rz @ ry @ rx


def points(size, sp, org, dirm):
    """physical points of every voxel, numpy (z,y,x) order, last axis = (x,y,z)"""
    D = len(size)
    idx = np.stack(np.meshgrid(*[np.arange(n, dtype=np.float64) for n in reversed(size)], indexing="ij"), -1)[..., ::-1]
    return np.asarray(org, float) + idx @ (np.asarray(dirm, float) * np.asarray(sp, float)).T


def cidx(p, size, sp, org, dirm):
    return ((p - np.asarray(org, float)) @ np.asarray(dirm, float)) / np.asarray(sp, float)


def inside(c, size):
    return np.all((c >= -0.5) & (c < np.asarray(size, float) - 0.5), axis=-1)


def scipy_linear(img, c):
    coords = [c[..., d] for d in reversed(range(c.shape[-1]))]
    return ndimage.map_coordinates(img.astype(np.float64), coords, order=1, mode="nearest")


@pytest.mark.parametrize("interp", ["linear", "nearestneighbor", "genericlabel"])
def test_identity_copies(interp):
    """Rounding modes copy exactly; linear copies up to the index -> point ->
    index round trip of non-dyadic spacings (1e-16 relative in the index)."""
    img = rng.integers(0, 5, (6, 7, 8)).astype(np.float32) + 0.5 * rng.standard_normal((6, 7, 8)).astype(np.float32)
    if interp != "linear":
        img = np.round(img)
    g = o.og_geom((8, 7, 6), (1.1, 0.9, 2.0), (3, -2, 1))
    for tl in ([], [("identity",)]):
        out = o.apply_transforms(g, g, img, tl, interp)
        if interp == "linear":
            np.testing.assert_allclose(out, img, rtol=0, atol=1e-6)
        else:
            np.testing.assert_array_equal(out, img)


def test_integer_translation_shifts_with_default():
    img = rng.standard_normal((5, 6, 7)).astype(np.float32)
    g = o.og_geom((7, 6, 5), (2.0, 1.0, 0.5))
    out = o.apply_transforms(g, g, img, [("affine", np.eye(3), [4.0, -1.0, 0.5], [0, 0, 0])], "linear", -9.0)
    ref = np.full_like(img, -9.0)
    ref[:4, 1:, :5] = img[1:, :5, 2:]   # x + 2 voxels, y - 1, z + 1
    np.testing.assert_array_equal(out, ref)


@pytest.mark.parametrize("D", [2, 3])
def test_general_affine_with_directions_matches_scipy(D):
    size_in = (23, 19, 17)[:D]
    size_ref = (20, 21, 15)[:D]
    img = ndimage.gaussian_filter(rng.standard_normal(tuple(reversed(size_in))), 1.5).astype(np.float32)
    R_in = rot3(0.2, -0.1, 0.3)[:D, :D] if D == 3 else rot3(0, 0, 0.25)[:2, :2]
    R_ref = rot3(-0.05, 0.15, -0.2)[:D, :D] if D == 3 else rot3(0, 0, -0.1)[:2, :2]
    sp_in, org_in = (1.1, 0.9, 1.3)[:D], (-3.0, 2.0, 1.0)[:D]
    sp_ref, org_ref = (1.2, 1.0, 1.4)[:D], (-1.0, 1.0, 2.0)[:D]
    A = np.eye(D) * 1.05 + 0.03 * rng.standard_normal((D, D))
    t, c = 0.7 * rng.standard_normal(D), np.array([5.0, 6.0, 4.0])[:D]
    gi, gr = o.og_geom(size_in, sp_in, org_in, R_in), o.og_geom(size_ref, sp_ref, org_ref, R_ref)
    out = o.apply_transforms(gr, gi, img, [("affine", A, t, c)], "linear", -5.0)
    p = points(size_ref, sp_ref, org_ref, R_ref)
    q = (p - c) @ A.T + t + c
    ci = cidx(q, size_in, sp_in, org_in, R_in)
    m = inside(ci, size_in)
    assert m.mean() > 0.2
    ref = np.where(m, scipy_linear(img, ci), -5.0)
    np.testing.assert_allclose(out, ref, atol=2e-6)


def test_constant_field_equals_translation_and_order_matters():
    img = ndimage.gaussian_filter(rng.standard_normal((12, 14, 16)), 1.0).astype(np.float32)
    g = o.og_geom((16, 14, 12), (1.0, 1.5, 0.8), (2.0, -1.0, 0.5))
    u = np.array([1.3, -0.4, 0.7])
    fld = np.broadcast_to(u, (12, 14, 16, 3)).astype(np.float32)
    by_field = o.apply_transforms(g, g, img, [("field", g, fld)], "linear")
    by_affine = o.apply_transforms(g, g, img, [("affine", np.eye(3), u.astype(np.float32), [0, 0, 0])], "linear")
    np.testing.assert_array_equal(by_field, by_affine)
    # composition: [field, affine] maps p -> A(p + u); [affine, field] maps p -> A p + u
    A = np.diag([1.1, 0.95, 1.05])
    p = points((16, 14, 12), (1.0, 1.5, 0.8), (2.0, -1.0, 0.5), np.eye(3))
    for order, q in ((["field", "affine"], (p + u) @ A.T), (["affine", "field"], p @ A.T + u)):
        tl = [("field", g, fld) if k == "field" else ("affine", A, [0, 0, 0], [0, 0, 0]) for k in order]
        out = o.apply_transforms(g, g, img, tl, "linear", 0.0)
        ci = cidx(q, (16, 14, 12), (1.0, 1.5, 0.8), (2.0, -1.0, 0.5), np.eye(3))
        m = inside(ci, (16, 14, 12))
        # the field is constant inside its buffer only: keep points whose field lookup is inside
        pf = p if order[0] == "field" else p @ A.T
        mf = inside(cidx(pf, (16, 14, 12), (1.0, 1.5, 0.8), (2.0, -1.0, 0.5), np.eye(3)), (16, 14, 12))
        sel = m & mf
        np.testing.assert_allclose(out[sel], scipy_linear(img, ci)[sel], atol=2e-6)


def test_nearest_neighbour_rounds_half_up():
    img = np.arange(4 * 5 * 6, dtype=np.float32).reshape(4, 5, 6)
    g = o.og_geom((6, 5, 4))
    out = o.apply_transforms(g, g, img, [("affine", np.eye(3), [0.5, 0.0, 0.0], [0, 0, 0])], "nearestneighbor", -1)
    ref = np.full_like(img, -1.0)
    ref[..., :5] = img[..., 1:]        # x + 0.5 rounds up to x + 1
    np.testing.assert_array_equal(out, ref)


def test_direction_permutation_transposes():
    img = rng.standard_normal((3, 7, 7)).astype(np.float32)
    swap = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 1]], float)
    gi, gr = o.og_geom((7, 7, 3)), o.og_geom((7, 7, 3), direction=swap)
    out = o.apply_transforms(gr, gi, img, [], "nearestneighbor")
    np.testing.assert_array_equal(out, np.transpose(img, (0, 2, 1)))


def test_generic_label_matches_brute_force():
    labels = np.array([0, 3, 7, 11], np.float32)
    coarse = rng.integers(0, 4, (5, 6, 7))
    img = labels[ndimage.zoom(coarse, 3, order=0)][:14, :16, :18].astype(np.float32)
    size = (18, 16, 14)
    g = o.og_geom(size, (1.0, 1.1, 0.9))
    A = rot3(0.1, 0.05, -0.12) * 1.02
    t, c = np.array([0.6, -0.3, 0.4]), np.array([9.0, 8.0, 6.0])
    out = o.apply_transforms(g, g, img, [("affine", A, t, c)], "genericlabel", -1)
    p = points(size, (1.0, 1.1, 0.9), (0, 0, 0), np.eye(3))
    ci = cidx((p - c) @ A.T + t + c, size, (1.0, 1.1, 0.9), (0, 0, 0), np.eye(3))
    m = inside(ci, size)
    vals = np.stack([scipy_linear((img == l).astype(np.float64), ci) for l in labels], -1)
    ref = np.where(m, labels[np.argmax(vals, -1)], -1.0)   # argmax: first (smallest) label on ties
    srt = np.sort(vals, -1)
    clear = (srt[..., -1] - srt[..., -2] > 1e-9) | ~m
    assert clear.mean() > 0.95
    np.testing.assert_array_equal(out[clear], ref[clear])
    # exact ties (a point midway between two labels) go to the smaller label
    two = np.zeros((1, 1, 2), np.float32)
    two[..., 1] = 5.0
    g2 = o.og_geom((2, 1, 1))
    half = o.apply_transforms(g2, g2, two, [("affine", np.eye(3), [0.5, 0, 0], [0, 0, 0])], "genericlabel", -1)
    assert half[0, 0, 0] == 0.0


def _bspline_case(D, seed):
    r = np.random.default_rng(seed)
    size = (23, 19, 17)[:D]
    img = ndimage.gaussian_filter(r.standard_normal(tuple(reversed(size))), 1.0).astype(np.float32)
    g = o.og_geom(size, (1.0,) * D, (0.0,) * D)
    return r, size, img, g


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("q", [0, 1, 2, 3, 4, 5])
def test_bspline_matches_scipy_mirror_spline(D, q):
    """-n BSpline[q] through an affine with rotation and scaling equals
    scipy's map_coordinates(order=q, mode='mirror') -- an independent
    implementation of the same mirror-boundary B-spline interpolation -- at
    every point inside the buffer, edges included; outside gets the default."""
    r, size, img, g = _bspline_case(D, 40 + q)
    A = np.eye(D) * 1.05 + 0.06 * r.standard_normal((D, D))
    t = 1.5 + 0.5 * r.standard_normal(D)
    c = np.asarray(size, float) / 2
    out = o.apply_transforms(g, g, img, [("affine", A, t, c)], f"bspline[{q}]", -5.0)
    pts = points(size, (1.0,) * D, (0.0,) * D, np.eye(D))
    mapped = (pts - c) @ A.T + t + c
    ci = cidx(mapped, size, (1.0,) * D, (0.0,) * D, np.eye(D))
    ins = inside(ci, size)
    coords = [ci[..., d] for d in reversed(range(D))]
    ref = ndimage.map_coordinates(img.astype(np.float64), coords, order=q, mode="mirror")
    assert 0.3 < ins.mean() < 1.0
    if q == 0:
        # nearest sample, half-integers rounding up (scipy may round them
        # the other way, so compare away from ties)
        tie = np.any(np.abs(ci - np.floor(ci) - 0.5) < 1e-9, axis=-1)
        m = ins & ~tie
        np.testing.assert_array_equal(out[m], ref[m].astype(np.float32))
    else:
        np.testing.assert_allclose(out[ins], ref[ins], rtol=0, atol=2e-6 * np.abs(img).max())
    assert np.all(out[~ins] == -5.0)


@pytest.mark.parametrize("q", [2, 3, 4, 5])
def test_bspline_interpolates_samples(q):
    """At voxel centres the spline passes through the samples, whatever the
    order (a wrong gain or pole leaves a smoothed image)."""
    _, size, img, g = _bspline_case(3, 7)
    out = o.apply_transforms(g, g, img, [], f"bspline[{q}]")
    np.testing.assert_allclose(out, img, rtol=0, atol=1e-6 * np.abs(img).max())


def test_bspline_reproduces_polynomials():
    """A cubic spline reproduces polynomials of degree <= 3 away from the
    mirror boundary exactly, so a quadratic ramp is interpolated exactly in
    the interior at any sub-voxel shift."""
    size = (40, 36)
    y, x = np.meshgrid(np.arange(36.0), np.arange(40.0), indexing="ij")
    img = (0.01 * (x - 20) ** 2 - 0.02 * (y - 15) * (x - 20) + 0.3 * y).astype(np.float32)
    g = o.og_geom(size, (1.0, 1.0), (0.0, 0.0))
    shift = np.array([0.37, -0.61])
    out = o.apply_transforms(g, g, img, [("affine", np.eye(2), shift, np.zeros(2))], "bspline[3]")
    xs, ys = x + shift[0], y + shift[1]
    want = 0.01 * (xs - 20) ** 2 - 0.02 * (ys - 15) * (xs - 20) + 0.3 * ys
    core = (slice(8, -8), slice(8, -8))
    np.testing.assert_allclose(out[core], want[core], rtol=0, atol=2e-5)


def test_compose_field_affine_closed_form():
    """-o [field,1] of an affine: A (p - c) + t + c - p at every reference
    point (oblique, anisotropic reference grid)."""
    r = np.random.default_rng(5)
    size, sp, org, R = (9, 8, 7), (1.1, 0.9, 1.3), (2.0, -1.0, 0.5), rot3(0.2, -0.1, 0.3)
    A = np.eye(3) + 0.05 * r.standard_normal((3, 3))
    t, c = r.standard_normal(3), np.array([3.0, 1.0, -2.0])
    g = o.og_geom(size, sp, org, R)
    out = o.compose_field(g, [("affine", A, t, c)])
    p = points(size, sp, org, R)
    want = (p - c) @ A.T + t + c - p
    np.testing.assert_allclose(out, want, rtol=0, atol=2e-6)
    assert out.shape == (7, 8, 9, 3)
    np.testing.assert_array_equal(o.compose_field(g, []), 0.0)


def test_compose_field_displacement_and_order():
    """A field on the reference grid itself comes back unchanged (its samples
    are read at their own points); [field, affine] is the affine applied to
    p + u(p); points the field grid does not cover pass unchanged."""
    r = np.random.default_rng(6)
    size = (10, 9, 8)
    g = o.og_geom(size, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0))
    u = (0.3 * r.standard_normal((8, 9, 10, 3))).astype(np.float32)
    np.testing.assert_allclose(o.compose_field(g, [("field", g, u)]), u, rtol=0, atol=1e-6)
    A = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])
    t, c = np.array([0.5, 0.0, -0.25]), np.zeros(3)
    out = o.compose_field(g, [("field", g, u), ("affine", A, t, c)])
    p = points(size, (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), np.eye(3))
    q = p + u
    np.testing.assert_allclose(out, q @ A.T + t - p, rtol=0, atol=2e-6)
    # a field grid shifted far away covers no reference point
    gfar = o.og_geom(size, (1.0, 1.0, 1.0), (100.0, 0.0, 0.0))
    np.testing.assert_array_equal(o.compose_field(g, [("field", gfar, u)]), 0.0)


_WINDOWS = {
    "cosine": lambda x: np.cos(np.pi * x / 6.0),
    "hamming": lambda x: 0.54 + 0.46 * np.cos(np.pi * x / 3.0),
    "lanczos": lambda x: np.sinc(x / 3.0),
    "blackman": lambda x: 0.42 + 0.5 * np.cos(np.pi * x / 3.0) + 0.08 * np.cos(2 * np.pi * x / 3.0),
    "welch": lambda x: 1.0 - x * x / 9.0,
}


def _sinc_brute(img, c, window):
    """sum over the 6^D samples floor(c)-2 .. floor(c)+3 of
    img * prod_d window(x_d) sinc(x_d), zero outside, delta when whole."""
    D = img.ndim
    out = 0.0
    fl = np.floor(c)
    axes = []
    for d in range(D):
        js = fl[d] - 2 + np.arange(6)
        x = c[d] - js
        w = np.where(c[d] == fl[d], (js == fl[d]).astype(float), _WINDOWS[window](x) * np.sinc(x))
        axes.append((js.astype(int), w))
    for idx in np.ndindex(*(6,) * D):
        j = [axes[d][0][idx[d]] for d in range(D)]
        w = np.prod([axes[d][1][idx[d]] for d in range(D)])
        inside = all(0 <= j[d] < img.shape[D - 1 - d] for d in range(D))
        if inside:
            out += img[tuple(reversed(j))] * w
    return out


@pytest.mark.parametrize("window", list(_WINDOWS))
@pytest.mark.parametrize("D", [2, 3])
def test_windowed_sinc_matches_definition(window, D):
    """The oracle's windowed sinc against a direct numpy evaluation of the
    definition (np.sinc) at points over the whole buffer, edges included,
    through a rotation and shift."""
    r = np.random.default_rng(60 + D)
    size = (13, 11, 9)[:D]
    img = r.standard_normal(tuple(reversed(size))).astype(np.float32)
    g = o.og_geom(size, (1.0,) * D, (0.0,) * D)
    A = np.eye(D) * 0.97 + 0.05 * r.standard_normal((D, D))
    t, c = 0.4 * r.standard_normal(D), np.asarray(size, float) / 2
    out = o.apply_transforms(g, g, img, [("affine", A, t, c)], window + "windowedsinc", -9.0)
    pts = points(size, (1.0,) * D, (0.0,) * D, np.eye(D))
    ci = (pts - c) @ A.T + t + c
    ins = inside(ci, size)
    assert 0.5 < ins.mean() < 1.0
    flat_c, flat_o = ci.reshape(-1, D), out.reshape(-1)
    for k in np.flatnonzero(ins.reshape(-1))[:: max(1, ins.sum() // 60)]:
        assert flat_o[k] == pytest.approx(_sinc_brute(img.astype(np.float64), flat_c[k], window), abs=2e-6)
    assert np.all(out[~ins] == -9.0)


@pytest.mark.parametrize("window", list(_WINDOWS))
def test_windowed_sinc_samples_and_bandlimited_signal(window):
    """Whole indices return the samples; a slowly varying cosine is
    reconstructed at half-voxel shifts in the interior to within the
    window's ripple (a weight set one sample off moves it by ~0.3)."""
    size = (48, 40)
    y, x = np.meshgrid(np.arange(40.0), np.arange(48.0), indexing="ij")
    f = lambda xx, yy: np.cos(2 * np.pi * 0.05 * xx) * np.cos(2 * np.pi * 0.04 * yy + 0.3)
    img = f(x, y).astype(np.float32)
    g = o.og_geom(size, (1.0, 1.0), (0.0, 0.0))
    name = window + "windowedsinc"
    whole = o.apply_transforms(g, g, img, [("affine", np.eye(2), np.array([2.0, -1.0]), np.zeros(2))], name, 0.0)
    np.testing.assert_array_equal(whole[1:, :-2], img[:-1, 2:])  # out(x, y) = img(x + 2, y - 1)
    half = o.apply_transforms(g, g, img, [("affine", np.eye(2), np.array([0.5, 0.5]), np.zeros(2))], name, 0.0)
    core = (slice(6, -6), slice(6, -6))
    err = np.abs(half - f(x + 0.5, y + 0.5))[core].max()
    print(f"\n{window}: band-limited reconstruction error {err:.4f}")
    assert err < 0.05


def test_transform_points_closed_forms():
    """Points through [field, affine]: a constant field inside its grid is a
    translation, points outside it pass unchanged, then the affine
    A (p - c) + t + c; at grid points the result equals the composite
    field's T(p) - p plus p."""
    g = o.og_geom((10, 9, 8), (1.5, 1.0, 2.0), (1.0, -2.0, 0.5))
    u = np.zeros((8, 9, 10, 3), np.float32)
    u[...] = [0.5, -0.25, 1.0]
    A = np.array([[0.9, 0.1, 0.0], [-0.1, 1.1, 0.05], [0.0, 0.02, 1.0]])
    t, c = np.array([1.0, 2.0, -1.0]), np.array([5.0, 3.0, 4.0])
    xs = [("field", g, u), ("affine", A, t, c)]
    inside_pts = np.array([[3.0, 1.0, 4.0], [10.0, 3.0, 10.0]])
    outside_pts = np.array([[-50.0, 0.0, 0.0]])
    got = o.transform_points(np.vstack([inside_pts, outside_pts]), xs)
    q = np.vstack([inside_pts + [0.5, -0.25, 1.0], outside_pts])
    np.testing.assert_allclose(got, (q - c) @ A.T + t + c, rtol=0, atol=1e-12)
    p = points((10, 9, 8), (1.5, 1.0, 2.0), (1.0, -2.0, 0.5), np.eye(3)).reshape(-1, 3)
    fld = o.compose_field(g, xs).reshape(-1, 3)
    np.testing.assert_allclose(o.transform_points(p, xs), p + fld, rtol=0, atol=2e-6)
    np.testing.assert_array_equal(o.transform_points(p[:0], xs).shape, (0, 3))


def _phys(size, sp, org, R):
    return points(size, sp, org, R)


@pytest.mark.parametrize("D", [2, 3])
def test_jacobian_linear_and_cubic_fields(D):
    """CreateJacobianDeterminantImage: u(p) = (A - I) p on an oblique,
    anisotropic grid gives det A at every voxel two or more from the faces
    (the five-point stencil is exact there; a missing direction rotation or a
    transposed one fails on the oblique grid); a cubic field gives its
    analytic Jacobian in the interior; the gradient output is I + du/dp."""
    r = np.random.default_rng(20 + D)
    size = (14, 12, 10)[:D]
    sp = (1.3, 0.8, 1.1)[:D]
    org = (2.0, -1.0, 3.0)[:D]
    R = rot3(0.3, -0.2, 0.4) if D == 3 else np.array([[np.cos(0.5), -np.sin(0.5)], [np.sin(0.5), np.cos(0.5)]])
    g = o.og_geom(size, sp, org, R)
    p = _phys(size, sp, org, R)
    A = np.eye(D) + 0.1 * r.standard_normal((D, D))
    u = (p @ (A - np.eye(D)).T).astype(np.float64)
    core = (slice(2, -2),) * D
    J = o.jacobian(g, u.astype(np.float32))
    np.testing.assert_allclose(J[core], np.linalg.det(A), rtol=0, atol=2e-5)
    G = o.jacobian(g, u.astype(np.float32), gradient=True)
    np.testing.assert_allclose(G[core], np.broadcast_to(A, G[core].shape), rtol=0, atol=2e-5)
    # cubic: u_0 = a p_1^3, derivative 3 a p_1^2 along axis 1
    a = 2e-4
    u3 = np.zeros_like(u)
    u3[..., 0] = a * p[..., 1] ** 3
    G3 = o.jacobian(g, u3.astype(np.float32), gradient=True)
    want = np.zeros(G3.shape)
    want[..., 0, 1] = 3 * a * p[..., 1] ** 2
    want += np.eye(D)
    np.testing.assert_allclose(G3[core], want[core], rtol=0, atol=2e-4 * np.abs(want).max())


def test_jacobian_edge_stencil_and_clamps():
    """At x = 0 with u_x = a x (unit spacing) the clamped stencil gives
    (-2a + 8a - 8*0 + 0) / 12 = a/2; folding fields clamp to 0, the log to
    log(0.001)."""
    g = o.og_geom((8, 6), (1.0, 1.0), (0.0, 0.0))
    u = np.zeros((6, 8, 2), np.float32)
    a = 0.25                                     # exact in float
    u[..., 0] = a * np.arange(8.0)
    G = o.jacobian(g, u, gradient=True)
    assert G[3, 0, 0, 0] == pytest.approx(1 + a / 2, abs=1e-12)
    assert G[3, 1, 0, 0] == pytest.approx(1 + (-3 * a + 8 * 2 * a - 8 * 0 + 0) / 12, abs=1e-12)
    assert G[3, 4, 0, 0] == pytest.approx(1 + a, abs=1e-12)
    u[..., 0] = -2.0 * np.arange(8.0)            # J = 1 - 2 < 0 in the interior
    assert o.jacobian(g, u)[3, 4] == 0.0
    assert o.jacobian(g, u, do_log=True)[3, 4] == pytest.approx(np.log(0.001))
    u[..., 0] = 0.5 * np.arange(8.0)             # J = 1.5
    assert o.jacobian(g, u, do_log=True)[3, 4] == pytest.approx(np.log(1.5), abs=1e-12)


@pytest.mark.parametrize("D", [2, 3])
def test_jacobian_geometric_closed_forms(D):
    """useGeometric: zero and uniform fields give 1 everywhere; a linear
    field (A - I) p gives |det A| wherever the simplex's interpolation
    corners stay inside (one voxel from the faces) on an oblique grid; on a
    smooth field it agrees with the gradient-tensor form to first order."""
    r = np.random.default_rng(30 + D)
    size = (12, 11, 10)[:D]
    sp = (1.2, 0.9, 1.1)[:D]
    org = (1.0, 2.0, -1.0)[:D]
    R = rot3(0.2, 0.3, -0.1) if D == 3 else np.array([[np.cos(0.4), -np.sin(0.4)], [np.sin(0.4), np.cos(0.4)]])
    g = o.og_geom(size, sp, org, R)
    shape = tuple(reversed(size)) + (D,)
    np.testing.assert_allclose(o.jacobian(g, np.zeros(shape, np.float32), geometric=True), 1.0, atol=1e-12)
    np.testing.assert_allclose(o.jacobian(g, np.full(shape, 0.7, np.float32), geometric=True), 1.0, atol=1e-6)
    p = points(size, sp, org, R)
    A = np.eye(D) + 0.08 * r.standard_normal((D, D))
    u = (p @ (A - np.eye(D)).T).astype(np.float32)
    core = (slice(1, -1),) * D
    Jg = o.jacobian(g, u, geometric=True)
    np.testing.assert_allclose(Jg[core], abs(np.linalg.det(A)), rtol=0, atol=2e-5)
    assert o.jacobian(g, u, geometric=True, do_log=True)[core] == pytest.approx(np.log(abs(np.linalg.det(A))), abs=2e-5)
    x = p[..., 0] / 10.0
    us = np.zeros(shape, np.float32)
    us[..., 0] = 0.3 * np.sin(x)
    Jt = o.jacobian(g, us)
    np.testing.assert_allclose(Jg[core] * 0 + o.jacobian(g, us, geometric=True)[core], Jt[core], rtol=0, atol=3e-3)
