"""Pins of the CPU oracle (oracle/syn_oracle.cpp) against closed forms,
scipy and brute force.  CPU only.

The reference holds no golden vectors for this path (its SyN regression data
is remote and uses Mattes, its CC tests have the metric checks commented out),
so each oracle function is pinned here by mathematics chosen so that a
dropped term, a wrong sign or index, or swapped roles fail a check.
"""
import numpy as np
import pytest
from scipy import ndimage, special

import oracle as o

rng = np.random.default_rng(1234)


# ---------------------------------------------------------------- Gaussian operator
@pytest.mark.parametrize("n", range(0, 9))
@pytest.mark.parametrize("x", [0.25, 1.0, 3.0, 3.75, 4.0, 6.5])
def test_bessel_matches_scipy(n, x):
    assert o.bessel_in(n, x) * np.exp(-x) == pytest.approx(special.ive(n, x), rel=2e-6, abs=1e-12)


@pytest.mark.parametrize("t,eps,radius", [(3.0, 0.001, 6), (4.0, 0.01, 5), (1.0, 0.01, 3), (0.5, 0.001, 3)])
def test_gaussian_kernel_radius_and_values(t, eps, radius):
    k = o.gaussian_kernel(t, eps, 1000)
    assert len(k) - 1 == radius
    ref = special.ive(np.arange(radius + 1), t)
    ref /= ref[0] + 2 * ref[1:].sum()
    np.testing.assert_allclose(k, ref, rtol=1e-6, atol=1e-9)
    assert k[0] + 2 * k[1:].sum() == pytest.approx(1.0, abs=1e-12)


def test_gaussian_kernel_tabulated_values():
    # e^-3 I_n(3) normalised, ITK GaussianOperator(variance 3, MaxError 0.001)
    k = o.gaussian_kernel(3.0, 0.001, 1000)
    np.testing.assert_allclose(k[:4], [0.243132, 0.196933, 0.111843, 0.047809], atol=2e-6)


def test_gaussian_kernels_match_golden_file():
    """tests/golden/gaussian_kernels.json (written by make_gaussian_kernels.py
    from scipy's e^-t I_n(t)): the oracle's kernels have the same radius and
    values for every variance / maximum error the stage uses."""
    import json
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "gaussian_kernels.json")
    cases = json.load(open(path))["kernels"]
    assert len(cases) >= 6
    for c in cases:
        k = o.gaussian_kernel(c["variance"], c["max_error"], 1000)
        assert len(k) - 1 == c["radius"], c
        np.testing.assert_allclose(k, c["half_kernel"], rtol=1e-6, atol=1e-9)


def test_gaussian_kernel_width_cap_and_zero():
    assert len(o.gaussian_kernel(0.0, 0.01, 32)) == 2  # identity: [1, 0]
    assert o.gaussian_kernel(0.0, 0.01, 32)[0] == 1.0
    k = o.gaussian_kernel(50.0, 0.001, 5)  # truncated at the max width
    assert len(k) == 6


# ---------------------------------------------------------------- smoothing
def _sep_correlate(img, k_axes):
    out = img.astype(np.float64)
    for ax_xyz, k in k_axes:
        full = np.concatenate([k[:0:-1], k])
        out = ndimage.correlate1d(out, full, axis=out.ndim - 1 - ax_xyz, mode="nearest")
    return out


def test_smooth_image_matches_scipy_3d():
    img = rng.standard_normal((9, 13, 17))
    g = o.Grid([17, 13, 9], [1.0, 1.0, 1.0])
    out = o.smooth_image(g, img, 2.0)
    k = o.gaussian_kernel(4.0, 0.01, 32)
    np.testing.assert_allclose(out, _sep_correlate(img, [(0, k), (1, k), (2, k)]), atol=1e-12)


def test_smooth_image_mm_uses_spacing():
    img = rng.standard_normal((11, 12))
    g = o.Grid([12, 11], [0.5, 2.0])
    out = o.smooth_image(g, img, 1.0, in_mm=True)
    kx = o.gaussian_kernel(1.0 / 0.25, 0.01, 32)
    ky = o.gaussian_kernel(1.0 / 4.0, 0.01, 32)
    np.testing.assert_allclose(out, _sep_correlate(img, [(0, kx), (1, ky)]), atol=1e-12)
    np.testing.assert_array_equal(o.smooth_image(g, img, 0.0), img)


def test_smooth_field_boundary_interior_and_blend():
    D = 3
    f = rng.standard_normal((8, 10, 12, D))
    g = o.Grid([12, 10, 8])
    out = o.smooth_field(g, f, 3.0)
    k = o.gaussian_kernel(3.0, 0.001, 1000)
    ref = np.stack([_sep_correlate(f[..., c], [(0, k), (1, k), (2, k)]) for c in range(D)], -1)
    inner = (slice(1, -1),) * 3
    np.testing.assert_allclose(out[inner], ref[inner], atol=1e-12)
    bmask = np.ones((8, 10, 12), bool)
    bmask[inner] = False
    assert np.all(out[bmask] == 0)
    # variance < 0.5 blends w1*smooth + (1-w1)*input, w1 = 1 - var/0.5
    out2 = o.smooth_field(g, f, 0.25)
    k2 = o.gaussian_kernel(0.25, 0.001, 1000)
    sm2 = np.stack([_sep_correlate(f[..., c], [(0, k2), (1, k2), (2, k2)]) for c in range(D)], -1)
    np.testing.assert_allclose(out2[inner], 0.5 * sm2[inner] + 0.5 * f[inner], atol=1e-12)
    # variance <= 0: unmodified copy, boundary kept
    np.testing.assert_array_equal(o.smooth_field(g, f, 0.0), f)


def test_smooth_impulse_response_2d():
    img = np.zeros((31, 31))
    img[15, 15] = 1.0
    g = o.Grid([31, 31])
    out = o.smooth_image(g, img, 1.0)
    k = o.gaussian_kernel(1.0, 0.01, 32)
    full = np.concatenate([k[:0:-1], k])
    np.testing.assert_allclose(out[15 - 3:15 + 4, 15 - 3:15 + 4], np.outer(full, full), atol=1e-15)
    assert out.sum() == pytest.approx(1.0, abs=1e-12)


# ---------------------------------------------------------------- interpolation / warp
def test_warp_matches_map_coordinates():
    img = rng.standard_normal((7, 9, 11))
    gi = o.Grid([11, 9, 7], [1.0, 1.0, 1.0], [0.0, 0.0, 0.0])
    go = o.Grid([11, 9, 7])
    u = rng.uniform(-2.5, 2.5, (7, 9, 11, 3))
    out = o.warp_image(gi, img, go, u)
    zz, yy, xx = np.meshgrid(np.arange(7), np.arange(9), np.arange(11), indexing="ij")
    cx, cy, cz = xx + u[..., 0], yy + u[..., 1], zz + u[..., 2]
    inside = (cx >= -0.5) & (cx < 10.5) & (cy >= -0.5) & (cy < 8.5) & (cz >= -0.5) & (cz < 6.5)
    ref = ndimage.map_coordinates(img, [cz, cy, cx], order=1, mode="nearest")
    np.testing.assert_allclose(out[inside], ref[inside], atol=1e-12)
    assert np.all(out[~inside] == 0)
    # half-open bounds: -0.5 is inside, n - 0.5 is outside
    u0 = np.zeros((7, 9, 11, 3))
    u0[0, 0, 0, 0] = -0.5
    u0[0, 0, 10, 0] = 0.5
    out0 = o.warp_image(gi, img, go, u0)
    assert out0[0, 0, 0] == img[0, 0, 0]
    assert out0[0, 0, 10] == 0.0


def test_warp_through_shrunk_grid_physical_points():
    img = rng.standard_normal((32, 32))
    gF = o.Grid([32, 32])
    gl = o.shrink_grid(gF, 4)
    out = o.warp_image(gF, img, gl, np.zeros(gl.shape + (2,)))
    # level voxel i sits at physical 1.5 + 4 i: half way between full-res voxels
    ref = ndimage.map_coordinates(img, np.meshgrid(1.5 + 4 * np.arange(8), 1.5 + 4 * np.arange(8), indexing="ij"),
                                  order=1)
    np.testing.assert_allclose(out, ref, atol=1e-12)


def test_compose_constant_and_affine():
    g = o.Grid([12, 10])
    d = np.broadcast_to(np.array([0.7, -0.4]), (10, 12, 2)).copy()
    w = np.broadcast_to(np.array([-1.25, 0.5]), (10, 12, 2)).copy()
    out = o.compose(g, d, w)
    yy, xx = np.meshgrid(np.arange(10), np.arange(12), indexing="ij")
    inside = (xx - 1.25 >= -0.5) & (xx - 1.25 < 11.5) & (yy + 0.5 >= -0.5) & (yy + 0.5 < 9.5)
    np.testing.assert_allclose(out[inside], (d + w)[inside], atol=1e-12)
    np.testing.assert_allclose(out[~inside], w[~inside], atol=1e-12)
    # affine displacement d(x) = A x + b composes exactly away from the edges
    A = np.array([[0.05, -0.02], [0.03, 0.01]])
    b = np.array([0.2, -0.1])
    dd = np.stack([A[0, 0] * xx + A[0, 1] * yy + b[0], A[1, 0] * xx + A[1, 1] * yy + b[1]], -1)
    ww = rng.uniform(-0.9, 0.9, (10, 12, 2))
    out = o.compose(g, dd, ww)
    px, py = xx + ww[..., 0], yy + ww[..., 1]
    ref = ww + np.stack([A[0, 0] * px + A[0, 1] * py + b[0], A[1, 0] * px + A[1, 1] * py + b[1]], -1)
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(out[inner], ref[inner], atol=1e-12)


def test_resample_field_linear_and_outside():
    gi = o.Grid([8, 9, 10], [4.0, 4.0, 4.0], [1.5, 1.5, 1.5])
    go = o.Grid([16, 18, 20], [2.0, 2.0, 2.0], [0.5, 0.5, 0.5])
    zz, yy, xx = np.meshgrid(1.5 + 4 * np.arange(10), 1.5 + 4 * np.arange(9), 1.5 + 4 * np.arange(8), indexing="ij")
    f = np.stack([0.1 * xx - 0.2 * yy, 0.05 * zz, 0.3 + 0 * xx], -1)
    out = o.resample_field(gi, f, go)
    oz, oy, ox = np.meshgrid(0.5 + 2 * np.arange(20), 0.5 + 2 * np.arange(18), 0.5 + 2 * np.arange(16), indexing="ij")
    ref = np.stack([0.1 * ox - 0.2 * oy, 0.05 * oz, 0.3 + 0 * ox], -1)
    core = (ox > 1.5) & (ox < 29.5) & (oy > 1.5) & (oy < 33.5) & (oz > 1.5) & (oz < 37.5)
    np.testing.assert_allclose(out[core], ref[core], atol=1e-12)
    # a coarse grid that covers only part of the fine one: points outside get 0
    gsmall = o.Grid([4, 4, 4], [4.0, 4.0, 4.0], [1.5, 1.5, 1.5])
    out2 = o.resample_field(gsmall, np.ones((4, 4, 4, 3)), go)
    assert out2[-1, -1, -1].tolist() == [0.0, 0.0, 0.0]
    assert out2[0, 0, 0].tolist() == [1.0, 1.0, 1.0]


# ---------------------------------------------------------------- CC metric
def _window(a, idx, r):
    sl = tuple(slice(max(i - r, 0), min(i + r, n - 1) + 1) for i, n in zip(idx, a.shape))
    return a[sl].ravel()


def _cc_pearson(F, M, idx, r):
    f, m = _window(F, idx, r), _window(M, idx, r)
    return np.corrcoef(f, m)[0, 1] ** 2


def _grad_rule(a, idx, axis_np, h):
    n = a.shape[axis_np]
    i = idx[axis_np]
    if i == n - 1:
        return 0.0
    up = list(idx)
    up[axis_np] += 1
    if i == 0:
        return (a[tuple(up)] - a[idx]) / (2 * h)
    dn = list(idx)
    dn[axis_np] -= 1
    return (a[tuple(up)] - a[tuple(dn)]) / (2 * h)


@pytest.mark.parametrize("shape,r", [((9, 7), 1), ((9, 7), 2), ((5, 6, 7), 1), ((6, 5, 7), 2)])
def test_cc_value_is_mean_squared_pearson(shape, r):
    F = rng.standard_normal(shape)
    M = 0.5 * F + rng.standard_normal(shape)
    D = len(shape)
    g = o.Grid(list(reversed(shape)), [1.0] * D)
    v, g1, g2 = o.cc(g, F, M, r)
    ref = np.mean([_cc_pearson(F, M, idx, r) for idx in np.ndindex(*shape)])
    assert v == pytest.approx(-ref, rel=1e-12)


@pytest.mark.parametrize("shape,r", [((9, 8), 2), ((6, 7, 5), 1), ((7, 6, 8), 2)])
def test_cc_derivative_is_fd_of_local_cc_times_gradient(shape, r):
    """g_f2m = d(localCC)/d(m_centre) * grad m ; g_m2m = d(localCC)/d(f_centre) * grad f."""
    F = rng.standard_normal(shape)
    M = 0.3 * F + rng.standard_normal(shape)
    D = len(shape)
    sp = [1.0, 2.0, 0.5][:D]
    g = o.Grid(list(reversed(shape)), sp)
    _, g1, g2 = o.cc(g, F, M, r)
    h = 1e-6
    for idx in [tuple(s // 2 for s in shape), (0,) * D, tuple(s - 1 for s in shape), (1,) + (0,) * (D - 1)]:
        Mp, Mm = M.copy(), M.copy()
        Mp[idx] += h
        Mm[idx] -= h
        dcc_dm = (_cc_pearson(F, Mp, idx, r) - _cc_pearson(F, Mm, idx, r)) / (2 * h)
        Fp, Fm = F.copy(), F.copy()
        Fp[idx] += h
        Fm[idx] -= h
        dcc_df = (_cc_pearson(Fp, M, idx, r) - _cc_pearson(Fm, M, idx, r)) / (2 * h)
        for c in range(D):  # component c is axis x,y,z -> numpy axis D-1-c
            gm = _grad_rule(M, idx, D - 1 - c, sp[c])
            gf = _grad_rule(F, idx, D - 1 - c, sp[c])
            assert g1[idx][c] == pytest.approx(dcc_dm * gm, rel=1e-5, abs=1e-8)
            assert g2[idx][c] == pytest.approx(dcc_df * gf, rel=1e-5, abs=1e-8)


def test_cc_closed_forms():
    F = rng.standard_normal((8, 9, 10))
    g = o.Grid([10, 9, 8])
    v, g1, g2 = o.cc(g, F, F, 2)
    assert v == pytest.approx(-1.0, abs=1e-12)
    v, g1, g2 = o.cc(g, F, 2.5 * F - 0.7, 2)
    assert v == pytest.approx(-1.0, abs=1e-12)
    assert np.abs(g1).max() < 1e-9 and np.abs(g2).max() < 1e-9
    # constant windows are degenerate: localCC = 1, zero derivative
    C0 = np.zeros((8, 9, 10))
    v, g1, g2 = o.cc(g, C0, F, 1)
    assert v == pytest.approx(-1.0, abs=1e-15)
    assert not g1.any() and not g2.any()
    # anti-correlated images also give localCC = 1 (squared correlation)
    v, _, _ = o.cc(g, F, -F, 1)
    assert v == pytest.approx(-1.0, abs=1e-12)


def test_cc_symmetric_roles():
    F = rng.standard_normal((6, 7, 8))
    M = rng.standard_normal((6, 7, 8))
    g = o.Grid([8, 7, 6])
    v1, a1, b1 = o.cc(g, F, M, 2)
    v2, a2, b2 = o.cc(g, M, F, 2)
    assert v1 == pytest.approx(v2, rel=1e-13)
    np.testing.assert_allclose(a1, b2, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b1, a2, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------- scale, invert
def test_scale_update():
    u = rng.standard_normal((5, 6, 2))
    g = o.Grid([6, 5], [2.0, 0.5])
    s, mx = o.scale_update(g, u, 0.25)
    ref = np.sqrt((u[..., 0] / 2.0) ** 2 + (u[..., 1] / 0.5) ** 2).max()
    assert mx == pytest.approx(ref)
    np.testing.assert_allclose(s, u * 0.25 / ref)
    z, mz = o.scale_update(g, np.zeros_like(u), 0.25)
    assert mz == 0 and not z.any()


def test_invert_constant_shift():
    g = o.Grid([20, 18])
    c = np.array([0.3, -0.2])
    fld = np.broadcast_to(c, (18, 20, 2)).copy()
    inv, k, mx, mn = o.invert(g, fld)
    assert k == 20  # boundary voxels stay 0, so max and mean never drop below the tolerances
    np.testing.assert_allclose(inv[2:-2, 2:-2], np.broadcast_to(-c, inv[2:-2, 2:-2].shape), atol=1e-6)
    assert np.all(inv[0] == 0) and np.all(inv[:, -1] == 0)


def test_invert_smooth_field_residual_and_stop_rule():
    g = o.Grid([24, 20, 16])
    f = ndimage.gaussian_filter(rng.standard_normal((16, 20, 24, 3)), (3, 3, 3, 0)) * 4.0
    f[0] = f[-1] = 0
    f[:, 0] = f[:, -1] = 0
    f[:, :, 0] = f[:, :, -1] = 0
    inv, k, mx, mn = o.invert(g, f)
    assert 1 <= k <= 20
    assert (mx <= 0.1 or mn <= 0.001 or k == 20)
    e = o.compose(g, f, inv)[2:-2, 2:-2, 2:-2]  # e = inv + f(x + inv)
    assert np.sqrt((e ** 2).sum(-1)).mean() < 0.05
    # the first step uses eps 0.75: one iteration from zero moves inv to -0.75 * clamp(field)
    inv1, k1, _, _ = o.invert(g, np.zeros_like(f))
    assert k1 == 1 and not inv1.any()


# ---------------------------------------------------------------- geometry, preprocessing
def test_shrink_grid_closed_forms():
    g = o.shrink_grid(o.Grid([192, 224, 192]), 4)
    assert g.size == [48, 56, 48] and g.spacing == [4.0, 4.0, 4.0] and g.origin == [1.5, 1.5, 1.5]
    g = o.shrink_grid(o.Grid([192, 224, 192]), 2)
    assert g.size == [96, 112, 96] and g.origin == [0.5, 0.5, 0.5]
    g = o.shrink_grid(o.Grid([193, 10, 7], [1.0, 1.0, 1.0], [5.0, -2.0, 0.0]), 4)
    assert g.size == [48, 2, 1]
    assert g.origin[0] == pytest.approx(5.0 + 96.0 - 23.5 * 4.0)
    # anisotropic: the finest axis gets the factor, others the closest factor (smaller on ties)
    g = o.shrink_grid(o.Grid([64, 64, 32], [1.0, 1.0, 2.0]), 4)
    assert g.spacing == [4.0, 4.0, 4.0] and g.size == [16, 16, 16]
    g = o.shrink_grid(o.Grid([64, 64, 32], [1.0, 1.0, 3.0]), 2)
    assert g.spacing[2] == 3.0  # |3*1 - 2| = 1 ties |3*2 - 2|? no: 4 > 1 -> factor 1


@pytest.mark.parametrize("seed", [0, 1, 2, 3])
def test_preprocess_window(seed):
    """Window [lo, hi] in float (lo = the minimum rounded to float, which may
    round above it: samples below lo clamp to 0), rescaled to [0, 1)."""
    x = np.random.default_rng(seed).uniform(-3, 5, 1000)
    out = o.preprocess(x)
    lo = np.float32(x.min())
    hi = np.float32(x.max() + (x.max() - x.min()) / 2560.0)
    want = np.clip((x - float(lo)) / (float(hi) - float(lo)), 0.0, 1.0)
    np.testing.assert_allclose(out, want, atol=1e-12)
    assert out.min() >= 0 and out.max() < 1


def test_nn_labels_round_half_up():
    lab = np.arange(1, 11, dtype=np.int32)[None, :].repeat(3, 0)  # (y=3, x=10)
    g = o.Grid([10, 3])
    u = np.zeros((3, 10, 2), np.float32)
    u[..., 0] = 0.5
    out = o.warp_labels_nn(g, u, g, lab)
    np.testing.assert_array_equal(out[:, :9], lab[:, 1:])  # x + 0.5 rounds up
    assert np.all(out[:, 9] == 0)  # 9.5 is outside the buffer
    u[..., 0] = -0.5
    out = o.warp_labels_nn(g, u, g, lab)
    np.testing.assert_array_equal(out, lab)  # x - 0.5 rounds up to x


# ---------------------------------------------------------------- convergence monitor
def test_convergence_monitor():
    assert o.conv_value([-0.5] * 5, 10) > 1e300  # fewer than window values
    assert o.conv_value([-0.5] * 10, 10) == pytest.approx(0.0, abs=1e-15)
    assert o.conv_value(-0.5 - 0.01 * np.arange(10), 10) > 0  # falling energy
    assert o.conv_value(-0.5 + 0.01 * np.arange(10), 10) < 0  # rising energy
    # window 2: the fit passes through both points, slope = E1 - E0 over total |E|
    assert o.conv_value([-1.0, -3.0], 2) == pytest.approx(-((-3.0) - (-1.0)) / 4.0)


# ---------------------------------------------------------------- full SyN stage
def test_syn_recovers_shift_2d():
    F = ndimage.gaussian_filter(np.random.default_rng(0).standard_normal((96, 96)), 3)
    M = ndimage.shift(F, (0, 3.0), order=1, mode="nearest")  # M(x) = F(x - 3)
    g = o.Grid([96, 96])
    P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=[30, 20, 10], learning_rate=0.25, radius=4,
                    force_iterations=True)
    r = o.syn_register(P, g, F, g, M)
    assert r.energy[-1] < r.energy[0] - 0.1
    core = r.fwd[16:-16, 16:-16]
    # the forward field maps fixed to moving points: F(x) = M(x + 3)
    assert np.median(core[..., 0]) == pytest.approx(3.0, abs=0.15)
    assert abs(np.median(core[..., 1])) < 0.1
    np.testing.assert_allclose(np.median(r.inv[16:-16, 16:-16], axis=(0, 1)), [-3.0, 0.0], atol=0.15)
    # symmetric normalisation: swapping the images swaps forward and inverse
    r2 = o.syn_register(P, g, M, g, F)
    np.testing.assert_allclose(r2.fwd, r.inv, atol=1e-6)
    np.testing.assert_allclose(r2.energy, r.energy, rtol=1e-9)


def _rot3(a, b, c):
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


@pytest.mark.parametrize("D", [2, 3])
def test_direction_equivariance_of_the_stage(D):
    """Both images sharing a direction R with isotropic spacing: every step
    (warps, CC derivatives turned into physical vectors, compositions,
    inversions, shrink origins) is equivariant and every norm invariant, so
    the stage gives the identity-direction energies and R times its fields.
    A dropped gradient rotation or a wrong transpose in the index/point
    maps breaks this."""
    from scipy import ndimage
    rng = np.random.default_rng(5)
    size = (26, 22, 20)[:D]
    shape = tuple(reversed(size))
    F = ndimage.gaussian_filter(rng.standard_normal(shape), 2.0)
    M = ndimage.shift(F, [0.8, -1.2, 0.6][:D], order=1, mode="nearest")
    F, M = (F - F.min()) / np.ptp(F), (M - M.min()) / np.ptp(M)
    R = _rot3(0.4, -0.3, 0.7) if D == 3 else _rot3(0, 0, 0.6)[:2, :2]
    R = R @ np.diag([-1.0] + [1.0] * (D - 1))            # improper, like an LPS flip
    P = o.SynParams(shrink=[2, 1], sigma=[1, 0], iterations=[5, 3], learning_rate=0.2, update_variance=3.0,
                    total_variance=0.0, radius=2, force_iterations=True)
    org = [3.0, -5.0, 2.0][:D]
    gI = o.Grid(list(size), [1.5] * D, org)
    gR = o.Grid(list(size), [1.5] * D, org, R)
    rI = o.syn_register(P, gI, F, gI, M)
    rR = o.syn_register(P, gR, F, gR, M)
    np.testing.assert_allclose(rR.energy, rI.energy, rtol=1e-9)
    scale = np.abs(rI.fwd).max()
    assert scale > 0.1
    np.testing.assert_allclose(rR.fwd, rI.fwd @ R.T, atol=1e-8 * scale)
    np.testing.assert_allclose(rR.inv, rI.inv @ R.T, atol=1e-8 * scale)
    # shrink: the level grid keeps the direction and the physical centre
    lr, li = o.shrink_grid(gR, 2), o.shrink_grid(gI, 2)
    np.testing.assert_allclose(lr.direction, R)
    cR = np.asarray(lr.origin) + R @ ((np.asarray(lr.size) - 1) / 2.0 * np.asarray(lr.spacing))
    cR0 = np.asarray(org) + R @ ((np.asarray(size) - 1) / 2.0 * 1.5)
    np.testing.assert_allclose(cR, cR0, atol=1e-12)
    np.testing.assert_allclose(li.size, lr.size)


# ------------------------------------------------- initial transforms of the stage
def _rigid(D, seed):
    rng = np.random.default_rng(seed)
    R = _rot3(*rng.uniform(-0.3, 0.3, 3)) if D == 3 else _rot3(0, 0, rng.uniform(-0.3, 0.3))[:2, :2]
    return R, rng.uniform(-2, 2, D), rng.uniform(5, 15, D)


def _header_through_inverse(g: "o.Grid", inverses):
    """Grid whose voxel idx sits at inv_k(...inv_1(point_g(idx))): sampling it
    at p equals sampling g at T_1(...T_k(p)).  Each inverse is (A, x -> A x + b)
    with A orthogonal-times-diagonal so the result is again a grid."""
    D = g.dim
    S = np.asarray(g.direction, float).reshape(D, D) @ np.diag(np.asarray(g.spacing, float))
    org = np.asarray(g.origin, float)
    for A, b in inverses:
        S = A @ S
        org = A @ org + b
    sp = np.linalg.norm(S, axis=0)
    return o.Grid(list(g.size), list(sp), list(org), S / sp)


def _stage_params():
    return o.SynParams(shrink=[2, 1], sigma=[1, 0], iterations=[5, 3], learning_rate=0.2, update_variance=3.0,
                       total_variance=0.0, radius=2, force_iterations=True)


def _pair(D, seed):
    rng = np.random.default_rng(seed)
    size = (26, 22, 20)[:D]
    shape = tuple(reversed(size))
    F = ndimage.gaussian_filter(rng.standard_normal(shape), 2.0)
    M = ndimage.shift(F, [0.8, -1.2, 0.6][:D], order=1, mode="nearest")
    return size, (F - F.min()) / np.ptp(F), (M - M.min()) / np.ptp(M)


@pytest.mark.parametrize("D", [2, 3])
def test_initial_rigid_transform_is_a_moving_header_change(D):
    """A rigid moving initial transform y = R (x - c) + t + c is the same as
    moving the moving image's header by its inverse x = R^T (y - t - c) + c:
    the stage must give the same energies and fields.  Applying the
    transform before the displacement, transposing R or dropping the centre
    breaks this."""
    size, F, M = _pair(D, 11)
    R, t, c = _rigid(D, 3)
    gF = o.Grid(list(size), [1.2] * D, [1.0, -2.0, 0.5][:D])
    gM = o.Grid(list(size), [1.2] * D, [0.0, -1.0, 1.5][:D])
    gMh = _header_through_inverse(gM, [(R.T, -R.T @ (t + c) + c)])
    P = _stage_params()
    r1 = o.syn_register(P, gF, F, gM, M, moving_xforms=[("affine", R, t, c)])
    r2 = o.syn_register(P, gF, F, gMh, M)
    assert r1.energy.max() < -0.15 and np.abs(r1.fwd).max() > 0.3
    np.testing.assert_allclose(r1.energy, r2.energy, rtol=1e-9)
    np.testing.assert_allclose(r1.fwd, r2.fwd, atol=1e-8)
    np.testing.assert_allclose(r1.inv, r2.inv, atol=1e-8)
    # the transform does change the result (the test is not vacuous)
    r0 = o.syn_register(P, gF, F, gM, M)
    assert np.abs(r0.energy - r1.energy).max() > 1e-3


def test_initial_scaling_and_composite_order():
    """Anisotropic scaling y = S x + b equals a header with spacing / S; a
    list [T0, T1] maps a point through T0 first, so it equals the header
    moved by T1^-1 and then T0^-1 (rotations about different centres do not
    commute, so the reversed order fails)."""
    D = 3
    size, F, M = _pair(D, 12)
    gF = o.Grid(list(size), [1.0] * D, [0.0] * D)
    gM = o.Grid(list(size), [1.0] * D, [0.5, -0.5, 0.0])
    P = _stage_params()
    s = np.array([1.04, 0.97, 1.02])
    b = np.array([-0.4, 0.3, 0.2])
    r1 = o.syn_register(P, gF, F, gM, M, moving_xforms=[("affine", np.diag(s), b, np.zeros(3))])
    r2 = o.syn_register(P, gF, F, _header_through_inverse(gM, [(np.diag(1 / s), -b / s)]), M)
    np.testing.assert_allclose(r1.energy, r2.energy, rtol=1e-9)
    np.testing.assert_allclose(r1.fwd, r2.fwd, atol=1e-8)
    R0, t0, c0 = _rigid(D, 4)
    R1, t1, c1 = _rigid(D, 5)
    inv0 = (R0.T, -R0.T @ (t0 + c0) + c0)
    inv1 = (R1.T, -R1.T @ (t1 + c1) + c1)
    xs = [("affine", R0, t0, c0), ("affine", R1, t1, c1)]
    ra = o.syn_register(P, gF, F, gM, M, moving_xforms=xs)
    rb = o.syn_register(P, gF, F, _header_through_inverse(gM, [inv1, inv0]), M)
    rc = o.syn_register(P, gF, F, _header_through_inverse(gM, [inv0, inv1]), M)
    np.testing.assert_allclose(ra.energy, rb.energy, rtol=1e-9)
    np.testing.assert_allclose(ra.fwd, rb.fwd, atol=1e-8)
    assert np.abs(ra.energy - rc.energy).max() > 1e-6


def test_fixed_initial_transform_mirrors_the_moving_one():
    """SyN is symmetric: swapping the images (same grid) swaps the forward
    and inverse fields.  With the transform moved along, a fixed initial
    transform must act on the fixed side exactly as a moving one does on
    the moving side."""
    D = 2
    size, F, M = _pair(D, 13)
    g = o.Grid(list(size), [1.0] * D, [0.0] * D)
    R, t, c = _rigid(D, 6)
    P = _stage_params()
    x = [("affine", R, t, c)]
    r1 = o.syn_register(P, g, F, g, M, fixed_xforms=x)
    r2 = o.syn_register(P, g, M, g, F, moving_xforms=x)
    np.testing.assert_allclose(r1.energy, r2.energy, rtol=1e-9)
    np.testing.assert_allclose(r1.fwd, r2.inv, atol=1e-8)
    np.testing.assert_allclose(r1.inv, r2.fwd, atol=1e-8)
    r3 = o.syn_register(P, g, F, g, M, fixed_xforms=[("identity",)], moving_xforms=[("identity",)])
    r4 = o.syn_register(P, g, F, g, M)
    np.testing.assert_array_equal(r3.fwd, r4.fwd)


# ------------------------------------------------------------ restored state
def test_restored_state_continues_the_stage_exactly():
    """A stage split in two -- levels 4x2 saving the state, then level 1
    restored from it -- is the same computation as the single 4x2x1 stage:
    the restored fields are resampled onto the new level-0 grid as the
    single run resamples them between levels.  Energies and fields equal
    bit for bit; without the restore the second part differs."""
    D = 3
    size, F, M = _pair(D, 14)
    g = o.Grid(list(size), [1.1, 1.0, 1.2], [2.0, -1.0, 0.5])
    kw = dict(learning_rate=0.2, update_variance=3.0, total_variance=0.0, radius=2, force_iterations=True)
    full = o.syn_register(o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=[4, 3, 2], **kw), g, F, g, M)
    a = o.syn_register(o.SynParams(shrink=[4, 2], sigma=[2, 1], iterations=[4, 3], **kw), g, F, g, M,
                       want_state=True)
    gS = o.shrink_grid(g, 2)
    P1 = o.SynParams(shrink=[1], sigma=[0], iterations=[2], **kw)
    b = o.syn_register(P1, g, F, g, M, restore=(gS, a.state))
    np.testing.assert_array_equal(b.energy, full.energy[-2:])
    np.testing.assert_array_equal(b.fwd, full.fwd)
    np.testing.assert_array_equal(b.inv, full.inv)
    cold = o.syn_register(P1, g, F, g, M)
    assert np.abs(cold.fwd - b.fwd).max() > 0.05
    # a state already on the level grid is taken as it is (resampling at grid
    # points, up to the rounding of the point-to-index map)
    st = o.syn_register(P1, g, F, g, M, restore=(gS, a.state), want_state=True).state
    c = o.syn_register(o.SynParams(shrink=[1], sigma=[0], iterations=[0], **kw), g, F, g, M, restore=(g, st),
                       want_state=True)
    np.testing.assert_allclose(c.state, st, rtol=0, atol=1e-12)


# ---------------------------------------------------------------- per-level results and the tie rule
def _small_pair_3d():
    from ants_amd import phantoms as ph
    return ph.phantom_pair_3d(3, size=(32, 36, 30), max_disp_mm=3.0, svf_sigma_mm=5.0)


def test_level_results_equal_truncated_schedules():
    """Level l's result is the output of the same schedule with no
    iterations after level l (fields carried through the later grids, then
    composed), bit for bit; the last level's is the stage output."""
    pair = _small_pair_3d()
    g = o.Grid(list(pair.size))
    iters = [6, 4, 2]
    P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=iters, learning_rate=0.1, radius=2,
                    force_iterations=True)
    full = o.syn_register(P, g, pair.fixed, g, pair.moving, want_levels=True)
    np.testing.assert_array_equal(full.level_fwd[-1], full.fwd)
    np.testing.assert_array_equal(full.level_inv[-1], full.inv)
    for lv in range(2):
        it = [n if k <= lv else 0 for k, n in enumerate(iters)]
        Pt = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=it, learning_rate=0.1, radius=2,
                         force_iterations=True)
        t = o.syn_register(Pt, g, pair.fixed, g, pair.moving)
        np.testing.assert_array_equal(full.level_fwd[lv], t.fwd)
        np.testing.assert_array_equal(full.level_inv[lv], t.inv)
        assert np.abs(full.level_fwd[lv] - full.fwd).max() > 1e-4  # later levels changed the field


def test_strict_tie_rule_only_differs_at_the_tie():
    """The guarded and strict inversion max tests differ only when the max
    norm lies within 1e-5 relative above 0.1: with a learning rate of 0.2 no
    inversion starts on that edge, and the two rules give the same bits."""
    pair = _small_pair_3d()
    g = o.Grid(list(pair.size))
    runs = []
    for rule in (0, 1):
        P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=[5, 3, 2], learning_rate=0.2, radius=2,
                        force_iterations=True, inv_tie_rule=rule)
        runs.append(o.syn_register(P, g, pair.fixed, g, pair.moving))
    np.testing.assert_array_equal(runs[0].fwd, runs[1].fwd)
    np.testing.assert_array_equal(runs[0].inv_iters, runs[1].inv_iters)


@pytest.mark.parametrize("D", [2, 3])
def test_initial_constant_field_is_a_translation(D):
    """A displacement-field initial transform whose field is a constant t on a
    grid covering every sampled point is the affine translation by t; a field
    then an affine equals the translation composed before the affine (the
    list order of the composite)."""
    size, F, M = _pair(D, 21)
    gF = o.Grid(list(size), [1.2] * D, [1.0, -2.0, 0.5][:D])
    gM = o.Grid(list(size), [1.2] * D, [0.0, -1.0, 1.5][:D])
    t = np.array([0.75, -0.375, 0.3125][:D])  # exact in float32, as the field stores it
    fsize = [int(s * 1.2 / 2.0) + 12 for s in size]
    fg = o.og_geom(fsize, [2.0] * D, [-12.0] * D)
    fld = np.broadcast_to(t.astype(np.float32), tuple(reversed(fsize)) + (D,)).copy()
    P = _stage_params()
    r_field = o.syn_register(P, gF, F, gM, M, moving_xforms=[("field", fg, fld)])
    r_aff = o.syn_register(P, gF, F, gM, M, moving_xforms=[("affine", np.eye(D), t, np.zeros(D))])
    np.testing.assert_allclose(r_field.energy, r_aff.energy, rtol=1e-9)
    np.testing.assert_allclose(r_field.fwd, r_aff.fwd, atol=1e-7)
    R, tr, c = _rigid(D, 5)
    r_fa = o.syn_register(P, gF, F, gM, M, moving_xforms=[("field", fg, fld), ("affine", R, tr, c)])
    # the translation first: y = R (x + t - c) + tr + c = R (x - c') + tr + c' with c' = c - t ... as one affine
    off = R @ (t - c) + tr + c
    r_one = o.syn_register(P, gF, F, gM, M, moving_xforms=[("affine", R, off, np.zeros(D))])
    np.testing.assert_allclose(r_fa.energy, r_one.energy, rtol=1e-9)
    np.testing.assert_allclose(r_fa.fwd, r_one.fwd, atol=1e-7)
