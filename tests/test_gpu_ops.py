"""Each gfx950 building block against the FP64 oracle on the same float32
inputs.  Tolerances are stated per test (FP32 device arithmetic against
FP64; integer label work is bit-exact)."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as o
import gpu_helpers as G
from ants_amd import _lib as L

pytestmark = pytest.mark.gpu
rng = np.random.default_rng(7)


def f32(a):
    return np.asarray(a, np.float32)


def grid_np(shape, spacing=None, origin=None):
    return o.Grid(list(reversed(shape)), spacing, origin)


def smooth_random_field(shape, D, sigma=2.0, amp=2.0):
    f = np.stack([ndimage.gaussian_filter(rng.standard_normal(shape), sigma) for _ in range(D)], -1)
    return f32(amp * f / np.abs(f).max())


@pytest.mark.parametrize("shape,sigma,in_mm,sp", [((33, 40), 2.0, False, None), ((17, 23, 29), 1.0, False, None),
                                                 ((16, 20, 12), 2.0, True, [1.0, 0.7, 1.3]),
                                                 ((5, 6, 7), 0.0, False, None),
                                                 # one fused 3-D pass (same taps on every axis), radii 1..6,
                                                 # ragged tiles; radius > 6 falls back to separable passes
                                                 ((40, 37, 35), 2.0, False, None), ((9, 70, 11), 1.5, False, None),
                                                 ((20, 22, 24), 1.0, True, [1.2, 1.2, 1.2]),
                                                 ((35, 33, 50), 0.6, False, None), ((3, 4, 40), 2.5, False, None),
                                                 ((30, 30, 30), 3.0, False, None)])
def test_smooth_image(shape, sigma, in_mm, sp):
    img = f32(rng.standard_normal(shape))
    out = G.smooth_image(img, sigma, in_mm, sp)
    ref = o.smooth_image(grid_np(shape, sp), img, sigma, in_mm)
    np.testing.assert_allclose(out, ref, atol=2e-6 * np.abs(img).max())


@pytest.mark.parametrize("shape,var", [((30, 41), 3.0), ((13, 17, 19), 3.0), ((9, 8, 70), 3.0), ((12, 14), 0.25),
                                       ((10, 11, 12), 0.0), ((13, 17, 19), 0.25), ((20, 9, 14), 0.4),
                                       ((11, 12, 10), 1.0)])
def test_smooth_field(shape, var):
    """Including 3-D fields with 0 < variance < 0.5 (the separable passes
    keep the input for the blend w1 * smooth + (1 - w1) * input)."""
    D = len(shape)
    f = f32(rng.standard_normal(shape + (D,)))
    out = G.smooth_field(f, var)
    ref = o.smooth_field(grid_np(shape), f, var)
    np.testing.assert_allclose(out, ref, atol=2e-6 * np.abs(f).max())


@pytest.mark.parametrize("shape", [(31, 29), (13, 17, 19)])
def test_warp_image_same_grid(shape):
    D = len(shape)
    img = f32(rng.standard_normal(shape))
    u = f32(rng.uniform(-3, 3, shape + (D,)))
    out = G.warp_image(img, shape, u)
    ref = o.warp_image(grid_np(shape), img, grid_np(shape), u)
    # float32 continuous indices: interpolation weights differ by ~1e-6
    np.testing.assert_allclose(out, ref, atol=2e-5 * np.abs(img).max())


def test_warp_image_to_level_grid():
    shape = (40, 44, 36)
    img = f32(rng.standard_normal(shape))
    gF = grid_np(shape)
    gl = o.shrink_grid(gF, 4)
    u = smooth_random_field(gl.shape, 3, 1.0, 6.0)
    out = G.warp_image(img, gl.shape, u, gi=G.desc_of(shape), go=L.desc(gl.size, gl.spacing, gl.origin))
    ref = o.warp_image(gF, img, gl, u)
    np.testing.assert_allclose(out, ref, atol=3e-5 * np.abs(img).max())


@pytest.mark.parametrize("shape,r", [((37, 41), 4), ((20, 18), 2), ((19, 23, 21), 4), ((24, 20, 40), 2),
                                     ((11, 9, 7), 3), ((70, 5, 67), 4), ((64, 4, 32), 4), ((3, 3, 3), 4)])
def test_cc(shape, r):
    D = len(shape)
    F = f32(ndimage.gaussian_filter(rng.standard_normal(shape), 1.0))
    M = f32(0.6 * F + 0.3 * ndimage.gaussian_filter(rng.standard_normal(shape), 1.0) + 0.5)
    F = F + 0.5
    v, g1, g2 = G.cc(F, M, r)
    vr, r1, r2 = o.cc(grid_np(shape), F, M, r)
    assert v == pytest.approx(vr, rel=2e-5)
    scale = max(np.abs(r1).max(), np.abs(r2).max())
    np.testing.assert_allclose(g1, r1, atol=2e-4 * scale)
    np.testing.assert_allclose(g2, r2, atol=2e-4 * scale)


@pytest.mark.parametrize("shape,r", [((19, 23, 21), 4), ((24, 20, 40), 2), ((70, 5, 67), 4), ((40, 36, 48), 4)])
def test_cc_float_shifted_sums(monkeypatch, shape, r):
    """The tiled CC with float box sums over block-shifted samples (the BF16
    pyramid form's precision) against the FP64 oracle, on flat noisy regions
    (intensity 0.6, noise 0.02: sum f^2 is about 900 times the window
    variance) next to an edge; bars written here."""
    monkeypatch.setenv("ANTS_SYN_CC_SUMS", "fp32")
    F = np.full(shape, 0.6)
    F[..., : shape[-1] // 3] = 0.2
    F = f32(F + 0.02 * rng.standard_normal(shape))
    M = f32(0.9 * F + 0.02 * rng.standard_normal(shape) + 0.05)
    v, g1, g2 = G.cc(F, M, r)
    vr, r1, r2 = o.cc(grid_np(shape), F, M, r)
    assert v == pytest.approx(vr, rel=1e-5)
    scale = max(np.abs(r1).max(), np.abs(r2).max())
    np.testing.assert_allclose(g1, r1, atol=1e-3 * scale)
    np.testing.assert_allclose(g2, r2, atol=1e-3 * scale)


def test_cc_degenerate_windows_and_spacing():
    shape = (20, 24, 28)
    F = f32(rng.uniform(0, 1, shape))
    M = F.copy()
    F[:, :, :10] = 0.0  # constant windows -> localCC 1, zero derivative
    M[:8] = 0.25
    sp = [1.0, 2.0, 0.5]
    v, g1, g2 = G.cc(F, M, 2, sp)
    vr, r1, r2 = o.cc(grid_np(shape, sp), F, M, 2)
    assert v == pytest.approx(vr, rel=2e-5)
    assert np.all(g1[:, :, :7] == 0)
    scale = max(np.abs(r1).max(), np.abs(r2).max())
    np.testing.assert_allclose(g1, r1, atol=2e-4 * scale)
    np.testing.assert_allclose(g2, r2, atol=2e-4 * scale)


@pytest.mark.parametrize("shape", [(30, 26), (14, 18, 22)])
def test_compose(shape):
    D = len(shape)
    d = smooth_random_field(shape, D, 2.0, 3.0)
    w = smooth_random_field(shape, D, 2.0, 3.0)
    out = G.compose(d, w)
    ref = o.compose(grid_np(shape), d, w)
    np.testing.assert_allclose(out, ref, atol=3e-5)


def test_scale_update():
    u = f32(rng.standard_normal((12, 14, 16, 3)))
    sp = [1.0, 2.0, 0.5]
    out, mx = G.scale_update(u, 0.25, sp)
    ref, mr = o.scale_update(grid_np((12, 14, 16), sp), u, 0.25)
    assert mx == pytest.approx(mr, rel=1e-6)
    np.testing.assert_allclose(out, ref, rtol=1e-6, atol=1e-7)


@pytest.mark.parametrize("shape,amp,sp", [((40, 36), 2.0, None), ((20, 24, 28), 2.5, None), ((12, 10, 14), 0.05, None),
                                         ((22, 18, 26), 2.0, [1.2, 0.9, 1.5])])
def test_invert(shape, amp, sp):
    D = len(shape)
    f = smooth_random_field(shape, D, 3.0, amp)
    init = smooth_random_field(shape, D, 3.0, 0.3 * amp)
    out, k, mx, mn = G.invert(f, init, sp)
    ref, kr, mxr, mnr = o.invert(grid_np(shape, sp), f, init)
    assert k == kr
    np.testing.assert_allclose(out, ref, atol=2e-4 * max(amp, 1.0))
    assert mx == pytest.approx(mxr, rel=1e-4, abs=1e-6)
    assert mn == pytest.approx(mnr, rel=1e-4, abs=1e-7)


@pytest.mark.parametrize("shape,amp", [((20, 24, 28), 2.5), ((33, 17, 29), 4.0), ((40, 36), 2.0)])
def test_invert_speculative_matches_plain(monkeypatch, shape, amp):
    """Speculative updates (forced on this small grid) give the plain
    scheme's bits, including large displacements where many voxels clamp,
    and ragged sizes whose last mask word is partial."""
    D = len(shape)
    f = smooth_random_field(shape, D, 3.0, amp)
    init = smooth_random_field(shape, D, 3.0, 0.3 * amp)
    monkeypatch.setenv("ANTS_SYN_INV_SPEC_MIN_VOXELS", "1000000000")
    plain = G.invert(f, init)
    monkeypatch.setenv("ANTS_SYN_INV_SPEC_MIN_VOXELS", "0")
    spec = G.invert(f, init)
    assert plain[1] == spec[1] and plain[1] > 1
    np.testing.assert_array_equal(spec[0], plain[0])
    assert spec[2] == plain[2] and spec[3] == plain[3]


@pytest.mark.parametrize("bf16", [False, True])
@pytest.mark.parametrize("fshape,mshape,shrink", [((33, 40, 28), (33, 40, 28), 2), ((29, 31, 37), (25, 35, 30), 1),
                                                   ((48, 40), (44, 46), 4)])
def test_warp_pair(bf16, fshape, mshape, shrink):
    """The iteration's warp kernel (both images through their inverse middle
    fields onto the level grid, 8-byte corner-pair loads for float images)
    against the oracle's ResampleImageFilter restatement, float and BF16
    images (the oracle then warps the BF16-rounded image)."""
    D = len(fshape)
    gF = grid_np(fshape, [1.1] * D, [0.5] * D)
    gM = grid_np(mshape, [0.9] * D, [-1.0] * D)
    gl = o.shrink_grid(gF, shrink)
    F = f32(rng.uniform(0, 1, fshape))
    M = f32(rng.uniform(0, 1, mshape))
    uF = smooth_random_field(gl.shape, D, 2.0, 3.0)
    uM = smooth_random_field(gl.shape, D, 2.0, 3.0)
    d = lambda g: L.desc(g.size, g.spacing, g.origin)
    oF, oM = G.warp_pair(d(gl), d(gF), F, uF, d(gM), M, uM, gl.shape, bf16)
    Fi, Mi = (G.round_bf16(F), G.round_bf16(M)) if bf16 else (F, M)
    rF = o.warp_image(gF, Fi, gl, uF)
    rM = o.warp_image(gM, Mi, gl, uM)
    # float continuous indices (non-integer spacing and origin) against double:
    # a few 1e-6 of a voxel times unit-range gradients
    np.testing.assert_allclose(oF, rF, atol=1e-5)
    np.testing.assert_allclose(oM, rM, atol=1e-5)
    assert (rF == 0).any() or (rM == 0).any()  # some points fall outside a buffer


@pytest.mark.parametrize("shape,margin,sp", [((20, 24, 28), 3, [1.0, 1.0, 1.0]), ((30, 26), 0, [0.8, 1.3]),
                                             ((17, 19, 23), 5, [1.2, 0.9, 1.1])])
def test_inverse_consistency_residual(shape, margin, sp):
    """|fwd(x) + inv(x + fwd(x))| in voxels (FP64 sums on the device) against
    the oracle's composition and numpy norms over the voxels `margin` from
    the faces; an exact inverse pair (constant shift) gives 0 inside."""
    D = len(shape)
    fwd = smooth_random_field(shape, D, 3.0, 2.0)
    inv = f32(-fwd + smooth_random_field(shape, D, 3.0, 0.2))
    mx, mn, n = G.inverse_consistency(fwd, inv, margin, sp)
    r = o.compose(grid_np(shape, sp), inv, fwd)
    q = np.sqrt(((r / np.asarray(sp)) ** 2).sum(-1))
    core = q[tuple(slice(margin, s - margin) for s in shape)]
    assert n == core.size
    assert mx == pytest.approx(core.max(), rel=1e-5)
    assert mn == pytest.approx(core.mean(), rel=1e-5)
    shift = np.zeros(shape + (D,), np.float32)
    shift[...] = 0.5
    mx0, mn0, _ = G.inverse_consistency(shift, -shift, 2, None)
    assert mx0 == 0.0 and mn0 == 0.0


def test_invert_zero_init_constant_shift():
    shape = (16, 18, 20)
    f = np.zeros(shape + (3,), np.float32)
    f[...] = [0.3, -0.2, 0.1]
    out, k, _, _ = G.invert(f)
    ref, kr, _, _ = o.invert(grid_np(shape), f)
    assert k == kr == 20
    np.testing.assert_allclose(out, ref, atol=1e-6)


def test_resample_field():
    gi = o.Grid([12, 14, 10], [4.0, 4.0, 4.0], [1.5, 1.5, 1.5])
    go = o.Grid([24, 28, 20], [2.0, 2.0, 2.0], [0.5, 0.5, 0.5])
    f = smooth_random_field(gi.shape, 3, 1.5, 2.0)
    out = G.resample_field(f, L.desc(gi.size, gi.spacing, gi.origin), L.desc(go.size, go.spacing, go.origin),
                           go.shape + (3,))
    ref = o.resample_field(gi, f, go)
    np.testing.assert_allclose(out, ref, atol=1e-5)


def test_preprocess():
    img = f32(rng.uniform(-2, 7, (9, 10, 11)))
    out = G.preprocess(img)
    ref = o.preprocess(img)
    np.testing.assert_allclose(out, ref, atol=1e-6)


@pytest.mark.parametrize("shape", [(25, 31), (17, 19, 23)])
def test_labels_nn_bit_exact(shape):
    D = len(shape)
    lab = rng.integers(0, 33, shape).astype(np.int32)
    u = f32(rng.uniform(-4, 4, shape + (D,)))
    # put some samples exactly on half-integers and on the buffer edges
    u.reshape(-1, D)[::7] = np.round(u.reshape(-1, D)[::7]) + 0.5
    g = grid_np(shape)
    out = G.labels_nn(u, lab, G.desc_of(shape), G.desc_of(shape), shape)
    ref = o.warp_labels_nn(g, u, g, lab)
    np.testing.assert_array_equal(out, ref)
