"""Linear stages, host side (no GPU): the library's sample points and
parameter scales against the oracle, and the oracle itself pinned by what
the mathematics fixes -- the MT19937 reference stream, Euler angle
conventions, Jacobians by finite differences, metric derivatives on images
where central differences are exact, closed-form metric values, the
golden-section search on known functions, and the scales' closed forms."""
import math

import numpy as np
import pytest
from scipy import interpolate as si
from scipy.spatial.transform import Rotation

import oracle as O
import linear_oracle as LO
from ants_amd import linear as LN
from ants_amd.apply_transforms import AffineTransform, Geometry


def rot3(a):
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])


def test_twister_reference_stream():
    """init_genrand(5489) starts 3499211612, 581869302, 3890346734 (the
    MT19937 reference output); a bounded draw stays in range."""
    tw = LO.Twister(5489)
    assert list(tw.ints(3)) == [3499211612, 581869302, 3890346734]
    tw = LO.Twister(1)
    v = [tw.int_below_or_equal(10) for _ in range(2000)]
    assert min(v) == 0 and max(v) == 10
    z = LO.Twister(2).normals(20000)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1.0) < 0.03


@pytest.mark.parametrize("samp,pct", [("Regular", 0.25), ("Random", 0.1), ("Regular", 0.3), ("None", 1.0)])
@pytest.mark.parametrize("D", [2, 3])
def test_sample_points_match_oracle(samp, pct, D):
    g = Geometry((23, 19, 17)[:D], (1.2, 0.9, 1.1)[:D], (3.0, -2.0, 1.0)[:D], rot3(0.3)[:D, :D])
    st = LN.LinearStage("Rigid", 0.1, "MeanSquares", sampling=samp, percentage=pct, iterations=(5, 5),
                        shrink=(2, 1), sigma=(1, 0), seed=7)
    og = O.Grid(g.size, g.spacing, g.origin, g.direction)
    for lev, f in ((0, 2), (1, 1)):
        a = LN.sample_points(st, lev, g)
        gv = O.shrink_grid(og, f)
        if samp == "None":
            assert a.shape == (0, D)   # the level grid itself, no explicit list
            continue
        b = LO.sample_points(gv, {"Regular": LO.REGULAR, "Random": LO.RANDOM}[samp], pct, 7)
        assert a.shape == b.shape
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)
        if samp == "Regular":
            step = math.ceil(1.0 / pct)
            assert a.shape[0] == -(-gv.count // step)
            # jitter: N(0, spacing/3) per physical axis around the voxel points
            vox = LO.grid_points(gv, np.arange(0, gv.count, step))
            z = (a - vox) / (np.asarray(gv.spacing) / 3.0)
            if a.shape[0] > 500:
                assert abs(z.mean()) < 0.1 and abs(z.std() - 1.0) < 0.1


def test_scales_match_oracle_and_closed_forms():
    g = Geometry((23, 19, 17), (1.2, 0.9, 1.1), (3.0, -2.0, 1.0), np.eye(3))
    og = O.Grid(g.size, g.spacing, g.origin, g.direction)
    c = np.array([10.0, 5.0, 3.0])
    init = [AffineTransform(np.array([[1.1, 0.1, 0], [0, 0.9, 0.05], [0.02, 0, 1.0]]), [1, 2, 3], [4, 5, 6])]
    chain = LO.affine_chain([t.parameters() for t in init], 3)
    corners = LO.corners(O.shrink_grid(og, 2))
    for kind, name, p in ((LO.TRANSLATION, "Translation", [0.1, 0.2, 0.3]),
                          (LO.RIGID, "Rigid", [0.1, -0.2, 0.05, 1, 2, 3]),
                          (LO.AFFINE, "Affine", list(np.eye(3).ravel()) + [1, 2, 3])):
        st = LN.LinearStage(name, 0.1, "MeanSquares", iterations=(5, 5), shrink=(2, 1), sigma=(1, 0))
        a = LN.estimate_scales(st, 0, g, p, init, c)
        b = LO.estimate_scales(kind, p, c, chain, corners, 3)
        np.testing.assert_allclose(a, b, rtol=1e-10)
    # closed forms with no initial transform: a translation moves every point
    # by its own size (scale 1); an affine entry A_ij by |x_j - c_j|; a small
    # rotation about axis x from zero by 2 sin(d/2) times the distance from
    # the axis through c
    ident = (np.eye(3), np.zeros(3))
    np.testing.assert_allclose(LO.estimate_scales(LO.TRANSLATION, [0, 0, 0], c, ident, corners, 3), 1.0, rtol=1e-9)
    sa = LO.estimate_scales(LO.AFFINE, list(np.eye(3).ravel()) + [0, 0, 0], c, ident, corners, 3)
    for i in range(3):
        for j in range(3):
            assert sa[i * 3 + j] == pytest.approx(np.max(np.abs(corners[:, j] - c[j])) ** 2, rel=1e-9)
    sr = LO.estimate_scales(LO.RIGID, [0] * 6, c, ident, corners, 3)
    rx = np.max(np.hypot(corners[:, 1] - c[1], corners[:, 2] - c[2]))
    assert sr[0] == pytest.approx((2 * math.sin(0.005) * rx / 0.01) ** 2, rel=1e-9)


def test_euler_convention_and_jacobians():
    """Euler3DTransform's R = Rz Rx Ry is scipy's extrinsic y-x-z sequence;
    the Jacobian forms equal finite differences of the stage map."""
    ax, ay, az = 0.3, -0.2, 0.5
    A, _ = LO.stage_matrix(LO.RIGID, [ax, ay, az, 0, 0, 0], 3)
    np.testing.assert_allclose(A, Rotation.from_euler("yxz", [ay, ax, az]).as_matrix(), atol=1e-15)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((5, 3)) * 10
    c = np.array([1.0, -2.0, 3.0])
    for kind, D, p in ((LO.RIGID, 3, [0.3, -0.2, 0.5, 1, 2, 3]), (LO.RIGID, 2, [0.4, 1, -1]),
                       (LO.AFFINE, 3, list(np.eye(3).ravel() + 0.1) + [1, 2, 3]), (LO.TRANSLATION, 2, [1, 2])):
        cc = c[:D]
        xx = x[:, :D]
        L, tau = LO.jacobian_forms(kind, p, D)

        def smap(q):
            Am, t = LO.stage_matrix(kind, q, D)
            return (xx - cc) @ Am.T + cc + t
        for i in range(len(p)):
            e = np.zeros(len(p))
            e[i] = 1e-6
            fd = (smap(np.asarray(p) + e) - smap(np.asarray(p) - e)) / 2e-6
            an = (xx - cc) @ L[i].T + tau[i]
            np.testing.assert_allclose(an, fd, rtol=0, atol=1e-8)


def _ramp_setup(metric, kind, D=3):
    """moving image = a linear ramp (central differences and linear
    interpolation exact); fixed = a smooth texture; interior points only"""
    g = O.Grid((30, 28, 26)[:D], (1.1, 0.9, 1.2)[:D], (2.0, -1.0, 0.5)[:D], np.eye(D))
    rng = np.random.default_rng(3)
    from scipy import ndimage
    F = ndimage.gaussian_filter(rng.standard_normal(tuple(reversed(g.size))), 2.0)
    idx = np.meshgrid(*[np.arange(n) for n in reversed(g.size)], indexing="ij")
    M = 0.2 + 0.01 * idx[-1] + 0.015 * idx[-2] + (0.007 * idx[0] if D == 3 else 0.0)
    inner = O.Grid([n - 12 for n in g.size], g.spacing, [o + 6 * s for o, s in zip(g.origin, g.spacing)], np.eye(D))
    pts = LO.grid_points(inner)[::3]
    c = np.asarray(g.origin) + 10.0
    S = LO.MetricSetup(kind, metric, g, F, g, M, pts, c, (np.eye(D), np.zeros(D)), (np.eye(D), np.zeros(D)), 24,
                       (float(F.min()), float(F.max()), float(M.min()), float(M.max())))
    return S


@pytest.mark.parametrize("metric", [LO.MEAN_SQUARES, LO.GC, LO.MATTES])
@pytest.mark.parametrize("kind", [LO.RIGID, LO.AFFINE])
def test_oracle_derivative_equals_finite_differences(metric, kind):
    """Where the gradient is exact the returned derivative is minus the
    derivative of the value: a sign, an index or a transposed Jacobian
    fails this."""
    S = _ramp_setup(metric, kind)
    p = np.array([0.02, -0.01, 0.015, 0.3, -0.2, 0.1]) if kind == LO.RIGID else \
        np.concatenate([(np.eye(3) + 0.01 * np.arange(9).reshape(3, 3) / 9).ravel(), [0.3, -0.2, 0.1]])
    v, d, n = LO.evaluate(S, p)
    assert n == len(S.points)
    fd = np.zeros_like(p)
    for i in range(len(p)):
        h = 1e-6 * max(1.0, abs(p[i]))
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = -(LO.evaluate(S, p + e, False)[0] - LO.evaluate(S, p - e, False)[0]) / (2 * h)
    np.testing.assert_allclose(d, fd, rtol=1e-5, atol=1e-7 * np.abs(fd).max())


def test_oracle_metric_values_closed_forms():
    g = O.Grid((20, 18, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), np.eye(3))
    rng = np.random.default_rng(1)
    F = rng.uniform(0.0, 1.0, tuple(reversed(g.size)))
    pts = LO.grid_points(g)
    ident = (np.eye(3), np.zeros(3))

    def setup(metric, M, bins=16):
        return LO.MetricSetup(LO.TRANSLATION, metric, g, F, g, M, pts, np.zeros(3), ident, ident, bins,
                              (float(F.min()), float(F.max()), float(M.min()), float(M.max())))
    assert LO.evaluate(setup(LO.MEAN_SQUARES, F), [0, 0, 0], False)[0] == 0.0
    M2 = 0.3 * F + 2.0
    assert LO.evaluate(setup(LO.GC, M2), [0, 0, 0], False)[0] == pytest.approx(-1.0, abs=1e-12)
    M3 = rng.uniform(0.0, 1.0, F.shape)
    r = np.corrcoef(F.ravel(), M3.ravel())[0, 1]
    assert LO.evaluate(setup(LO.GC, M3), [0, 0, 0], False)[0] == pytest.approx(-r * r, rel=1e-10)
    # Mattes against an independent Parzen-window sum with scipy's cubic
    # B-spline basis element
    B = 16
    S = setup(LO.MATTES, M3, B)
    v = LO.evaluate(S, [0, 0, 0], False)[0]
    b3 = si.BSpline.basis_element([-2, -1, 0, 1, 2], extrapolate=False)
    f, m = F.ravel(), M3.ravel()
    fbs, mbs = (f.max() - f.min()) / (B - 4), (m.max() - m.min()) / (B - 4)
    fi = np.clip(np.floor(f / fbs - (f.min() / fbs - 2)).astype(int), 2, B - 3)
    mt = m / mbs - (m.min() / mbs - 2)
    H = np.zeros((B, B))
    for k in range(B):
        w = np.nan_to_num(b3(k - mt))
        np.add.at(H[:, k], fi, w)
    p = H / f.size
    pf = np.bincount(fi, minlength=B) / f.size
    pm = p.sum(0)
    nz = p > 0
    mi = np.sum(p[nz] * np.log(p[nz] / np.outer(pf, pm)[nz]))
    assert v == pytest.approx(-mi, rel=1e-9)


def test_central_difference_at_points_hand_values():
    """CentralDifferenceImageFunction::Evaluate(point): samples half a
    spacing either side, divided by their distance.  On the impulse
    0, 0, 1, 0, 0 (spacing 2) at x = 4.5 (index 2.25): 0.25 at index 2.75,
    0.75 at 1.75 -> -0.5 / 2 = -0.25 (whole-spacing neighbours would give
    -0.0625); at index 0.1 the lower neighbour leaves the buffer -> 0."""
    g = O.Grid((5, 1, 1), (2.0, 1.0, 1.0), (0.0, 0.0, 0.0), np.eye(3))
    img = np.zeros((1, 1, 5))
    img[0, 0, 2] = 1.0
    pts = np.array([[4.5, 0.0, 0.0], [0.2, 0.0, 0.0], [3.0, 0.0, 0.0]])
    gr = LO.gradient(g, img, pts)
    assert gr[0, 0] == pytest.approx(-0.25, abs=1e-15)
    assert gr[1, 0] == 0.0
    assert gr[2, 0] == pytest.approx(0.5, abs=1e-15)    # index 1.5: (1 at index 2 - 0 at 1) / 2 (not 0.125)
    assert np.all(gr[:, 1:] == 0.0)                      # one-voxel axes: neighbours outside


def test_golden_section_search():
    """On a unimodal function the search brackets the minimum to its
    tolerance; on a decreasing one it runs to the upper end."""
    for s_star in (0.3, 1.0, 1.7):
        calls = []

        def f(s):
            calls.append(s)
            return (s - s_star) ** 2
        r = LO._LineSearch(f).search(0.0, 1.0, 2.0)
        assert abs(r - s_star) < 0.2 * (1.0 + s_star)
        assert len(calls) <= 23
    r = LO._LineSearch(lambda s: -s).search(0.0, 1.0, 2.0)
    assert r > 1.6
    r = LO._LineSearch(lambda s: s).search(0.0, 1.0, 2.0)
    assert r < 0.4


def test_parse_linear_stage():
    st = LN.LinearStage.from_args("Rigid[0.1]", "Mattes[f.nii,m.nii,1,32,Regular,0.25]",
                                  "[1000x500x250x100,1e-6,10]", "8x4x2x1", "3x2x1x0vox", seed=3)
    assert (st.transform, st.gradient_step, st.metric, st.bins, st.sampling, st.percentage) == \
        ("Rigid", 0.1, "Mattes", 32, "regular", 0.25)
    assert list(st.iterations) == [1000, 500, 250, 100] and list(st.shrink) == [8, 4, 2, 1]
    assert list(st.sigma) == [3, 2, 1, 0] and not st.sigma_in_mm and st.seed == 3
    st = LN.LinearStage.from_args("Affine[0.1]", "GC[f,m,1,1,Random,0.1]", "[10,1e-7,5]", "1", "0mm")
    assert (st.metric, st.sampling, st.window, st.threshold, st.sigma_in_mm) == ("GC", "random", 5, 1e-7, True)
    with pytest.raises(ValueError):
        LN.LinearStage.from_args("Similarity[0.1]", "GC[f,m]", "[10]", "1", "0")
    # antsRegistration's names: MI is Mattes, MSQ / SSD are MeanSquares
    # (StringToMetricType); MI2, the joint-histogram metric, is not provided
    st = LN.LinearStage.from_args("Rigid[0.1]", "MI[f,m,1,32,Regular,0.25]", "[10]", "1", "0")
    assert (st.metric, st.bins, st.sampling, st.percentage) == ("Mattes", 32, "regular", 0.25)
    assert LN.LinearStage.from_args("Rigid[0.1]", "SSD[f,m]", "[10]", "1", "0").metric == "MeanSquares"
    assert LN.LinearStage.from_args("Rigid[0.1]", "msq[f,m]", "[10]", "1", "0").metric == "MeanSquares"
    with pytest.raises(ValueError):
        LN.LinearStage.from_args("Rigid[0.1]", "MI2[f,m,1,32]", "[10]", "1", "0")
    with pytest.raises(ValueError):
        LN.LinearStage.from_args("Rigid[0.1]", "CC[f,m,1,4]", "[10]", "1", "0")
    with pytest.raises(ValueError):
        LN.LinearStage.from_args("Rigid[0.1]", "GC[f,m,1,1,Sparse,0.1]", "[10]", "1", "0")
    assert LN.LinearStage.is_linear("Translation[1]") and not LN.LinearStage.is_linear("SyN[0.1,3,0]")
