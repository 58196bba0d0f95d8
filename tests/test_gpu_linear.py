"""Linear stages on the GPU against the oracle's restatement
(oracle/linear_oracle.py): metric value and derivative passes (MeanSquares,
GC, Mattes; 2-D and 3-D; Translation, Rigid, Affine; full, regular and random
sampling; moving and fixed initial affines; oblique grids) and whole stages
(the same iteration counts, records and parameters), then recovery of known
transforms and the CLI's multi-stage runs."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
import linear_oracle as LO
from ants_amd import linear as LN
from ants_amd.apply_transforms import AffineTransform, Geometry

pytestmark = pytest.mark.gpu

KINDS = {"Translation": LO.TRANSLATION, "Rigid": LO.RIGID, "Affine": LO.AFFINE}
METRICS = {"MeanSquares": LO.MEAN_SQUARES, "GC": LO.GC, "Mattes": LO.MATTES}
SAMPLING = {"None": LO.NONE, "Regular": LO.REGULAR, "Random": LO.RANDOM}


def rot(D, a):
    if D == 2:
        return np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])


def phantom(D, seed=0, oblique=False):
    rng = np.random.default_rng(seed)
    size = (36, 30, 26)[:D]
    sp = (1.1, 0.9, 1.2)[:D]
    org = (4.0, -3.0, 2.0)[:D]
    dirm = rot(D, 0.2) if oblique else np.eye(D)
    g = Geometry(size, sp, org, dirm)
    img = ndimage.gaussian_filter(rng.standard_normal(g.shape), 2.5)
    img = ((img - img.min()) / (img.max() - img.min())).astype(np.float32)
    return g, img


def blob(D, seed):
    """smooth non-negative texture inside the middle half of the grid, zero
    around it: a moved copy stays inside the field of view"""
    rng = np.random.default_rng(seed)
    g = Geometry((40, 36, 32)[:D], (1.1, 0.9, 1.2)[:D], (4.0, -3.0, 2.0)[:D], np.eye(D))
    img = np.zeros(g.shape)
    inner = tuple(slice(s // 4, s - s // 4) for s in g.shape)
    img[inner] = rng.standard_normal(img[inner].shape)
    img = np.clip(ndimage.gaussian_filter(img, 2.0), 0, None)
    return g, (img / img.max()).astype(np.float32)


def moved(g, img, A, t, c):
    """moving(y) = fixed(S^-1 y) for S(x) = A (x - c) + c + t"""
    D = g.dim
    pts = LO.grid_points(O.Grid(g.size, g.spacing, g.origin, g.direction))
    q = (pts - c - t) @ np.linalg.inv(A).T + c
    v, _ = LO.interpolate(O.Grid(g.size, g.spacing, g.origin, g.direction), img, q)
    return v.reshape(g.shape).astype(np.float32)


def og(g):
    return O.Grid(g.size, g.spacing, g.origin, g.direction)


def aff_tuple(t):
    A, tr, c = t.parameters()
    return A, tr, c


def setup(D, kind, metric, sampling, pct, level_shrink, sigma, init=True, fixed_init=False, oblique=False):
    gF, F = phantom(D, 1, oblique)
    A = rot(D, 0.05) * 1.02
    c = np.asarray(gF.origin) + 10.0
    gM = Geometry(tuple(s + 3 for s in gF.size), tuple(s * 1.05 for s in gF.spacing),
                  tuple(o - 1.0 for o in gF.origin), gF.direction)
    M = moved(gF, F, A, np.full(D, 0.8), c)
    M = ndimage.zoom(M, [a / b for a, b in zip(gM.shape, gF.shape)], order=1).astype(np.float32)
    mi = [AffineTransform(rot(D, -0.03), np.full(D, 0.4), c)] if init else []
    fi = [AffineTransform(np.eye(D) * 1.01, np.full(D, -0.2), c)] if fixed_init else []
    st = LN.LinearStage(kind, 0.2, metric, bins=24, sampling=sampling, percentage=pct, iterations=(4,),
                        shrink=(level_shrink,), sigma=(sigma,), seed=3, preprocess=False)
    return gF, F, gM, M, mi, fi, c, st


def oracle_setup(st, gF, F, gM, M, mi, fi, c, level=0):
    gv = O.shrink_grid(og(gF), int(st.shrink[level]))
    Fs = O.smooth_image(og(gF), F.astype(np.float64), st.sigma[level]) if st.sigma[level] > 0 else F.astype(np.float64)
    Ms = O.smooth_image(og(gM), M.astype(np.float64), st.sigma[level]) if st.sigma[level] > 0 else M.astype(np.float64)
    pts = LO.sample_points(gv, SAMPLING[st.sampling], st.percentage, st.seed)
    D = gF.dim
    return LO.MetricSetup(KINDS[st.transform], METRICS[st.metric], og(gF), Fs, og(gM), Ms, pts, np.asarray(c, float),
                          LO.affine_chain([aff_tuple(t) for t in mi], D), LO.affine_chain([aff_tuple(t) for t in fi], D),
                          st.bins, LO.mattes_ranges(Fs, Ms))


def params_for(kind, D, rng):
    P = LO.n_params(KINDS[kind], D)
    p = 0.02 * rng.standard_normal(P)
    if kind == "Affine":
        p[: D * D] += np.eye(D).ravel()
    if kind == "Rigid":
        p[(1 if D == 2 else 3):] *= 20
    if kind == "Translation":
        p *= 20
    return p


CASES = [
    (3, "Rigid", "MeanSquares", "None", 1.0, 2, 0.0, True, False, False),
    (3, "Affine", "GC", "Regular", 0.25, 1, 0.0, True, True, False),
    (3, "Translation", "Mattes", "Random", 0.2, 1, 0.0, False, False, True),
    (3, "Rigid", "Mattes", "Regular", 0.25, 2, 0.0, True, False, True),
    (2, "Affine", "MeanSquares", "Regular", 0.3, 1, 0.0, True, False, True),
    (2, "Rigid", "GC", "None", 1.0, 1, 0.0, False, True, False),
    (2, "Affine", "Mattes", "None", 1.0, 1, 0.0, True, False, False),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-{c[1]}-{c[2]}-{c[3]}" for c in CASES])
def test_metric_pass_matches_oracle(case):
    D, kind, metric, samp, pct, shrink, sigma, init, finit, obl = case
    gF, F, gM, M, mi, fi, c, st = setup(D, kind, metric, samp, pct, shrink, sigma, init, finit, obl)
    S = oracle_setup(st, gF, F, gM, M, mi, fi, c)
    rng = np.random.default_rng(5)
    for trial in range(2):
        p = params_for(kind, D, rng)
        v, d, n = LN.metric_value(st, 0, F, gF, M, gM, p, mi, fi, c, True)
        rv, rd, rn = LO.evaluate(S, p, True)
        print(f"\n{case[:4]} value {v:.10e} vs {rv:.10e}  n {n}/{rn}  |dd| {np.abs(d - rd).max():.2e}"
              f" of {np.abs(rd).max():.2e}")
        assert n == rn
        assert abs(v - rv) <= 1e-10 * max(1.0, abs(rv))
        np.testing.assert_allclose(d, rd, rtol=0, atol=1e-8 * np.abs(rd).max())
        v2, _, _ = LN.metric_value(st, 0, F, gF, M, gM, p, mi, fi, c, False)
        assert v2 == v


@pytest.mark.parametrize("metric", ["MeanSquares", "GC", "Mattes"])
def test_stage_matches_oracle(metric):
    """Two levels (shrink 2 with smoothing, then full resolution) of a Rigid
    stage under a moving initial affine, images windowed and smoothed (in
    float on the GPU, in double in the oracle): the same iteration counts,
    and every record's value within 1e-6 relative up to the first iteration
    whose line search picks a learning rate more than 1e-3 apart -- the
    float/double image difference (~1e-7) can flip a golden-section
    comparison between two nearly equal trial values, after which the two
    paths differ by the step taken -- which comes no earlier than half-way;
    after it every value within 1e-3 relative and the parameters within
    0.01.  With identical inputs the stage agrees to rounding throughout
    (next test)."""
    D = 3
    gF, F, gM, M, mi, fi, c, _ = setup(D, "Rigid", metric, "Regular", 0.25, 1, 0.0)
    st = LN.LinearStage("Rigid", 0.25, metric, bins=24, sampling="Regular", percentage=0.25, iterations=(12, 8),
                        shrink=(2, 1), sigma=(1.0, 0.0), seed=11, preprocess=True)
    res = LN.register_linear(st, F, gF, M, gM, mi, (), c)
    P = LO.LinearParams(kind=LO.RIGID, gradient_step=0.25, metric=METRICS[metric], bins=24, sampling=LO.REGULAR,
                        percentage=0.25, seed=11, iterations=(12, 8), shrink=(2, 1), sigma=(1.0, 0.0), preprocess=True)
    ref = LO.linear_register(P, og(gF), F, og(gM), M, [aff_tuple(t) for t in mi], (), c)
    print(f"\n{metric}: gpu iterations {res.iterations} params {res.params}\n oracle {ref.iterations} {ref.params}")
    assert res.iterations == ref.iterations
    assert len(res.records) == len(ref.records)
    split = next((k for k, (a, b) in enumerate(zip(res.records, ref.records))
                  if abs(a[4] - b[4]) > 1e-3 * max(abs(a[4]), abs(b[4]))), len(ref.records))
    print(f"  line searches agree on {split} of {len(ref.records)} iterations")
    assert split >= len(ref.records) // 2
    for k, (a, b) in enumerate(zip(res.records, ref.records)):
        assert a[:2] == b[:2]
        tol = 1e-6 if k <= split else 1e-3
        assert abs(a[2] - b[2]) <= tol * max(1e-3, abs(b[2])), (k, a, b)
    np.testing.assert_allclose(res.params, ref.params, rtol=0, atol=1e-5 if split == len(ref.records) else 1e-2)


@pytest.mark.parametrize("D,kind,metric", [(3, "Affine", "GC"), (2, "Rigid", "Mattes"), (3, "Translation", "MeanSquares")])
def test_stage_bitwise_inputs_match_oracle(D, kind, metric):
    """With the same input images on both sides (no windowing or smoothing)
    the whole stage agrees to rounding: identical iteration counts, record
    values within 1e-10, learning rates within 1e-8 relative (each level's
    first rate is step / the largest shift under the derivative, whose
    Mattes form sums a fixed-point joint histogram on the GPU: ~1e-9
    relative) or both below 1e-6 (where the line search compares values
    equal to rounding and either choice moves the parameters by less than
    that), parameters within 1e-9."""
    gF, F, gM, M, mi, fi, c, _ = setup(D, kind, metric, "Regular", 0.25, 1, 0.0, True, D == 2)
    st = LN.LinearStage(kind, 0.3, metric, bins=20, sampling="Regular", percentage=0.25, iterations=(10, 6),
                        shrink=(2, 1), sigma=(0.0, 0.0), seed=5, preprocess=False)
    res = LN.register_linear(st, F, gF, M, gM, mi, fi, c)
    P = LO.LinearParams(kind=KINDS[kind], gradient_step=0.3, metric=METRICS[metric], bins=20, sampling=LO.REGULAR,
                        percentage=0.25, seed=5, iterations=(10, 6), shrink=(2, 1), sigma=(0.0, 0.0))
    ref = LO.linear_register(P, og(gF), F, og(gM), M, [aff_tuple(t) for t in mi], [aff_tuple(t) for t in fi], c)
    print(f"\n{D}-D {kind}/{metric}: iterations {res.iterations} / {ref.iterations}, "
          f"|dp| {np.abs(res.params - ref.params).max():.2e}")
    assert res.iterations == ref.iterations
    for a, b in zip(res.records, ref.records):
        assert a[:2] == b[:2]
        assert abs(a[2] - b[2]) <= 1e-10 * max(1.0, abs(b[2]))
        assert abs(a[4] - b[4]) <= 1e-8 * max(1.0, abs(b[4])) or max(a[4], b[4]) < 1e-6, (a, b)
    np.testing.assert_allclose(res.params, ref.params, rtol=0, atol=1e-9)


@pytest.mark.parametrize("kind,metric,tol", [("Rigid", "MeanSquares", 0.1), ("Rigid", "Mattes", 0.1),
                                             ("Affine", "GC", 0.15), ("Affine", "Mattes", 0.3)])
def test_recovers_known_transform(kind, metric, tol):
    """A moving image made from the fixed one by a known rotation and
    translation about c: the stage finds them to within tol mm over the
    middle half of the grid.  The oracle's runs of the same cases land at
    0.027, 0.025, 0.083 and 0.24 mm: the 12-parameter fits stop where the
    line search no longer finds descent along the central-difference
    gradient, the Mattes one on a texture with a large flat background."""
    D = 3
    gF, F = blob(D, 2)
    A = rot(D, 0.08)
    t = np.array([2.0, -1.5, 1.0])
    c = np.asarray(gF.origin) + np.asarray(gF.spacing) * (np.asarray(gF.size) - 1) / 2
    M = moved(gF, F, A, t, c)
    st = LN.LinearStage(kind, 0.5, metric, bins=32, sampling="Regular", percentage=0.5, iterations=(60, 40, 30),
                        shrink=(4, 2, 1), sigma=(2.0, 1.0, 0.0), seed=1, threshold=1e-8)
    res = LN.register_linear(st, F, gF, M, gF, (), (), c)
    Ar, tr, cr = res.transform.parameters()
    pts = LO.corners(og(gF)) * 0.5 + c * 0.5
    want = (pts - c) @ A.T + c + t
    got = (pts - cr) @ Ar.T + cr + tr
    err = np.abs(want - got).max()
    print(f"\n{kind}/{metric}: iterations {res.iterations}, max point error {err:.4f} mm")
    assert err < tol


def test_cli_linear_stages_then_syn(tmp_path):
    """antsRegistrationSyN.sh's stage layout from files: -r [f,m,1], Rigid,
    Affine, then SyN.  The output composite (0GenericAffine.mat, 1Warp,
    1InverseWarp) equals the library run stage by stage on the same arrays."""
    from ants_amd import cli
    from ants_amd import io as aio
    from ants_amd.apply_transforms import compose_affines
    from ants_amd.registration import SyNRegistration, SyNStage, initial_transform_from_images
    gF, F = blob(3, 4)
    c = np.asarray(gF.origin) + np.asarray(gF.spacing) * (np.asarray(gF.size) - 1) / 2
    M = moved(gF, F, rot(3, 0.06) * 1.03, np.array([1.5, -1.0, 0.5]), c)
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, F, gF)
    aio.write_nifti(m, M, gF)
    F, gf = aio.read_nifti(f)
    M, gm = aio.read_nifti(m)
    prefix = str(tmp_path / "out_")
    rigid = ["Rigid[0.2]", f"GC[{f},{m},1,1,Regular,0.5]", "[20x10,1e-6,10]", "2x1", "1x0vox"]
    affine = ["Affine[0.2]", f"Mattes[{f},{m},1,32,Regular,0.5]", "[10x5,1e-6,10]", "2x1", "1x0vox"]
    syn = ["SyN[0.2,3,0]", f"CC[{f},{m},1,2]", "[5x3,1e-6,10]", "2x1", "1x0vox"]
    args = ["-d", "3", "-r", f"[{f},{m},1]", "--random-seed", "3"]
    for t, mm, cc, ff, ss in (rigid, affine, syn):
        args += ["-t", t, "-m", mm, "-c", cc, "-f", ff, "-s", ss]
    args += ["-o", f"[{prefix},{prefix}Warped.nii.gz]"]
    assert cli.main(["registration"] + args) == 0
    init = initial_transform_from_images(F, gf, M, gm, 1)
    chain = [init]
    for t, mm, cc, ff, ss in (rigid, affine):
        st = LN.LinearStage.from_args(t, mm, cc, ff, ss, seed=3)
        r = LN.register_linear(st, F, gf, M, gm, chain)
        chain = [r.transform] + chain
    want = compose_affines(chain)
    got = aio.read_transform(prefix + "0GenericAffine.mat")
    pts = LO.corners(og(gf))
    Aw, tw, cw = want.parameters()
    Ag, tg, cg = got.parameters()
    np.testing.assert_allclose((pts - cg) @ Ag.T + cg + tg, (pts - cw) @ Aw.T + cw + tw, rtol=0, atol=1e-9)
    reg = SyNRegistration(SyNStage.from_args(*syn), 3)
    reg.set_initial_transforms(chain, [])
    res = reg.register(F, M, gf.spacing, gf.origin, gm.spacing, gm.origin, gf.direction, gm.direction)
    reg.close()
    fwd, _ = aio.read_nifti(prefix + "1Warp.nii.gz")
    np.testing.assert_array_equal(fwd, res.fwd)
    # a linear-only run writes the transform alone
    p2 = str(tmp_path / "lin_")
    args = ["-d", "3", "-t", "Translation[0.5]", "-m", f"MeanSquares[{f},{m},1,0,None,1]", "-c", "[15,1e-6,10]",
            "-f", "1", "-s", "0vox", "-o", p2]
    assert cli.main(["registration"] + args) == 0
    st = LN.LinearStage.from_args("Translation[0.5]", f"MeanSquares[{f},{m},1,0,None,1]", "[15,1e-6,10]", "1", "0vox")
    r = LN.register_linear(st, F, gf, M, gm)
    got = aio.read_transform(p2 + "0GenericAffine.mat")
    np.testing.assert_allclose(got.parameters()[1], r.transform.parameters()[1], rtol=0, atol=1e-12)


def test_point_cache_reuse_is_exact():
    """Sample points kept on the device between stages are the points a
    fresh stage makes: a stage run cold (cache released), again warm, and
    one with another seed in between (its own points, not the first seed's)
    give the first stage's parameters bit for bit, and the other seed gives
    a different result."""
    D = 3
    gF, F, gM, M, mi, fi, c, _ = setup(D, "Rigid", "Mattes", "Regular", 0.25, 1, 0.0)
    st = LN.LinearStage("Rigid", 0.25, "Mattes", bins=24, sampling="Random", percentage=0.3, iterations=(6, 4),
                        shrink=(2, 1), sigma=(1.0, 0.0), seed=21, preprocess=True)
    other = LN.LinearStage("Rigid", 0.25, "Mattes", bins=24, sampling="Random", percentage=0.3, iterations=(6, 4),
                           shrink=(2, 1), sigma=(1.0, 0.0), seed=22, preprocess=True)
    LN.release_points()
    cold = LN.register_linear(st, F, gF, M, gM, mi, (), c)
    oth = LN.register_linear(other, F, gF, M, gM, mi, (), c)
    warm = LN.register_linear(st, F, gF, M, gM, mi, (), c)
    LN.release_points()
    again = LN.register_linear(st, F, gF, M, gM, mi, (), c)
    np.testing.assert_array_equal(warm.params, cold.params)
    np.testing.assert_array_equal(again.params, cold.params)
    assert warm.records == cold.records
    assert not np.array_equal(oth.params, cold.params)
