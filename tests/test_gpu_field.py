"""Dense field-transform stages (GaussianDisplacementField, Exponential with
the CC metric) on the GPU against the oracle's restatement
(oracle/field_oracle.py): the metric pass (value and derivative field at a
given displacement, with initial affines and oblique grids), the scaling-
and-squaring exponential, and whole stages (the same records within the
parity bars, fields within 0.05 voxel), then recovery of a known shift."""
import numpy as np
import pytest
from scipy import ndimage

import oracle as O
import field_oracle as FO
from ants_amd import fieldmodels as FM
from ants_amd.apply_transforms import AffineTransform, Geometry

pytestmark = pytest.mark.gpu


def rot(D, a):
    if D == 2:
        return np.array([[np.cos(a), -np.sin(a)], [np.sin(a), np.cos(a)]])
    return np.array([[np.cos(a), -np.sin(a), 0], [np.sin(a), np.cos(a), 0], [0, 0, 1.0]])


def og(g):
    return O.Grid(g.size, g.spacing, g.origin, g.direction)


def images(D, seed, oblique=False, size=None):
    rng = np.random.default_rng(seed)
    size = size or (34, 30, 26)[:D]
    g = Geometry(size, (1.1, 0.9, 1.2)[:D], (4.0, -3.0, 2.0)[:D], rot(D, 0.15) if oblique else np.eye(D))
    F = ndimage.gaussian_filter(rng.standard_normal(g.shape), 2.0)
    M = ndimage.shift(F, [0.8, -0.6, 0.5][:D], order=1) + 0.05 * ndimage.gaussian_filter(rng.standard_normal(g.shape), 2.0)
    F = ((F - F.min()) / (F.max() - F.min())).astype(np.float32)
    M = ((M - M.min()) / (M.max() - M.min())).astype(np.float32)
    return g, F, M


def smooth_field(shape, D, seed, amp):
    rng = np.random.default_rng(seed)
    u = np.stack([ndimage.gaussian_filter(rng.standard_normal(shape), 3.0) for _ in range(D)], axis=-1)
    return (amp * u / np.abs(u).max()).astype(np.float32)


def chain(ts, D):
    return [t.parameters() for t in ts]


CASES = [
    # D, shrink, sigma, radius, moving initial, fixed initial, oblique
    (3, 1, 0.0, 4, False, False, False),
    (3, 2, 1.0, 2, True, False, False),
    (3, 1, 0.0, 2, True, True, True),
    (2, 1, 0.0, 4, False, False, False),
    (2, 2, 1.0, 3, True, True, True),
]


@pytest.mark.parametrize("case", CASES, ids=[f"{c[0]}d-s{c[1]}-r{c[3]}-mi{int(c[4])}-fi{int(c[5])}-ob{int(c[6])}"
                                             for c in CASES])
def test_metric_pass_matches_oracle(case):
    D, shrink, sigma, radius, mi, fi, obl = case
    g, F, M = images(D, 1, obl)
    c = np.asarray(g.origin) + 8.0
    minit = [AffineTransform(rot(D, 0.04) * 1.01, np.full(D, 0.3), c)] if mi else []
    finit = [AffineTransform(np.eye(D) * 0.99, np.full(D, -0.2), c)] if fi else []
    st = FM.FieldStage(FM.GAUSSIAN_DISPLACEMENT, radius=radius, iterations=(1,), shrink=(shrink,), sigma=(sigma,),
                       preprocess=False)
    shape = FM.level_shape(st, g, 0)
    u = smooth_field(shape, D, 3, 1.5)
    v, d = FM.field_metric(st, 0, F, g, M, g, u, minit, finit)
    gl = O.shrink_grid(og(g), shrink)
    Fs = O.smooth_image(og(g), F.astype(np.float64), sigma) if sigma > 0 else F.astype(np.float64)
    Ms = O.smooth_image(og(g), M.astype(np.float64), sigma) if sigma > 0 else M.astype(np.float64)
    import linear_oracle as LO
    m = FO._Metric(gl, og(g), Fs, og(g), Ms, radius, LO.affine_chain(chain(minit, D), D),
                   LO.affine_chain(chain(finit, D), D))
    rv, rd = m(u.astype(np.float64))
    print(f"\n{case}: value {v:.10f} vs {rv:.10f}, |dd| {np.abs(d - rd).max():.2e} of {np.abs(rd).max():.2e}")
    assert abs(v - rv) <= 1e-6 * abs(rv)
    np.testing.assert_allclose(d, rd, rtol=0, atol=2e-4 * np.abs(rd).max())
    v0, _ = FM.field_metric(st, 0, F, g, M, g, u, minit, finit, derivative=False)
    assert v0 == v


@pytest.mark.parametrize("D,steps", [(3, 0), (3, 5), (2, 0)])
def test_exponential_matches_oracle(D, steps):
    g = Geometry((40, 36, 30)[:D], (1.0, 0.8, 1.25)[:D], (1.0, 2.0, 3.0)[:D], np.eye(D))
    v = smooth_field(g.shape, D, 7, 4.0)
    for inverse in (False, True):
        out, n = FM.exponential(g, v, steps, inverse)
        ref, rn = FO.exponentiate(og(g), v.astype(np.float64), steps, inverse)
        print(f"\nD={D} steps={steps} inverse={inverse}: {n} squarings, |d| {np.abs(out - ref).max():.2e}")
        assert n == rn
        np.testing.assert_allclose(out, ref, rtol=0, atol=2e-4)


def run_both(model, D, iterations, shrink, sigma, lr, uvar, tvar, seed=2, steps=0, minit=(), umesh=(0, 0, 0),
             tmesh=(0, 0, 0), order=3):
    g, F, M = images(D, seed)
    st = FM.FieldStage(model, lr, uvar, tvar, steps, 2, iterations, 1e-6, 10, shrink, sigma, False, True, True,
                       update_mesh=umesh, total_mesh=tmesh, spline_order=order)
    res = FM.register_field(st, F, g, M, g, minit)
    P = FO.FieldParams(model=model, learning_rate=lr, update_variance=uvar, total_variance=tvar,
                       integration_steps=steps, radius=2, iterations=iterations, shrink=shrink, sigma=sigma,
                       force_iterations=True, update_mesh=umesh, total_mesh=tmesh, spline_order=order)
    ref = FO.field_register(P, og(g), F, og(g), M, chain(minit, D))
    return g, res, ref


STAGES = [(FM.GAUSSIAN_DISPLACEMENT, 3), (FM.EXPONENTIAL, 3), (FM.GAUSSIAN_DISPLACEMENT, 2), (FM.EXPONENTIAL, 2),
          (FM.BSPLINE_DISPLACEMENT, 3), (FM.BSPLINE_EXPONENTIAL, 3), (FM.BSPLINE_DISPLACEMENT, 2),
          (FM.BSPLINE_EXPONENTIAL, 2)]


@pytest.mark.parametrize("model,D", STAGES, ids=[f"model{m}-{d}d" for m, d in STAGES])
def test_stage_matches_oracle(model, D):
    """Two levels of the stage with preprocessing, pyramid smoothing and
    both field smoothings (Gaussian, or B-spline fits of orders 2 and 3 on
    meshes doubled per level): the same records (levels, iterations), values
    within 1e-4 relative, learning rates within 1e-3 relative, the final
    fields within 0.05 voxel (mean within 0.005)."""
    expo = model in (FM.EXPONENTIAL, FM.BSPLINE_EXPONENTIAL)
    lr = 15.0 if expo else 0.4
    c = np.array([10.0, 8.0, 6.0][:D])
    minit = [AffineTransform(rot(D, 0.02), np.full(D, 0.2), c)] if D == 3 else []
    order = 2 if D == 2 else 3
    # B-spline forms: update meshes of 4 spans per axis (8 and 16 on the two
    # levels); the exponential form also fits its velocity field (2 spans)
    umesh = (4, 4, 4)[:D]
    tmesh = (2, 2, 2)[:D] if model == FM.BSPLINE_EXPONENTIAL else (0, 0, 0)[:D]
    g, res, ref = run_both(model, D, (6, 4), (2, 1), (1.0, 0.0), lr, 2.0, 0.5, minit=minit, umesh=umesh,
                           tmesh=tmesh, order=order)
    assert res.iterations == ref.iterations
    assert len(res.records) == len(ref.records)
    for a, b in zip(res.records, ref.records):
        assert a[:2] == b[:2]
        assert abs(a[2] - b[2]) <= 1e-4 * abs(b[2]), (a, b)
        assert abs(a[4] - b[4]) <= 1e-3 * abs(b[4]), (a, b)
    sp = np.asarray(g.spacing)
    dev = np.abs(res.field - ref.field) / sp
    print(f"\nmodel {model} D={D}: energies {ref.records[0][2]:.5f} -> {ref.records[-1][2]:.5f}; field max dev "
          f"{dev.max():.2e}, mean {dev.mean():.2e} voxel")
    assert dev.max() < 0.05 and dev.mean() < 0.005
    if expo:
        assert np.abs(res.inverse - ref.inverse).max() / sp.min() < 0.05
        assert np.abs(res.velocity - ref.velocity).max() / sp.min() < 0.05


@pytest.mark.parametrize("model", [FM.GAUSSIAN_DISPLACEMENT, FM.EXPONENTIAL])
def test_stage_recovers_translation(model):
    """A blob moved by a known sub-voxel translation: the field over the
    blob's core approaches that translation."""
    rng = np.random.default_rng(3)
    g = Geometry((32, 30, 28), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), np.eye(3))
    img = np.zeros(g.shape)
    inner = tuple(slice(s // 4, s - s // 4) for s in g.shape)
    img[inner] = rng.standard_normal(img[inner].shape)
    F = ndimage.gaussian_filter(img, 1.5)
    shift = np.array([1.2, -0.8, 0.6])
    M = ndimage.shift(F, shift, order=1)
    lr = 0.5 if model == FM.GAUSSIAN_DISPLACEMENT else 20.0
    st = FM.FieldStage(model, lr, 2.0, 0.5, 0, 2, (30, 20), 1e-6, 10, (2, 1), (1.0, 0.0), False, True, True)
    res = FM.register_field(st, F.astype(np.float32), g, M.astype(np.float32), g)
    core = tuple(slice(s // 2 - 3, s // 2 + 3) for s in g.shape)
    mean = res.field[core].reshape(-1, 3).mean(axis=0)
    print(f"\nmodel {model}: mean core displacement {mean}, want {shift[::-1]}")
    assert np.abs(mean - shift[::-1]).max() < 0.35


@pytest.mark.parametrize("model", ["GaussianDisplacementField[0.4,2,0.5]", "Exponential[15,2,0.5,0]",
                                   "BSplineExponential[15,3x3x2,2x2x1,0,3]"])
def test_cli_field_stage(tmp_path, model):
    """cli registration with a rigid stage and then a field stage writes
    0GenericAffine.mat and 1Warp.nii.gz (Exponential also 1InverseWarp),
    equal to the library run on the same arrays; the warped image is the
    moving image through [field, affine]."""
    from ants_amd import cli
    from ants_amd import io as aio
    from ants_amd import linear as LN
    from ants_amd.apply_transforms import DisplacementFieldTransform, apply_transforms
    g, F, M = images(3, 5)
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, F, g)
    aio.write_nifti(m, M, g)
    F, gf = aio.read_nifti(f)
    M, gm = aio.read_nifti(m)
    prefix = str(tmp_path / "out_")
    rigid = ["Rigid[0.2]", f"GC[{f},{m},1,1,Regular,0.5]", "[6,1e-6,10]", "1", "0vox"]
    fld = [model, f"CC[{f},{m},1,2]", "[4x3,1e-6,10]", "2x1", "1x0vox"]
    args = ["-d", "3", "--random-seed", "2"]
    for t, mm, cc, ff, ss in (rigid, fld):
        args += ["-t", t, "-m", mm, "-c", cc, "-f", ff, "-s", ss]
    args += ["-o", f"[{prefix},{prefix}Warped.nii.gz]"]
    assert cli.main(["registration"] + args) == 0
    st = LN.LinearStage.from_args(*rigid, seed=2)
    r = LN.register_linear(st, F, gf, M, gm)
    fs = FM.FieldStage.from_args(*fld)
    res = FM.register_field(fs, F, gf, M, gm, [r.transform])
    fwd, _ = aio.read_nifti(prefix + "1Warp.nii.gz")
    np.testing.assert_array_equal(fwd, res.field)
    if fs.expo:
        inv, _ = aio.read_nifti(prefix + "1InverseWarp.nii.gz")
        np.testing.assert_array_equal(inv, res.inverse)
    else:
        import os
        assert not os.path.exists(prefix + "1InverseWarp.nii.gz")
    warped, _ = aio.read_nifti(prefix + "Warped.nii.gz")
    want = apply_transforms(gf, M, gm, [DisplacementFieldTransform(res.field, gf), r.transform], "Linear")
    np.testing.assert_array_equal(warped, want)
