"""The command lines end to end from NIfTI files: registration writes the
ANTs-named warp files, apply reads them back; results equal the library
calls on the same arrays."""
import numpy as np
import pytest

from ants_amd import cli
from ants_amd import io as aio
from ants_amd import phantoms as ph
from ants_amd.apply_transforms import DisplacementFieldTransform, Geometry, apply_transforms
from ants_amd.registration import SyNRegistration, SyNStage

pytestmark = pytest.mark.gpu


def test_registration_and_apply_from_files(tmp_path):
    pair = ph.phantom_pair_3d(2, size=(48, 56, 40), max_disp_mm=4.0, svf_sigma_mm=6.0)
    g = Geometry(pair.size, pair.spacing, (1.0, -2.0, 3.0), np.eye(3))
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, g)
    aio.write_nifti(str(tmp_path / "labels.nii.gz"), pair.labels_moving.astype(np.int16), g)
    prefix = str(tmp_path / "out_")
    args = ["-d", "3", "-m", f"CC[{f},{m},1,2]", "-t", "SyN[0.2,3,0]", "-c", "[10x5x2,1e-6,10]", "-f", "4x2x1",
            "-s", "2x1x0vox", "-o", f"[{prefix},{prefix}Warped.nii.gz]"]
    assert cli.main(["registration"] + args) == 0
    fwd, gw = aio.read_nifti(prefix + "0Warp.nii.gz")
    inv, _ = aio.read_nifti(prefix + "0InverseWarp.nii.gz")
    assert fwd.shape == g.shape + (3,) and gw.size == g.size
    # the library on the same arrays gives the same fields
    st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[10x5x2,1e-6,10]", "4x2x1", "2x1x0vox")
    reg = SyNRegistration(st, 3)
    res = reg.register(pair.fixed, pair.moving, g.spacing, g.origin, g.spacing, g.origin)
    reg.close()
    np.testing.assert_array_equal(fwd, res.fwd)
    np.testing.assert_array_equal(inv, res.inv)
    warped, _ = aio.read_nifti(prefix + "Warped.nii.gz")
    np.testing.assert_array_equal(warped, apply_transforms(g, pair.moving, g, [DisplacementFieldTransform(fwd, g)]))
    # apply: labels through the warp, GenericLabel, integer type kept
    out = str(tmp_path / "labels_warped.nii.gz")
    assert cli.main(["apply", "-d", "3", "-i", str(tmp_path / "labels.nii.gz"), "-r", f, "-o", out,
                     "-t", prefix + "0Warp.nii.gz", "-n", "GenericLabel[Linear]"]) == 0
    lw, _ = aio.read_nifti(out)
    assert lw.dtype == np.int16
    ref = apply_transforms(g, pair.labels_moving.astype(np.float32), g, [DisplacementFieldTransform(fwd, g)],
                           "GenericLabel")
    np.testing.assert_array_equal(lw, ref.astype(np.int16))
    assert set(np.unique(lw)) <= set(np.unique(pair.labels_moving))


def test_registration_with_initial_moving_transform(tmp_path):
    """-r affine.mat: the stage runs with the moving initial transform, the
    outputs are 0GenericAffine.mat + 1Warp / 1InverseWarp, and the warped
    image goes through [warp, affine]."""
    from ants_amd.apply_transforms import AffineTransform
    pair = ph.phantom_pair_3d(4, size=(40, 44, 36), max_disp_mm=4.0, svf_sigma_mm=6.0)
    g = Geometry(pair.size, pair.spacing, (0.0, 0.0, 0.0), np.eye(3))
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, g)
    c = np.cos(0.05), np.sin(0.05)
    A = np.array([[c[0], -c[1], 0.0], [c[1], c[0], 0.0], [0.0, 0.0, 1.02]])
    aff = AffineTransform(A, [0.6, -0.3, 0.4], [20.0, 22.0, 18.0])
    aio.write_transform(str(tmp_path / "init.mat"), aff)
    prefix = str(tmp_path / "out_")
    args = ["-d", "3", "-m", f"CC[{f},{m},1,2]", "-t", "SyN[0.2,3,0]", "-c", "[6x4x2,1e-6,10]", "-f", "4x2x1",
            "-s", "2x1x0vox", "-r", str(tmp_path / "init.mat"),
            "-o", f"[{prefix},{prefix}Warped.nii.gz,{prefix}InverseWarped.nii.gz]"]
    assert cli.main(["registration"] + args) == 0
    fwd, _ = aio.read_nifti(prefix + "1Warp.nii.gz")
    st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[6x4x2,1e-6,10]", "4x2x1", "2x1x0vox")
    reg = SyNRegistration(st, 3)
    reg.set_initial_transforms([aff])
    res = reg.register(pair.fixed, pair.moving, g.spacing, g.origin, g.spacing, g.origin)
    reg.close()
    np.testing.assert_array_equal(fwd, res.fwd)
    back = aio.read_transform(prefix + "0GenericAffine.mat")
    pts = np.random.default_rng(0).uniform(0, 40, (5, 3))
    A0, t0, c0 = aff.parameters()
    A1, t1, c1 = back.parameters()
    np.testing.assert_allclose([A1 @ (p - c1) + t1 + c1 for p in pts], [A0 @ (p - c0) + t0 + c0 for p in pts],
                               atol=1e-9)
    warped, _ = aio.read_nifti(prefix + "Warped.nii.gz")
    np.testing.assert_array_equal(warped, apply_transforms(g, pair.moving, g,
                                                           [DisplacementFieldTransform(fwd, g), aff]))


def test_registration_with_initial_fixed_transform(tmp_path):
    """-q affine.mat: the fixed image's points go through the transform
    (SetFixedInitialTransform); the warp files equal the library's."""
    from ants_amd.apply_transforms import AffineTransform
    pair = ph.phantom_pair_3d(5, size=(36, 40, 32), max_disp_mm=3.0, svf_sigma_mm=6.0)
    g = Geometry(pair.size, pair.spacing, (0.0, 0.0, 0.0), np.eye(3))
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, g)
    aff = AffineTransform(np.diag([1.02, 0.99, 1.0]), [0.3, -0.2, 0.1], [18.0, 20.0, 16.0])
    aio.write_transform(str(tmp_path / "q.mat"), aff)
    prefix = str(tmp_path / "out_")
    args = ["-d", "3", "-m", f"CC[{f},{m},1,2]", "-t", "SyN[0.2,3,0]", "-c", "[4x2,1e-6,10]", "-f", "2x1",
            "-s", "1x0vox", "-q", str(tmp_path / "q.mat"), "-o", f"[{prefix}]"]
    assert cli.main(["registration"] + args) == 0
    fwd, _ = aio.read_nifti(prefix + "0Warp.nii.gz")
    st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[4x2,1e-6,10]", "2x1", "1x0vox")
    reg = SyNRegistration(st, 3)
    reg.set_initial_transforms([], [aff])
    res = reg.register(pair.fixed, pair.moving, g.spacing, g.origin, g.spacing, g.origin)
    reg.close()
    np.testing.assert_array_equal(fwd, res.fwd)


def test_registration_bspline_syn_knot_spacing(tmp_path):
    """`-t BSplineSyN[0.1,26,0,3]` (antsRegistrationSyN.sh -t b): the knot
    spacing becomes a mesh with the fixed image's geometry; the warp files
    equal the library run with that mesh."""
    pair = ph.phantom_pair_3d(4, size=(80, 88, 72), spacing=(1.0, 1.2, 1.0), max_disp_mm=4.0, svf_sigma_mm=8.0)
    g = Geometry(pair.size, pair.spacing, (0.0, 0.0, 0.0), np.eye(3))
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, g)
    prefix = str(tmp_path / "b_")
    args = ["-d", "3", "-m", f"CC[{f},{m},1,2]", "-t", "BSplineSyN[0.1,26,0,3]", "-c", "[8x4x2,1e-6,10]",
            "-f", "4x2x1", "-s", "2x1x0vox", "-o", prefix]
    assert cli.main(["registration"] + args) == 0
    fwd, _ = aio.read_nifti(prefix + "0Warp.nii.gz")
    _, gr = aio.read_nifti(f)           # the geometry as stored (float32 header fields)
    mesh = [int(np.ceil((n - 1) * s / 26)) for n, s in zip(gr.size, gr.spacing)]
    st = SyNStage.from_args(f"BSplineSyN[0.1,{'x'.join(map(str, mesh))},0,3]", "CC[f,m,1,2]", "[8x4x2,1e-6,10]",
                            "4x2x1", "2x1x0vox")
    reg = SyNRegistration(st, 3)
    res = reg.register(pair.fixed, pair.moving, gr.spacing, gr.origin, gr.spacing, gr.origin, gr.direction,
                       gr.direction)
    reg.close()
    np.testing.assert_array_equal(fwd, res.fwd)
    assert np.abs(fwd).max() > 0.05


def test_composite_outputs_and_state_files(tmp_path):
    """-a 1 writes Composite.h5 / InverseComposite.h5 that apply reads back
    (equal to the warp fields); --save-state after -f 4x2 and --restore-state
    for -f 1 reproduce the single 4x2x1 stage bit for bit."""
    from ants_amd import h5transform
    pair = ph.phantom_pair_3d(6, size=(48, 40, 32), max_disp_mm=3.0, svf_sigma_mm=6.0)
    g = Geometry(pair.size, pair.spacing, (0.0, 0.0, 0.0), np.eye(3))
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, g)
    base = ["-d", "3", "-m", f"CC[{f},{m},1,2]", "-t", "SyN[0.2,3,0]"]
    # single stage, composite outputs
    one = str(tmp_path / "one_")
    assert cli.main(["registration"] + base + ["-c", "[6x4x3,1e-6,10]", "-f", "4x2x1", "-s", "2x1x0vox",
                                                "-o", one, "-a", "1"]) == 0
    comp = h5transform.read_composite(one + "Composite.h5")
    icomp = h5transform.read_composite(one + "InverseComposite.h5")
    assert len(comp) == 1 and len(icomp) == 1
    ref = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[6x4x3,1e-6,10]", "4x2x1", "2x1x0vox")
    reg = SyNRegistration(ref, 3)
    res = reg.register(pair.fixed, pair.moving, g.spacing, g.origin, g.spacing, g.origin)
    reg.close()
    np.testing.assert_array_equal(comp[0].field, res.fwd)
    np.testing.assert_array_equal(icomp[0].field, res.inv)
    out = str(tmp_path / "w.nii.gz")
    assert cli.main(["apply", "-d", "3", "-i", m, "-r", f, "-o", out, "-t", one + "Composite.h5"]) == 0
    w, _ = aio.read_nifti(out)
    np.testing.assert_array_equal(w, apply_transforms(g, pair.moving, g, [DisplacementFieldTransform(res.fwd, g)]))
    # the same stage split in two through a state file
    a_ = str(tmp_path / "a_")
    st = str(tmp_path / "state.h5")
    assert cli.main(["registration"] + base + ["-c", "[6x4,1e-6,10]", "-f", "4x2", "-s", "2x1vox", "-o", a_,
                                                "--save-state", st]) == 0
    lin, fields, gst = h5transform.read_syn_state(st)
    assert lin == [] and len(fields) == 4 and tuple(gst.size) == (24, 20, 16)
    b_ = str(tmp_path / "b_")
    assert cli.main(["registration"] + base + ["-c", "[3,1e-6,10]", "-f", "1", "-s", "0vox", "-o", b_,
                                                "--restore-state", st]) == 0
    fwd_b, _ = aio.read_nifti(b_ + "0Warp.nii.gz")
    np.testing.assert_array_equal(fwd_b, res.fwd)
    with pytest.raises(SystemExit):
        cli.main(["registration"] + base + ["-c", "[3,1e-6,10]", "-f", "1", "-s", "0vox", "-o", b_,
                                            "--restore-state", st, "-r", one + "Composite.h5"])


def test_registration_initial_transform_from_images(tmp_path):
    """-r [fixed,moving,1]: the centre-of-mass initial transform (written
    composed into 0GenericAffine.mat) equals initial_transform_from_images,
    and the stage runs with it."""
    from ants_amd.registration import initial_transform_from_images
    pair = ph.phantom_pair_3d(7, size=(40, 44, 36), max_disp_mm=3.0, svf_sigma_mm=6.0)
    g = Geometry(pair.size, pair.spacing, (0.0, 0.0, 0.0), np.eye(3))
    gm = Geometry(pair.size, pair.spacing, (2.5, -1.5, 1.0), np.eye(3))
    f, m = str(tmp_path / "fixed.nii.gz"), str(tmp_path / "moving.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, gm)
    prefix = str(tmp_path / "c_")
    assert cli.main(["registration", "-d", "3", "-m", f"CC[{f},{m},1,2]", "-t", "SyN[0.2,3,0]", "-c", "[4x2,1e-6,10]",
                     "-f", "2x1", "-s", "1x0vox", "-o", prefix, "-r", f"[{f},{m},1]"]) == 0
    aff = aio.read_transform(prefix + "0GenericAffine.mat")
    fi, gfi = aio.read_nifti(f)
    mi, gmi = aio.read_nifti(m)
    ref = initial_transform_from_images(np.asarray(fi, np.float32), gfi, np.asarray(mi, np.float32), gmi, 1)
    A, t, c = ref.parameters()
    p = np.array([3.0, -4.0, 5.0])
    np.testing.assert_allclose(aff.matrix @ (p - aff.center) + aff.translation + aff.center, A @ (p - c) + t + c,
                               atol=1e-5)
    assert np.abs(t).max() > 1.0


def test_apply_transform_outputs(tmp_path):
    """apply -o [field,1] / DisplacementField[x.h5] / CompositeTransform[x.h5]:
    the written field equals compose_transforms_to_field of the same list,
    the .h5 forms read back to the same field and transforms, and resampling
    through either equals resampling through the list."""
    from ants_amd.apply_transforms import AffineTransform, compose_transforms_to_field
    from ants_amd.h5transform import read_composite
    rng = np.random.default_rng(8)
    g = Geometry((30, 28, 24), (1.2, 1.0, 1.1), (2.0, -3.0, 1.0), np.eye(3))
    ref = str(tmp_path / "ref.nii.gz")
    img = rng.standard_normal(g.shape).astype(np.float32)
    aio.write_nifti(ref, img, g)
    fld = (0.8 * rng.standard_normal(g.shape + (3,))).astype(np.float32)
    fpath, apath = str(tmp_path / "w.nii.gz"), str(tmp_path / "a.mat")
    aio.write_nifti(fpath, fld, g, vector=True)
    aff = AffineTransform(np.eye(3) * 1.02, [0.5, -0.4, 0.3], [15.0, 12.0, 10.0])
    aio.write_transform(apath, aff)
    # the geometries as the files hold them (NIfTI stores float32 headers)
    _, g = aio.read_nifti(ref)
    xf = cli.load_transforms(fpath) + cli.load_transforms(apath)
    want = compose_transforms_to_field(g, xf)
    out = str(tmp_path / "composite_field.nii.gz")
    assert cli.main(["apply", "-d", "3", "-r", ref, "-o", f"[{out},1]", "-t", fpath, "-t", apath]) == 0
    got, gg = aio.read_nifti(out)
    np.testing.assert_array_equal(got, want)
    assert gg.size == g.size
    h5 = str(tmp_path / "composite_field.h5")
    assert cli.main(["apply", "-d", "3", "-r", ref, "-o", f"DisplacementField[{h5}]", "-t", fpath, "-t", apath]) == 0
    (t,) = read_composite(h5)
    np.testing.assert_array_equal(np.asarray(t.field), want)
    comp = str(tmp_path / "composite.h5")
    assert cli.main(["apply", "-d", "3", "-o", f"CompositeTransform[{comp}]", "-t", fpath, "-t", apath]) == 0
    warped = str(tmp_path / "warped.nii.gz")
    assert cli.main(["apply", "-d", "3", "-i", ref, "-r", ref, "-o", warped, "-t", comp]) == 0
    a, _ = aio.read_nifti(warped)
    # the composite lists the same transforms in the same order (its affine
    # parameters stored as float, as --float 1 writes them)
    xc = cli.load_transforms(comp)
    assert [type(t) for t in xc] == [type(t) for t in xf]
    np.testing.assert_array_equal(np.asarray(xc[0].field), np.asarray(xf[0].field))
    for u, v in zip(xc[1].parameters(), xf[1].parameters()):
        np.testing.assert_allclose(u, v, rtol=1e-6, atol=1e-6)
    np.testing.assert_array_equal(a, apply_transforms(g, img, g, xc))


def test_apply_points_csv(tmp_path):
    """apply-points: a CSV with x,y,z,t,label,comment columns through
    [affine, inverse affine] comes back unchanged, and through a field and an
    affine equals the library call; the header and extra columns are kept."""
    from ants_amd.apply_transforms import AffineTransform, apply_transforms_to_points
    rng = np.random.default_rng(9)
    pin = str(tmp_path / "pts.csv")
    pts = np.column_stack([rng.uniform(0, 20, (40, 3)), np.zeros(40), np.arange(40.0)])
    with open(pin, "w") as fh:
        fh.write("x,y,z,t,label,comment\n")
        for r in pts:
            fh.write(",".join(repr(float(v)) for v in r) + ",pt\n")
    apath = str(tmp_path / "a.mat")
    aio.write_transform(apath, AffineTransform(np.eye(3) * 1.1, [1.0, -2.0, 0.5], [10.0, 10.0, 10.0]))
    out = str(tmp_path / "out.csv")
    assert cli.main(["apply-points", "-d", "3", "-i", pin, "-o", out, "-t", apath, "-t", f"[{apath},1]"]) == 0
    names, got = aio.read_points_csv(out)
    assert names == ["x", "y", "z", "t", "label", "comment"]
    np.testing.assert_allclose(got[:, :5], pts, rtol=0, atol=1e-12)
    assert np.isnan(got[:, 5]).all()
    g = Geometry((24, 24, 24), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), np.eye(3))
    fld = (0.5 * rng.standard_normal(g.shape + (3,))).astype(np.float32)
    fpath = str(tmp_path / "w.nii.gz")
    aio.write_nifti(fpath, fld, g, vector=True)
    assert cli.main(["apply-points", "-d", "3", "-i", pin, "-o", out, "-t", fpath, "-t", apath]) == 0
    _, got = aio.read_points_csv(out)
    want = apply_transforms_to_points(got[:, :5] * 0 + pts, cli.load_transforms(fpath) + cli.load_transforms(apath), 3)
    np.testing.assert_array_equal(got[:, :5], want)


def test_jacobian_cli(tmp_path):
    """jacobian D field out [log] [0] [gradient]: files equal the library
    call; the gradient form is written as a D*D-component vector image."""
    from ants_amd.analysis import create_jacobian_determinant_image
    rng = np.random.default_rng(12)
    g = Geometry((20, 18, 16), (1.1, 1.0, 0.9), (0.0, 0.0, 0.0), np.eye(3))
    fld = (0.5 * rng.standard_normal(g.shape + (3,))).astype(np.float32)
    fpath = str(tmp_path / "w.nii.gz")
    aio.write_nifti(fpath, fld, g, vector=True)
    fld, g = aio.read_nifti(fpath)
    out = str(tmp_path / "jac.nii.gz")
    for log in (0, 1):
        assert cli.main(["jacobian", "3", fpath, out, str(log)]) == 0
        got, _ = aio.read_nifti(out)
        np.testing.assert_array_equal(got, create_jacobian_determinant_image(fld, g, do_log=bool(log)))
    assert cli.main(["jacobian", "3", fpath, out, "0", "0", "1"]) == 0
    got, _ = aio.read_nifti(out)
    want = create_jacobian_determinant_image(fld, g, deformation_gradient=True)
    np.testing.assert_array_equal(got, want.reshape(g.shape + (9,)))
    assert cli.main(["jacobian", "3", fpath, out, "0", "1"]) == 0
    got, _ = aio.read_nifti(out)
    np.testing.assert_array_equal(got, create_jacobian_determinant_image(fld, g, geometric=True))
    with pytest.raises(SystemExit):
        cli.main(["jacobian", "3", fpath, out, "0", "1", "1"])


def test_apply_time_series(tmp_path):
    """apply -e 3: each volume of a 4-D series through [field, affine] equals
    the scalar resample of that volume; the output has the reference's
    spatial grid and the input's time spacing and offset."""
    from ants_amd.apply_transforms import AffineTransform, apply_transforms_series
    rng = np.random.default_rng(13)
    gi = Geometry((22, 20, 18), (1.2, 1.1, 1.3), (1.0, 2.0, -3.0), np.eye(3))
    gr = Geometry((18, 16, 15), (1.4, 1.3, 1.5), (2.0, 3.0, -2.0), np.eye(3))
    series = rng.standard_normal((3,) + gi.shape).astype(np.float32)
    sp, rp = str(tmp_path / "ts.nii.gz"), str(tmp_path / "ref.nii.gz")
    aio.write_nifti_series(sp, series, gi, 2.0, 0.5)
    aio.write_nifti(rp, np.zeros(gr.shape, np.float32), gr)
    apath = str(tmp_path / "a.mat")
    aio.write_transform(apath, AffineTransform(np.eye(3) * 0.98, [0.4, -0.3, 0.2], [12.0, 12.0, 10.0]))
    out = str(tmp_path / "out.nii.gz")
    assert cli.main(["apply", "-d", "3", "-e", "3", "-i", sp, "-r", rp, "-o", out, "-t", apath,
                     "-n", "BSpline[3]"]) == 0
    got, g2, tsp, tor = aio.read_nifti_series(out)
    assert got.shape == (3,) + gr.shape and (tsp, tor) == (2.0, 0.5) and g2.size == gr.size
    s_arr, gs, _, _ = aio.read_nifti_series(sp)
    _, grf = aio.read_nifti(rp)
    xf = cli.load_transforms(apath)
    for k in range(3):
        np.testing.assert_array_equal(got[k], apply_transforms(grf, s_arr[k], gs, xf, "BSpline[3]"))
    np.testing.assert_array_equal(apply_transforms_series(grf, s_arr, gs, xf, "BSpline[3]"), got)


def test_similarity_cc(tmp_path, capsys):
    """similarity -m CC[f,m,1,r]: the printed value is the oracle's CC
    energy of the pair (same grid) for r = 2 and 4; a moving image on
    another grid is resampled onto the fixed grid first."""
    import oracle as o
    from ants_amd.analysis import measure_cc_similarity
    from ants_amd.apply_transforms import apply_transforms as AT_apply
    pair = ph.phantom_pair_3d(4, size=(40, 36, 32), max_disp_mm=3.0, svf_sigma_mm=5.0)
    g = Geometry(pair.size, pair.spacing, (0.0, 0.0, 0.0), np.eye(3))
    f, m = str(tmp_path / "f.nii.gz"), str(tmp_path / "m.nii.gz")
    aio.write_nifti(f, pair.fixed, g)
    aio.write_nifti(m, pair.moving, g)
    F, gF = aio.read_nifti(f)
    M, _ = aio.read_nifti(m)
    og_ = o.Grid(tuple(gF.size), tuple(gF.spacing), tuple(gF.origin))
    for r in (2, 4):
        capsys.readouterr()
        assert cli.main(["similarity", "-d", "3", "-m", f"CC[{f},{m},1,{r}]"]) == 0
        val = float(capsys.readouterr().out.strip())
        want, _, _ = o.cc(og_, F.astype(np.float64), M.astype(np.float64), r)
        assert val == pytest.approx(want, rel=1e-5)
    gm2 = Geometry((30, 28, 24), (1.3, 1.2, 1.3), (1.0, 0.5, 0.0), np.eye(3))
    M2 = AT_apply(gm2, M, gF, [], "Linear")
    v2 = measure_cc_similarity(F, gF, M2, gm2, 2)
    want2, _, _ = o.cc(og_, F.astype(np.float64), AT_apply(gF, M2, gm2, [], "Linear").astype(np.float64), 2)
    assert v2 == pytest.approx(want2, rel=1e-5)


def test_smooth_field_cli(tmp_path):
    """smooth-field: a variance runs the Gaussian displacement-field
    smoothing (boundary zero), AxBxC the one-level B-spline fit; the files
    equal the ops' outputs on the same field."""
    rng = np.random.default_rng(14)
    g = Geometry((24, 22, 20), (1.0, 1.2, 0.9), (0.0, 0.0, 0.0), np.eye(3))
    fld = rng.standard_normal(g.shape + (3,)).astype(np.float32)
    fin, fout = str(tmp_path / "w.nii.gz"), str(tmp_path / "s.nii.gz")
    aio.write_nifti(fin, fld, g, vector=True)
    fld, g = aio.read_nifti(fin)
    assert cli.main(["smooth-field", "3", fin, fout, "2.0"]) == 0
    got, _ = aio.read_nifti(fout)
    from ants_amd.analysis import smooth_displacement_field
    np.testing.assert_array_equal(got, smooth_displacement_field(fld, g, variance=2.0))
    assert not got[0].any() and not got[:, 0].any() and not got[..., -1, :].any()
    assert 0 < np.abs(got).max() < np.abs(fld).max()
    assert cli.main(["smooth-field", "3", fin, fout, "4x4x3", "1", "3"]) == 0
    got, _ = aio.read_nifti(fout)
    np.testing.assert_array_equal(got, smooth_displacement_field(fld, g, mesh=[4, 4, 3], order=3))
    with pytest.raises(SystemExit):
        cli.main(["smooth-field", "3", fin, fout, "4x4x3", "2"])
