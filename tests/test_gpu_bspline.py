"""BSplineSyN on the GPU against the FP64 oracle: the B-spline field fit
(ants_op_bspline_field, the separable fit/evaluate passes the stage runs) and
whole BSplineSyN stages.

Bars: the fit within 2e-5 of the field's scale (float passes against the
oracle's double point-by-point scatter); stages as the SyN stage's parity
bars (energy 1e-4 relative, fields 0.05 voxel)."""
import numpy as np
import pytest

import oracle as o
from ants_amd import _lib as L
from ants_amd import phantoms as ph
from ants_amd.registration import SyNRegistration, SyNStage
from gpu_helpers import bspline_field

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(11)
ENERGY_RTOL = 1e-4
FIELD_ATOL_VOX = 0.05


@pytest.mark.parametrize("size,mesh,q", [((24, 28, 20), (8, 9, 6), 3), ((48, 56, 48), (16, 18, 16), 3),
                                         ((19, 13, 17), (5, 24, 4), 3), ((30, 22, 26), (7, 5, 6), 2),
                                         ((21, 25, 18), (6, 8, 5), 1), ((64, 48), (20, 11), 3),
                                         ((33, 41), (9, 30), 2), ((3000, 6), (40, 6), 3)])
def test_bspline_field_vs_oracle(size, mesh, q):
    """(3000, 6): rows too long for the LDS-staged x fit (the strided pass
    takes over)."""
    D = len(size)
    shape = tuple(reversed(size)) + (D,)
    sp = [1.0, 1.25, 0.8][:D]
    fld = rng.standard_normal(shape).astype(np.float32)
    got, mx = bspline_field(fld, mesh, q, spacing=sp)
    ref = o.bspline_fit_field(o.Grid(list(size), sp), fld.astype(np.float64), mesh, q)
    scale = np.abs(ref).max()
    assert scale > 1e-2
    err = np.abs(got - ref).max() / scale
    print(f"bspline fit {size} mesh {mesh} order {q}: max err {err:.2e} of scale {scale:.3f}")
    assert err < 2e-5
    ref_max = np.sqrt(((ref / np.asarray(sp)) ** 2).sum(-1)).max()
    assert mx == pytest.approx(ref_max, rel=1e-5)


def test_bspline_field_refusals():
    fld = np.zeros((6, 7, 8, 3), np.float32)
    for mesh, order in (((0, 2, 2), 3), ((2, 2, 2), 4), ((2, 2, 2), 0)):
        with pytest.raises(L.AntsError):
            bspline_field(fld, mesh, order)
    with pytest.raises(L.AntsError):          # an axis of one voxel
        bspline_field(np.zeros((1, 7, 8, 3), np.float32), (2, 2, 2))


def run_pair(stage, pair, force=True):
    reg = SyNRegistration(stage, len(pair.size), force_iterations=force)
    res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.close()
    return res


def run_oracle(stage, pair, force=True):
    t = stage.transform
    P = o.SynParams(shrink=stage.shrink, sigma=stage.sigmas, iterations=stage.convergence.iterations,
                    learning_rate=t.gradient_step, radius=stage.metric.radius,
                    conv_threshold=stage.convergence.threshold, conv_window=stage.convergence.window,
                    sigma_in_mm=stage.sigmas_in_mm, force_iterations=force, model=1,
                    update_mesh=t.update_mesh, total_mesh=t.total_mesh, spline_order=t.spline_order)
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    return o.syn_register(P, g, pair.fixed, g, pair.moving)


def compare(res, ref, spacing):
    assert list(np.bincount(res.levels, minlength=3)) == list(np.bincount(ref.level, minlength=3))
    rel = np.abs(res.energy - ref.energy) / np.abs(ref.energy)
    sp = np.asarray(spacing)
    dfwd = (np.abs(res.fwd - ref.fwd) / sp).max()
    dinv = (np.abs(res.inv - ref.inv) / sp).max()
    assert rel.max() < ENERGY_RTOL, f"energy rel err {rel.max():.3e} at record {rel.argmax()}"
    assert dfwd < FIELD_ATOL_VOX and dinv < FIELD_ATOL_VOX, (dfwd, dinv)
    return rel.max(), dfwd, dinv


@pytest.mark.parametrize("total", ["0", "4x5x4"])
def test_bspline_syn_3d_vs_oracle(total):
    """BSplineSyN[0.1, 8x9x8, total, 3] on a 48x56x48 phantom (forced 12x8x4,
    4x2x1 levels: meshes 8x9x8 .. 32x36x32, about 1.5 voxels per span at
    every level), with and without a total-field fit."""
    pair = ph.phantom_pair_3d(3, size=(48, 56, 48), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args(f"BSplineSyN[0.1,8x9x8,{total},3]", "CC[f,m,1,4]", "[12x8x4,1e-6,10]", "4x2x1",
                               "2x1x0vox")
    res = run_pair(stage, pair)
    ref = run_oracle(stage, pair)
    print(f"bsplinesyn 3d total={total} (energy rel, fwd vox, inv vox):", compare(res, ref, pair.spacing))
    assert np.abs(ref.fwd).max() > 0.2          # the fit moved the images


def test_bspline_syn_2d_convergence_vs_oracle():
    """2-D, knot spacing converted with the fixed geometry, quadratic
    splines, convergence monitor on (same per-level iteration counts)."""
    pair = ph.blob_pair_2d(0)
    stage = SyNStage.from_args("BSplineSyN[0.25,16,0,2]", "CC[f,m,1,4]", "[60x40x20,1e-6,10]", "4x2x1", "2x1x0vox",
                               fixed_size=pair.size, fixed_spacing=pair.spacing)
    assert stage.transform.update_mesh == [int(np.ceil((n - 1) * s / 16)) for n, s in zip(pair.size, pair.spacing)]
    res = run_pair(stage, pair, force=False)
    ref = run_oracle(stage, pair, force=False)
    print("bsplinesyn 2d (energy rel, fwd vox, inv vox):", compare(res, ref, pair.spacing))


def test_bspline_syn_params_refused():
    stage = SyNStage.from_args("BSplineSyN[0.1,4x4x4,0,3]", "CC[f,m,1,4]", "[2x2,1e-6,10]", "2x1", "1x0vox")
    for change in ({"spline_order": 4}, {"total_mesh": [4, 0, 4]}, {"update_mesh": [4, 0, 4]}):
        for k, v in change.items():
            setattr(stage.transform, k, v)
        with pytest.raises(L.AntsError):
            SyNRegistration(stage, 3)
        stage = SyNStage.from_args("BSplineSyN[0.1,4x4x4,0,3]", "CC[f,m,1,4]", "[2x2,1e-6,10]", "2x1", "1x0vox")


@pytest.mark.parametrize("mode", ["pipeline", "graphs"])
def test_bspline_syn_same_bits_pipelined_and_replayed(monkeypatch, mode):
    """BSplineSyN with update and total fits (two lattice buffers, one per
    stream) and the observer's full-scale CC: the two-stream pipelined loop
    and replayed iteration graphs give the bits of the default launches."""
    pair = ph.phantom_pair_3d(9, size=(48, 56, 40), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("BSplineSyN[0.15,6x7x5,3x3x3,3]", "CC[f,m,1,2]", "[20x10x4,1e-6,10]", "4x2x1",
                               "2x1x0vox")
    outs = {}
    for on in ("0", "1"):
        monkeypatch.setenv("ANTS_SYN_PIPELINE" if mode == "pipeline" else "ANTS_SYN_GRAPHS", on)
        reg = SyNRegistration(stage, 3)
        reg.set_full_scale_cc_interval(4)
        res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        res2 = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        outs[on] = (res, res2, reg.stats())
        reg.close()
    (a, a2, sa), (b, b2, sb) = outs["0"], outs["1"]
    if mode == "graphs":
        assert sb["graph_launches"] > 0
    for x, y in ((a, b), (a, a2), (b, b2)):
        np.testing.assert_array_equal(x.energy, y.energy)
        np.testing.assert_array_equal(x.fwd, y.fwd)
        np.testing.assert_array_equal(x.inv, y.inv)
        assert [r.full_scale_cc for r in x.trace] == [r.full_scale_cc for r in y.trace]
    assert sum(1 for r in a.trace if r.full_scale_cc != 0.0) >= 6


def test_bspline_syn_split_stage_with_initial_affine():
    """BSplineSyN with a moving initial affine: 4x2 (meshes 4, 8) saving the
    state, then -f 1 with mesh 16 restored from it, equals the 4x2x1 stage
    (meshes 4, 8, 16) bit for bit."""
    from ants_amd.apply_transforms import AffineTransform
    pair = ph.phantom_pair_3d(14, size=(40, 44, 36), max_disp_mm=4.0, svf_sigma_mm=6.0)
    c = np.array(pair.size, float) / 2
    aff = AffineTransform(np.array([[1.02, 0.01, 0.0], [-0.01, 0.99, 0.02], [0.0, 0.01, 1.01]]), [0.4, -0.3, 0.2], c)
    st = lambda mesh, it, f, s_: SyNStage.from_args(f"BSplineSyN[0.2,{mesh},0,3]", "CC[f,m,1,2]",  # noqa: E731
                                                    f"[{it},1e-6,10]", f, s_)

    def run(stage, state=None, grid=None):
        reg = SyNRegistration(stage, 3, force_iterations=True)
        reg.set_initial_transforms(moving=[aff])
        if state is not None:
            reg.set_initial_state(state, *grid)
        res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        lg = reg.level_geometry(pair.size, pair.spacing, pair.origin)
        size = [int(lg.size[k]) for k in range(3)]
        out = reg.state(tuple(reversed(size)) + (3,)), (size, [lg.spacing[k] for k in range(3)],
                                                        [lg.origin[k] for k in range(3)])
        reg.close()
        return res, out

    full, _ = run(st("4x4x4", "5x4x3", "4x2x1", "2x1x0vox"))
    _, (state, grid) = run(st("4x4x4", "5x4", "4x2", "2x1vox"))
    b, _ = run(st("16x16x16", "3", "1", "0vox"), state, grid)
    np.testing.assert_array_equal(b.energy, full.energy[-3:])
    np.testing.assert_array_equal(b.fwd, full.fwd)
    np.testing.assert_array_equal(b.inv, full.inv)
    assert np.abs(full.fwd).max() > 0.1
