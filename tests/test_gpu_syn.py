"""Whole SyN stage on the GPU against the FP64 oracle.

Bars: per-iteration CC energy within
1e-4 relative, forward and inverse fields within 0.05 voxel max-abs, and
nearest-neighbour label warps bit-exact given the same field.  The full
192x224x192 workload is checked through size-independent properties
(inverse consistency, energy descent, determinism)."""
import ctypes as C

import numpy as np
import pytest

import oracle as o
from ants_amd import _lib as L
from ants_amd import phantoms as ph
from ants_amd.registration import SyNRegistration, SyNStage

pytestmark = pytest.mark.gpu

ENERGY_RTOL = 1e-4
FIELD_ATOL_VOX = 0.05


def run_pair(stage, pair, force=False):
    reg = SyNRegistration(stage, len(pair.size), force_iterations=force)
    res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.close()
    return res


def run_oracle(stage, pair, force=False):
    P = o.SynParams(shrink=stage.shrink, sigma=stage.sigmas, iterations=stage.convergence.iterations,
                    learning_rate=stage.transform.gradient_step, update_variance=stage.transform.update_variance,
                    total_variance=stage.transform.total_variance, radius=stage.metric.radius,
                    conv_threshold=stage.convergence.threshold, conv_window=stage.convergence.window,
                    sigma_in_mm=stage.sigmas_in_mm, force_iterations=force)
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    return o.syn_register(P, g, pair.fixed, g, pair.moving)


def compare(res, ref, spacing):
    lv = res.levels
    assert list(np.bincount(lv, minlength=3)) == list(np.bincount(ref.level, minlength=3)), \
        "iteration counts per level differ (convergence decision)"
    e, er = res.energy, ref.energy
    rel = np.abs(e - er) / np.abs(er)
    assert rel.max() < ENERGY_RTOL, f"energy rel err {rel.max():.3e} at record {rel.argmax()}"
    sp = np.asarray(spacing)
    dfwd = np.abs(res.fwd - ref.fwd) / sp
    dinv = np.abs(res.inv - ref.inv) / sp
    assert dfwd.max() < FIELD_ATOL_VOX, f"fwd field deviation {dfwd.max():.4f} voxel"
    assert dinv.max() < FIELD_ATOL_VOX, f"inv field deviation {dinv.max():.4f} voxel"
    return rel.max(), dfwd.max(), dinv.max()


def test_syn_2d_blobs_full_schedule():
    pair = ph.blob_pair_2d(0)
    stage = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,4]", "[100x70x50,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair)
    ref = run_oracle(stage, pair)
    print("2d parity (energy rel, fwd vox, inv vox):", compare(res, ref, pair.spacing))
    # inner inversion step counts: the 0.1 / 0.001 stop thresholds are compared in
    # float on the GPU and in double in the oracle, so a count may differ by one
    kin = np.array([r.inv_iters for r in res.trace])
    diff = np.abs(kin - ref.inv_iters)
    assert diff.max() <= 1 and (diff > 0).mean() < 0.05


def test_syn_3d_phantom_small():
    pair = ph.phantom_pair_3d(3, size=(48, 56, 48), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[20x10x5,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair, force=True)
    ref = run_oracle(stage, pair, force=True)
    print("3d parity (energy rel, fwd vox, inv vox):", compare(res, ref, pair.spacing))


def test_syn_3d_anisotropic_spacing_sigmas_in_mm():
    """Non-power-of-two anisotropic spacing (continuous indices by reciprocal
    spacing, norms in voxel units) and smoothing sigmas given in mm."""
    pair = ph.phantom_pair_3d(3, size=(40, 44, 36), spacing=(1.2, 0.9, 1.5), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[10x6x3,1e-6,10]", "4x2x1", "2x1x0mm")
    res = run_pair(stage, pair, force=True)
    ref = run_oracle(stage, pair, force=True)
    print("3d anisotropic parity:", compare(res, ref, pair.spacing))


def test_syn_2d_direct_cc_total_smoothing_spacing():
    """2-D with CC radius 3 (the untiled CC kernel), a total-field variance
    (the second smoothing and its blend) and anisotropic spacing."""
    pair = ph.blob_pair_2d(4, n=96)
    pair.spacing = (1.3, 0.7)
    stage = SyNStage.from_args("SyN[0.25,3,1]", "CC[f,m,1,3]", "[20x10x5,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair, force=True)
    ref = run_oracle(stage, pair, force=True)
    print("2d r3 total-variance parity:", compare(res, ref, pair.spacing))


def test_syn_3d_radius2_quick_schedule():
    pair = ph.phantom_pair_3d(5, size=(40, 40, 40), max_disp_mm=3.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,2]", "[40x20x0,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair)
    ref = run_oracle(stage, pair)
    print("3d r2 quick parity:", compare(res, ref, pair.spacing))


def test_label_warp_bit_exact_given_same_field():
    pair = ph.phantom_pair_3d(3, size=(48, 56, 48), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[10x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair)
    import gpu_helpers as G
    g = G.desc_of(pair.labels_moving.shape)
    out = G.labels_nn(res.fwd, pair.labels_moving, g, g, pair.labels_moving.shape)
    og = o.Grid(list(pair.size))
    ref = o.warp_labels_nn(og, res.fwd, og, pair.labels_moving)
    np.testing.assert_array_equal(out, ref)
    # the registration brings the moving labels closer to the fixed labels
    assert (out == pair.labels_fixed).mean() > (pair.labels_moving == pair.labels_fixed).mean()


def test_deterministic_rerun():
    pair = ph.blob_pair_2d(2, n=128)
    stage = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,4]", "[20x10x5,1e-6,10]", "4x2x1", "2x1x0vox")
    a = run_pair(stage, pair)
    b = run_pair(stage, pair)
    np.testing.assert_array_equal(a.fwd, b.fwd)
    np.testing.assert_array_equal(a.inv, b.inv)
    np.testing.assert_array_equal(a.energy, b.energy)


def test_full_size_properties():
    """192x224x192 config: energy falls at every level, forward and inverse
    fields are inverse-consistent in the interior, displacement is bounded."""
    pair = ph.phantom_pair_3d(11)
    stage = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[100x70x20,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair)
    lv = res.levels
    for l in range(3):
        e = res.energy[lv == l]
        assert len(e) > 0 and e[-1] < e[0]
    # inverse consistency: fwd(x) + inv(x + fwd(x)) ~ 0 away from the boundary
    og = o.Grid(list(pair.size))
    import gpu_helpers as G
    comp = G.compose(res.inv, res.fwd)
    core = comp[16:-16, 16:-16, 16:-16]
    resid = np.sqrt((core ** 2).sum(-1))
    assert np.percentile(resid, 99) < 0.5 and resid.mean() < 0.1
    assert np.abs(res.fwd).max() < 20.0


def test_inversion_launch_split_is_exact(monkeypatch):
    """Launching only the first K inversion steps per batch with the
    iteration's sequence, and finishing unconverged inversions from the host,
    gives bit-identical results for every K (K=1 finishes every inversion
    from the host, K=21 launches all steps), with and without graphs."""
    pair = ph.phantom_pair_3d(5, size=(64, 72, 56))
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,3]", "[15x10x4,1e-6,10]", "4x2x1", "2x1x0vox")
    outs = {}
    for k, graphs in ((21, "1"), (1, "1"), (2, "1"), (4, "1"), (3, "0")):
        monkeypatch.setenv("ANTS_SYN_INV_LAUNCHES", str(k))
        monkeypatch.setenv("ANTS_SYN_GRAPHS", graphs)
        reg = SyNRegistration(stage, 3)
        res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        st = reg.stats()
        reg.close()
        outs[k] = (res, st)
        print(f"K={k} graphs={graphs}: inversion tails {st['inv_tails']}, iterations {st['iterations']}", flush=True)
    ref, st_ref = outs[21]
    assert st_ref["inv_tails"] == 0
    assert outs[1][1]["inv_tails"] == sum(st_ref["iterations"])
    for k in (1, 2, 4, 3):
        res, st = outs[k]
        np.testing.assert_array_equal(res.energy, ref.energy)
        np.testing.assert_array_equal(res.fwd, ref.fwd)
        np.testing.assert_array_equal(res.inv, ref.inv)
        assert st["inv_steps"] == st_ref["inv_steps"]


def _rot3(a, b, c):
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


def test_syn_3d_oblique_directions_vs_oracle():
    """Fixed and moving with different orthonormal directions (both improper,
    as an LPS flip makes them) and anisotropic spacing."""
    pair = ph.phantom_pair_3d(6, size=(40, 44, 36), spacing=(1.2, 1.0, 1.1), max_disp_mm=4.0, svf_sigma_mm=6.0)
    # an LPS-style flip and an oblique acquisition; the moving image is turned
    # a few degrees further; both centred at the same physical point
    Rf = _rot3(0.3, -0.2, 0.4) @ np.diag([-1.0, -1.0, 1.0])
    Rm = Rf @ _rot3(0.03, -0.04, 0.05)
    half = (np.asarray(pair.size) - 1) / 2.0 * np.asarray(pair.spacing)
    centre = np.array([5.0, -3.0, 2.0])
    of, om = centre - Rf @ half, centre - Rm @ half
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[8x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    reg = SyNRegistration(stage, 3, force_iterations=True)
    res = reg.register(pair.fixed, pair.moving, pair.spacing, of, pair.spacing, om, Rf, Rm)
    reg.close()
    P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=[8, 5, 3], learning_rate=0.2, update_variance=3.0,
                    total_variance=0.0, radius=2, force_iterations=True)
    gf = o.Grid(list(pair.size), list(pair.spacing), list(of), Rf)
    gm = o.Grid(list(pair.size), list(pair.spacing), list(om), Rm)
    ref = o.syn_register(P, gf, pair.fixed, gm, pair.moving)
    assert ref.energy.max() < -0.3 and np.abs(ref.fwd).max() > 0.5, "degenerate case: the images must overlap"
    print("3d oblique parity:", compare(res, ref, pair.spacing))


def test_direction_equivariance_on_gpu():
    """Shared direction R, isotropic spacing: the GPU stage gives the
    identity-direction energies and R times its fields."""
    pair = ph.phantom_pair_3d(8, size=(40, 36, 44), max_disp_mm=4.0, svf_sigma_mm=6.0)
    R = _rot3(-0.35, 0.2, 0.5) @ np.diag([1.0, -1.0, 1.0])
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[8x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    out = []
    for d in (None, R):
        reg = SyNRegistration(stage, 3, force_iterations=True)
        out.append(reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin, d, d))
        reg.close()
    rI, rR = out
    assert rI.energy.max() < -0.3 and np.abs(rI.fwd).max() > 0.5
    np.testing.assert_allclose(rR.energy, rI.energy, rtol=1e-5)
    np.testing.assert_allclose(rR.fwd, rI.fwd @ R.T, atol=2e-3)
    np.testing.assert_allclose(rR.inv, rI.inv @ R.T, atol=2e-3)


def test_trace_and_observer():
    """The engine keeps the iteration records itself: a run without an
    observer returns the same records as a run with a live observer, and an
    observer returning True cancels after that iteration with
    ANTS_ERR_CANCELLED, leaving the records up to it."""
    pair = ph.phantom_pair_3d(2, size=(48, 56, 40))
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[6x4,1e-6,10]", "2x1", "1x0vox")
    reg = SyNRegistration(stage, 3)
    quiet = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    seen = []
    live = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin,
                        observer=lambda r: seen.append(r) and False)
    assert len(quiet.trace) == sum(quiet.stats["iterations"]) > 0
    assert seen == live.trace == quiet.trace
    np.testing.assert_array_equal(live.fwd, quiet.fwd)
    with pytest.raises(L.AntsError) as exc:
        reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin,
                     observer=lambda r: r.iteration == 3)
    assert exc.value.code == L.ANTS_ERR_CANCELLED
    cut = reg.trace()
    assert cut == quiet.trace[:3]
    reg.close()


def test_profiling_stride_and_launch_modes(monkeypatch):
    """Timing spans on every iteration or every 4th, and graph replay or
    kernel-by-kernel launches, leave the results bit for bit; the stride
    times the first iteration of each level and every 4th after it."""
    pair = ph.phantom_pair_3d(3, size=(48, 56, 40))
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[9x6,1e-6,10]", "2x1", "1x0vox")
    inv = 1 << L.KC_NAMES.index("inv")
    out = {}
    for name, mask, every, graphs in (("all", inv, 1, "0"), ("stride", inv, 4, "0"), ("none", 0, 1, "0"),
                                      ("graphs", 0, 1, "1")):
        monkeypatch.setenv("ANTS_SYN_GRAPHS", graphs)
        reg = SyNRegistration(stage, 3, force_iterations=True)
        reg.set_profiling(mask, every=every)
        res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        out[name] = (res, reg.kernel_stats()["inv"], reg.stats())
        reg.close()
    ref = out["none"][0]
    for name in ("all", "stride", "graphs"):
        np.testing.assert_array_equal(out[name][0].fwd, ref.fwd)
        np.testing.assert_array_equal(out[name][0].energy, ref.energy)
    assert out["graphs"][2]["graph_launches"] == 15 and out["none"][2]["eager_iterations"] == 15
    # spans: one per timed iteration (plus host tails); iterations 1, 5, 9 of 9 and 1, 5 of 6 with stride 4
    n_all, n_stride = out["all"][1]["launches"], out["stride"][1]["launches"]
    assert 0 < n_stride < n_all
    assert out["none"][1]["launches"] == 0


def test_speculative_inversion_is_exact(monkeypatch):
    """The inversion's speculative updates (large grids) and the plain
    recompute-in-the-next-launch scheme give the same bits: forced on and
    off for the same small registration, with graphs, and with the launch
    split that finishes every inversion from the host."""
    pair = ph.phantom_pair_3d(6, size=(72, 80, 64))
    stage = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,3]", "[12x8,1e-6,10]", "2x1", "1x0vox")
    outs = {}
    for name, thresh, k in (("plain", "1000000000", None), ("spec", "0", None), ("spec_k1", "0", "1")):
        monkeypatch.setenv("ANTS_SYN_INV_SPEC_MIN_VOXELS", thresh)
        if k is None:
            monkeypatch.delenv("ANTS_SYN_INV_LAUNCHES", raising=False)
        else:
            monkeypatch.setenv("ANTS_SYN_INV_LAUNCHES", k)
        reg = SyNRegistration(stage, 3)
        res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        outs[name] = (res, reg.stats())
        reg.close()
        print(f"{name}: inversion steps {outs[name][1]['inv_steps']}, tails {outs[name][1]['inv_tails']}", flush=True)
    ref, st_ref = outs["plain"]
    assert sum(st_ref["inv_steps"]) > 0
    for name in ("spec", "spec_k1"):
        res, st = outs[name]
        np.testing.assert_array_equal(res.energy, ref.energy)
        np.testing.assert_array_equal(res.fwd, ref.fwd)
        np.testing.assert_array_equal(res.inv, ref.inv)
        assert st["inv_steps"] == st_ref["inv_steps"]


# ------------------------------------------------- initial transforms of the stage
def _affine_case(D, seed):
    """A general affine (rotation, anisotropic scale, shear) about a centre."""
    rng = np.random.default_rng(seed)
    R = _rot3(*rng.uniform(-0.15, 0.15, 3)) if D == 3 else _rot3(0, 0, rng.uniform(-0.15, 0.15))[:2, :2]
    Sh = np.eye(D)
    Sh[0, 1] = 0.03
    A = R @ np.diag(rng.uniform(0.96, 1.04, D)) @ Sh
    return A, rng.uniform(-1.5, 1.5, D), rng.uniform(10, 20, D)


def _run_with_initial(stage, pair, moving=(), fixed=(), precision="fp32", mgrid=None):
    from ants_amd.apply_transforms import AffineTransform
    D = len(pair.size)
    reg = SyNRegistration(stage, D, precision=precision, force_iterations=True)
    reg.set_initial_transforms([AffineTransform(*t) for t in moving], [AffineTransform(*t) for t in fixed])
    sp, og, dr = (pair.spacing, pair.origin, None) if mgrid is None else mgrid
    res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, sp, og, None, dr)
    reg.close()
    return res


def _oracle_with_initial(stage, pair, moving=(), fixed=()):
    P = o.SynParams(shrink=stage.shrink, sigma=stage.sigmas, iterations=stage.convergence.iterations,
                    learning_rate=stage.transform.gradient_step, update_variance=stage.transform.update_variance,
                    total_variance=stage.transform.total_variance, radius=stage.metric.radius,
                    conv_threshold=stage.convergence.threshold, conv_window=stage.convergence.window,
                    sigma_in_mm=stage.sigmas_in_mm, force_iterations=True)
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    return o.syn_register(P, g, pair.fixed, g, pair.moving, fixed_xforms=[("affine",) + tuple(t) for t in fixed],
                          moving_xforms=[("affine",) + tuple(t) for t in moving])


def test_initial_transforms_3d_vs_oracle():
    """Moving initial transform = a composite of two affines (the -r
    transform and a preceding affine stage), fixed initial = a small rigid
    one (-q), against the oracle."""
    pair = ph.phantom_pair_3d(9, size=(40, 44, 36), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[8x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    mv = [_affine_case(3, 1), _affine_case(3, 2)]
    R, t, c = _affine_case(3, 3)
    fx = [(_rot3(0.02, -0.03, 0.01), t * 0.3, c)]
    res = _run_with_initial(stage, pair, mv, fx)
    ref = _oracle_with_initial(stage, pair, mv, fx)
    plain = _oracle_with_initial(stage, pair)
    assert np.abs(plain.energy - ref.energy).max() > 1e-3, "the transforms must change the problem"
    print("3d initial-transform parity:", compare(res, ref, pair.spacing))


def test_initial_transform_2d_vs_oracle():
    pair = ph.blob_pair_2d(1)
    stage = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,4]", "[20x10x5,1e-6,10]", "4x2x1", "2x1x0vox")
    mv = [_affine_case(2, 4)]
    res = _run_with_initial(stage, pair, mv)
    ref = _oracle_with_initial(stage, pair, mv)
    print("2d initial-transform parity:", compare(res, ref, pair.spacing))


def test_initial_rigid_transform_equals_header_change_on_gpu():
    """A rigid moving initial transform equals moving the moving header by
    its inverse; the GPU's two coordinate paths (folded affine map, oblique
    grid) agree."""
    pair = ph.phantom_pair_3d(10, size=(40, 36, 44), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[8x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    R = _rot3(0.05, -0.08, 0.06)
    t, c = np.array([0.7, -0.4, 0.9]), np.array([20.0, 18.0, 22.0])
    res = _run_with_initial(stage, pair, [(R, t, c)])
    org = R.T @ (np.asarray(pair.origin) - t - c) + c
    hdr = _run_with_initial(stage, pair, mgrid=(pair.spacing, org, R.T))
    np.testing.assert_allclose(res.energy, hdr.energy, rtol=1e-5)
    np.testing.assert_allclose(res.fwd, hdr.fwd, atol=5e-3)
    np.testing.assert_allclose(res.inv, hdr.inv, atol=5e-3)


def test_initial_transform_bf16_and_clearing():
    pair = ph.phantom_pair_3d(11, size=(32, 36, 28), max_disp_mm=3.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[6x4x2,1e-6,10]", "4x2x1", "2x1x0vox")
    mv = [_affine_case(3, 7)]
    ref = _oracle_with_initial(stage, pair, mv)
    b = _run_with_initial(stage, pair, mv, precision="bf16")
    np.testing.assert_allclose(b.energy, ref.energy, rtol=5e-3)
    # transforms stay on the handle until cleared; a cleared handle runs the plain stage
    from ants_amd.apply_transforms import AffineTransform
    reg = SyNRegistration(stage, 3, force_iterations=True)
    reg.set_initial_transforms([AffineTransform(*mv[0])])
    r1 = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    r2 = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.set_initial_transforms()
    r3 = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.close()
    plain = run_pair(stage, pair, force=True)
    np.testing.assert_array_equal(r1.fwd, r2.fwd)
    np.testing.assert_array_equal(r3.fwd, plain.fwd)
    assert np.abs(r1.fwd - r3.fwd).max() > 1e-3


def test_initial_transform_errors():
    from ants_amd.apply_transforms import AffineTransform, DisplacementFieldTransform, Geometry
    stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[2x1,1e-6,10]", "2x1", "1x0vox")
    reg = SyNRegistration(stage, 3)
    geo = Geometry(size=(4, 4, 4), spacing=(1, 1, 1), origin=(0, 0, 0), direction=np.eye(3))
    fld = DisplacementFieldTransform(np.zeros((4, 4, 4, 3), np.float32), geo)
    # a list with a displacement field keeps one slot for the SyN field: at most 7
    with pytest.raises(L.AntsError) as e:
        reg.set_initial_transforms([fld] + [AffineTransform(np.eye(3), [0, 0, 0])] * 7)
    assert e.value.code == 4
    bad = DisplacementFieldTransform(np.zeros((4, 4, 3), np.float32), Geometry(size=(4, 4), spacing=(1, 1)))
    with pytest.raises((L.AntsError, ValueError)):
        reg.set_initial_transforms([bad])
    with pytest.raises(ValueError):
        reg.set_initial_transforms([AffineTransform(np.eye(3), [0, 0, 0])] * 9)
    reg.close()


def _smooth_field(shape, D, amp, seed):
    from scipy import ndimage
    rng = np.random.default_rng(seed)
    f = np.stack([ndimage.gaussian_filter(rng.standard_normal(shape), 3.0) for _ in range(D)], -1)
    return np.ascontiguousarray(amp * f / np.abs(f).max(), np.float32)


@pytest.mark.parametrize("D", [3, 2])
def test_displacement_field_initial_transforms_vs_oracle(D):
    """Initial transforms holding displacement fields (an earlier deformable
    stage, or a restored composite): moving = [affine, field on its own
    coarser oblique-free grid], fixed = [field], against the oracle applying
    the same lists point by point."""
    from ants_amd.apply_transforms import AffineTransform, DisplacementFieldTransform, Geometry
    if D == 3:
        pair = ph.phantom_pair_3d(13, size=(40, 44, 36), max_disp_mm=3.0, svf_sigma_mm=6.0)
        stage = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[8x5x3,1e-6,10]", "4x2x1", "2x1x0vox")
    else:
        pair = ph.blob_pair_2d(5, n=96)
        stage = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,4]", "[10x6x3,1e-6,10]", "4x2x1", "2x1x0vox")
    A, t, c = _affine_case(D, 8)
    fsize = tuple(int(s) // 2 + 4 for s in pair.size)
    fsp, forg = [2.0] * D, [-3.0] * D
    fm = _smooth_field(tuple(reversed(fsize)), D, 2.0, 1)
    ff = _smooth_field(tuple(reversed(fsize)), D, 1.0, 2)
    geo = Geometry(size=fsize, spacing=fsp, origin=forg)
    reg = SyNRegistration(stage, D, force_iterations=True)
    reg.set_initial_transforms([AffineTransform(A, t, c), DisplacementFieldTransform(fm, geo)],
                               [DisplacementFieldTransform(ff, geo)])
    res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.close()
    P = o.SynParams(shrink=stage.shrink, sigma=stage.sigmas, iterations=stage.convergence.iterations,
                    learning_rate=stage.transform.gradient_step, radius=stage.metric.radius, force_iterations=True)
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    og = o.og_geom(fsize, fsp, forg)
    ref = o.syn_register(P, g, pair.fixed, g, pair.moving, fixed_xforms=[("field", og, ff)],
                         moving_xforms=[("affine", A, t, c), ("field", og, fm)])
    plain = o.syn_register(P, g, pair.fixed, g, pair.moving)
    assert np.abs(plain.energy - ref.energy).max() > 1e-3, "the transforms must change the problem"
    print(f"{D}-d displacement-field initial transforms parity:", compare(res, ref, pair.spacing))


# ------------------------------------------------------------ restored state
def _level_grid(reg, pair):
    g = reg.level_geometry(pair.size, pair.spacing, pair.origin)
    size = tuple(int(g.size[k]) for k in range(3))
    return size, tuple(g.spacing[k] for k in range(3)), tuple(g.origin[k] for k in range(3))


def test_restored_state_split_equals_single_stage():
    """4x2 saving the state, then 1 restored from it, equals the 4x2x1 stage
    bit for bit on the GPU; the restored part matches the oracle fed the
    same state."""
    pair = ph.phantom_pair_3d(12, size=(40, 44, 36), spacing=(1.1, 1.0, 1.2), max_disp_mm=4.0, svf_sigma_mm=6.0)
    full_st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[6x4x3,1e-6,10]", "4x2x1", "2x1x0vox")
    a_st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[6x4,1e-6,10]", "4x2", "2x1vox")
    b_st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[3,1e-6,10]", "1", "0vox")
    full = run_pair(full_st, pair, force=True)
    reg = SyNRegistration(a_st, 3, force_iterations=True)
    reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    size, sp, og = _level_grid(reg, pair)
    state = reg.state(tuple(reversed(size)) + (3,))
    reg.close()
    reg = SyNRegistration(b_st, 3, force_iterations=True)
    reg.set_initial_state(state, size, sp, og)
    b = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.close()
    np.testing.assert_array_equal(b.energy, full.energy[-3:])
    np.testing.assert_array_equal(b.fwd, full.fwd)
    np.testing.assert_array_equal(b.inv, full.inv)
    # the oracle restored from the same state
    P = o.SynParams(shrink=[1], sigma=[0], iterations=[3], learning_rate=0.2, update_variance=3.0,
                    total_variance=0.0, radius=4, force_iterations=True)
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    ref = o.syn_register(P, g, pair.fixed, g, pair.moving, restore=(o.Grid(list(size), list(sp), list(og)),
                                                                     np.stack(state)))
    print("restored-state parity:", compare(b, ref, pair.spacing))


def test_restored_state_errors_and_clearing():
    pair = ph.phantom_pair_3d(13, size=(32, 36, 28), max_disp_mm=3.0, svf_sigma_mm=6.0)
    st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[4x2,1e-6,10]", "2x1", "1x0vox")
    reg = SyNRegistration(st, 3, force_iterations=True)
    plain = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    shape = tuple(reversed(pair.size)) + (3,)
    fields = [L.DeviceArray.from_host(np.zeros(shape, np.float32)) for _ in range(4)]
    g = L.desc(pair.size, pair.spacing, pair.origin)
    with pytest.raises(L.AntsError):
        L.check(L.lib().ants_syn_set_initial_state_device(reg.handle, C.byref(g), fields[0].ptr, None,
                                                          fields[2].ptr, fields[3].ptr))
    # a zero state is the cold start; clearing returns to it as well
    reg.set_initial_state([f.to_host() for f in fields], pair.size, pair.spacing, pair.origin)
    z = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.set_initial_state(None)
    c = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
    reg.close()
    np.testing.assert_array_equal(z.fwd, plain.fwd)
    np.testing.assert_array_equal(c.fwd, plain.fwd)


# ------------------------------------------------------------ edge cases
def _edge_case(case):
    import dataclasses
    if case == "ragged3d":
        # sizes that fill no tile evenly; the coarsest grid (9x7x5) is
        # smaller than the CC window and the smoothing kernels
        pair = ph.phantom_pair_3d(14, size=(37, 29, 23), spacing=(1.3, 1.0, 1.1), max_disp_mm=3.0,
                                  svf_sigma_mm=6.0)
        st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[6x4x2,1e-6,10]", "4x2x1", "2x1x0vox")
        return pair, st, True
    if case == "zero_iteration_level":
        pair = ph.phantom_pair_3d(15, size=(40, 36, 32), max_disp_mm=3.0, svf_sigma_mm=6.0)
        st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,2]", "[6x0x3,1e-6,10]", "4x2x1", "2x1x0vox")
        return pair, st, True
    if case == "constant_moving":
        # every window is degenerate: CC 1, zero derivatives, zero fields,
        # and the energy window converges after 10 iterations per level
        pair = ph.phantom_pair_3d(16, size=(32, 28, 24), max_disp_mm=3.0, svf_sigma_mm=6.0)
        pair = dataclasses.replace(pair, moving=np.full_like(pair.moving, 0.3))
        st = SyNStage.from_args("SyN[0.2,3,0]", "CC[f,m,1,4]", "[20x20x20,1e-6,10]", "4x2x1", "2x1x0vox")
        return pair, st, False
    if case == "small2d":
        pair = ph.blob_pair_2d(2, n=40, n_blobs=6, max_disp=2.0)
        st = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,4]", "[10x6x4,1e-6,10]", "4x2x1", "2x1x0vox")
        return pair, st, True
    raise ValueError(case)


@pytest.mark.parametrize("case", ["ragged3d", "zero_iteration_level", "constant_moving", "small2d"])
def test_syn_edge_cases_vs_oracle(case):
    pair, stage, force = _edge_case(case)
    res = run_pair(stage, pair, force=force)
    ref = run_oracle(stage, pair, force=force)
    print(f"{case} parity:", compare(res, ref, pair.spacing))
    if case == "constant_moving":
        assert np.all(res.energy == -1.0) and not res.fwd.any() and not res.inv.any()
        assert list(np.bincount(res.levels, minlength=3)) == [10, 10, 10]
    if case == "zero_iteration_level":
        assert 1 not in set(res.levels.tolist())


def test_strict_tie_rule_against_oracle(monkeypatch):
    """ITK's inversion max test as written (max > 0.1, ANTS_SYN_INV_STRICT_TIE=1)
    against the default guarded rule, on the GPU and in the oracle, on the
    SyN[0.1,...] case whose first inversions start on the 0.1 edge.  With the
    strict rule the tie is decided by rounding (float on the GPU, double in
    the oracle), so GPU and oracle may take different paths: reported, not
    asserted.  With the guard both take the same path (asserted)."""
    pair = ph.phantom_pair_3d(7, size=(48, 56, 48), max_disp_mm=4.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[100x70x20,1e-6,10]", "4x2x1", "2x1x0vox")
    g = o.Grid(list(pair.size))
    guarded = run_pair(stage, pair)
    ref_guarded = run_oracle(stage, pair)
    print("guarded GPU vs guarded oracle:", compare(guarded, ref_guarded, pair.spacing), flush=True)
    monkeypatch.setenv("ANTS_SYN_INV_STRICT_TIE", "1")
    strict = run_pair(stage, pair)
    P = o.SynParams(shrink=[4, 2, 1], sigma=[2, 1, 0], iterations=[100, 70, 20], learning_rate=0.1, radius=4,
                    inv_tie_rule=1)
    ref_strict = o.syn_register(P, g, pair.fixed, g, pair.moving)

    def diff(a_fwd, a_inv, b_fwd, b_inv):
        return max(np.abs(a_fwd - b_fwd).max(), np.abs(a_inv - b_inv).max())

    print(f"strict GPU: iterations {np.bincount(strict.levels).tolist()}, first inner steps "
          f"{[list(r.inv_iters) for r in strict.trace[:1]]}; vs guarded GPU {diff(strict.fwd, strict.inv, guarded.fwd, guarded.inv):.2e} mm",
          flush=True)
    print(f"strict oracle: iterations {np.bincount(ref_strict.level).tolist()}, first inner steps "
          f"{ref_strict.inv_iters[:1].tolist()}; vs guarded oracle "
          f"{diff(ref_strict.fwd, ref_strict.inv, ref_guarded.fwd, ref_guarded.inv):.2e} mm; vs strict GPU "
          f"{diff(ref_strict.fwd, ref_strict.inv, strict.fwd, strict.inv):.2e} mm", flush=True)


@pytest.mark.parametrize("case", ["3d_converging", "3d_host_tails", "2d_total_variance", "3d_bf16"])
def test_pipelined_iterations_are_exact(monkeypatch, case):
    """The next iteration's front end (warp, CC, smoothing) on a second
    stream beside this iteration's inversions (ANTS_SYN_PIPELINE=1) gives the same bits as the
    strictly sequential launches: with the convergence monitor stopping
    levels early (a front end enqueued past the last iteration is dropped),
    with host tails on every iteration (one inversion launch per batch, so
    the front end is run again after the first batch's tail), in 2-D with a
    total-field variance (the separable smoothing's buffer swaps), and in
    the BF16 form."""
    if case == "2d_total_variance":
        pair = ph.blob_pair_2d(6, n=96)
        stage = SyNStage.from_args("SyN[0.25,3,1]", "CC[f,m,1,4]", "[30x20x10,1e-6,10]", "4x2x1", "2x1x0vox")
    else:
        pair = ph.phantom_pair_3d(8, size=(56, 64, 48), max_disp_mm=4.0, svf_sigma_mm=6.0)
        stage = SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[60x30x10,1e-6,10]", "4x2x1", "2x1x0vox")
    if case == "3d_host_tails":
        monkeypatch.setenv("ANTS_SYN_INV_LAUNCHES", "1")
    prec = "bf16" if case == "3d_bf16" else "fp32"
    outs = {}
    for pipe in ("0", "1"):
        monkeypatch.setenv("ANTS_SYN_PIPELINE", pipe)
        reg = SyNRegistration(stage, len(pair.size), precision=prec)
        res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin)
        outs[pipe] = (res, reg.stats())
        reg.close()
    (a, sa), (b, sb) = outs["0"], outs["1"]
    print(f"{case}: iterations {sa['iterations']}, host tails {sb['inv_tails']}", flush=True)
    np.testing.assert_array_equal(a.energy, b.energy)
    np.testing.assert_array_equal(a.fwd, b.fwd)
    np.testing.assert_array_equal(a.inv, b.inv)
    assert sa["iterations"] == sb["iterations"] and sa["inv_steps"] == sb["inv_steps"]
    if case == "3d_host_tails":
        assert sb["inv_tails"] == sum(sb["iterations"])
    if case == "3d_converging":
        assert sa["iterations"][0] < 60  # the monitor stopped a level early


def test_syn_3d_small_variances_vs_oracle():
    """3-D stage with update and total variances below 0.5 (the smoothing
    blends the smoothed field with its input; off the fused 13-tap path)."""
    pair = ph.phantom_pair_3d(12, size=(40, 44, 36), max_disp_mm=3.0, svf_sigma_mm=6.0)
    stage = SyNStage.from_args("SyN[0.2,0.3,0.4]", "CC[f,m,1,2]", "[10x6x3,1e-6,10]", "4x2x1", "2x1x0vox")
    res = run_pair(stage, pair, force=True)
    ref = run_oracle(stage, pair, force=True)
    print("3d small-variance parity:", compare(res, ref, pair.spacing))


@pytest.mark.parametrize("shrink,iters", [("1", "7"), ("2", "7")])
def test_full_scale_cc_matches_oracle(shrink, iters):
    """--print-similarity-measure-interval 3: computed at iterations 1, 3, 6
    and the last (7), zero elsewhere; the last value equals the oracle's CC
    r=4 between the ORIGINAL images resampled onto the fixed grid through the
    run's final inverse middle fields (on the half-resolution level grid when
    the schedule is -f 2: interpolated, zero outside the field's buffer)."""
    pair = ph.phantom_pair_3d(5, size=(40, 48, 36), max_disp_mm=3.0, svf_sigma_mm=5.0)
    stage = SyNStage.from_args("SyN[0.25,3,0]", "CC[f,m,1,2]", f"[{iters},1e-6,10]", shrink, "0vox")
    reg = SyNRegistration(stage, 3, force_iterations=True)
    reg.set_full_scale_cc_interval(3)
    seen = []
    res = reg.register(pair.fixed, pair.moving, pair.spacing, pair.origin, pair.spacing, pair.origin,
                       observer=lambda r: seen.append(r.full_scale_cc) and False)
    lg = reg.level_geometry(pair.size, pair.spacing, pair.origin)
    lshape = tuple(int(lg.size[k]) for k in reversed(range(3)))
    f2m_f, f2m_i, m2m_f, m2m_i = reg.state(lshape + (3,))
    reg.close()
    fs = [r.full_scale_cc for r in res.trace]
    assert fs == seen
    assert [i + 1 for i, v in enumerate(fs) if v != 0.0] == [1, 3, 6, 7]
    assert any("e-" in ln and ln.count(",") == 4 for ln in res.diagnostic_lines())
    gref = o.og_geom(pair.size, pair.spacing, pair.origin)
    gfld = o.og_geom([int(lg.size[k]) for k in range(3)], [lg.spacing[k] for k in range(3)],
                     [lg.origin[k] for k in range(3)])
    Fw = o.apply_transforms(gref, gref, pair.fixed, [("field", gfld, f2m_i)])
    Mw = o.apply_transforms(gref, gref, pair.moving, [("field", gfld, m2m_i)])
    g = o.Grid(list(pair.size), list(pair.spacing), list(pair.origin))
    ref, _, _ = o.cc(g, Fw.astype(np.float64), Mw.astype(np.float64), 4)
    print(f"full-scale CC -f {shrink}: gpu {fs[-1]:.9f} oracle {ref:.9f}")
    assert fs[-1] == pytest.approx(ref, rel=1e-5)
