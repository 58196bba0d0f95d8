"""The C-ABI library builds, loads and exports every declared symbol; the
host-side parsers mirror antsRegistration's option grammar.  CPU only: no
compute call is made here."""
import ctypes as C

import numpy as np
import pytest

import oracle as o
from ants_amd import _lib as L
from ants_amd import registration as R


def test_library_exports_every_header_symbol():
    names = L.header_functions()
    assert "ants_syn_register" in names and "ants_op_cc" in names and len(names) > 30
    lib = L.lib()
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_level_geometry_matches_oracle():
    p = L.SynParamsC()
    p.dim = 3
    p.n_levels = 3
    for i, (f, s, it) in enumerate([(4, 2.0, 100), (2, 1.0, 70), (1, 0.0, 20)]):
        p.shrink[i], p.sigma[i], p.iterations[i] = f, s, it
    p.learning_rate, p.update_variance, p.cc_radius, p.convergence_window = 0.1, 3.0, 4, 10
    for size, sp, org in [((192, 224, 192), (1, 1, 1), (0, 0, 0)), ((97, 64, 33), (1.0, 0.8, 2.5), (3, -7, 11))]:
        for lvl in range(3):
            out = L.ImageDesc()
            L.check(L.lib().ants_syn_level_geometry(C.byref(p), C.byref(L.desc(size, sp, org)), lvl, C.byref(out)))
            ref = o.shrink_grid(o.Grid(list(size), list(sp), list(org)), p.shrink[lvl])
            assert [out.size[k] for k in range(3)] == ref.size
            np.testing.assert_allclose([out.spacing[k] for k in range(3)], ref.spacing)
            np.testing.assert_allclose([out.origin[k] for k in range(3)], ref.origin)


def test_invalid_parameters_are_reported():
    p = L.SynParamsC()
    p.dim = 4
    h = C.c_void_p()
    rc = L.lib().ants_syn_create(C.byref(p), C.byref(h))
    assert rc == 1 and b"dim" in L.lib().ants_last_error()
    p.dim, p.n_levels = 3, 0
    assert L.lib().ants_syn_create(C.byref(p), C.byref(h)) == 1
    p.n_levels, p.shrink[0], p.iterations[0], p.cc_radius, p.learning_rate, p.convergence_window = 1, 1, 1, 4, 0.1, 10
    out = L.ImageDesc()
    bad = L.desc((8, 8, 8))
    bad.direction[1] = 0.5
    assert L.lib().ants_syn_level_geometry(C.byref(p), C.byref(bad), 0, C.byref(out)) == 4


def test_cli_grammar():
    st = R.SyNStage.from_args("SyN[0.1,3,0]", "CC[fixed.nii.gz,moving.nii.gz,1,4]", "[100x70x20,1e-6,10]",
                              "4x2x1", "2x1x0vox")
    assert st.transform == R.SyNTransform(0.1, 3.0, 0.0)
    assert st.metric.radius == 4 and st.convergence.iterations == [100, 70, 20]
    assert st.shrink == [4, 2, 1] and st.sigmas == [2.0, 1.0, 0.0] and not st.sigmas_in_mm
    st = R.SyNStage.from_args("SyN[0.25]", "CC[a,b,1]", "40x20x0", "4x2x1", "1mm")
    assert st.transform.update_variance == 3.0 and st.transform.total_variance == 0.0
    assert st.metric.radius == 4 and st.convergence.threshold == 1e-6 and st.convergence.window == 10
    assert st.sigmas == [1.0, 1.0, 1.0] and st.sigmas_in_mm
    with pytest.raises(ValueError):
        R.SyNStage.from_args("SyN[0.1]", "CC[a,b,1,4]", "100x70", "4x2x1", "2x1x0")
    with pytest.raises(ValueError):
        R.SyNTransform.parse("Affine[0.1]")


def _c_layout(tmp_path, exprs):
    """sizeof/offsetof values of the public structs as the C compiler lays
    them out from include/ants_syn.h."""
    import os
    import subprocess
    src = tmp_path / "layout.c"
    body = "\n".join(f'  printf("%zu\\n", (size_t)({e}));' for e in exprs)
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "ants_syn.h"\nint main(void) {\n'
                   + body + "\n  return 0;\n}\n")
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(L.HERE), "include")
    subprocess.run(["gcc", "-std=c99", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    return [int(v) for v in out]


def test_ctypes_structs_match_c_layout(tmp_path):
    from ants_amd import serving as S
    from ants_amd import linear as LN
    from ants_amd import fieldmodels as FM
    structs = {"ants_image_desc": L.ImageDesc, "ants_syn_params": L.SynParamsC, "ants_syn_stats": L.SynStatsC,
               "ants_syn_kernel_stats": L.KernelStatsC, "ants_batch_config": S.BatchConfigC,
               "ants_syn_request": S.RequestC, "ants_request_result": S.ResultC, "ants_transform": L.TransformC,
               "ants_syn_iteration_record": L.IterRecordC, "ants_linear_params": LN.LinearParamsC,
               "ants_linear_record": LN.LinearRecordC, "ants_linear_result": LN.LinearResultC,
               "ants_field_params": FM.FieldParamsC}
    exprs, want = [], []
    for cname, py in structs.items():
        exprs.append(f"sizeof({cname})")
        want.append(C.sizeof(py))
        for fname, _ in py._fields_:
            exprs.append(f"offsetof({cname}, {fname})")
            want.append(getattr(py, fname).offset)
    assert _c_layout(tmp_path, exprs) == want


def test_footprint_closed_form_and_errors():
    from ants_amd import serving as S
    stage = R.SyNStage.from_args("SyN[0.1,3,0]", "CC[f,m,1,4]", "[100x70x20,1e-6,10]", "4x2x1", "2x1x0vox")
    p = S.stage_params(stage, 3)
    d = L.desc((192, 224, 192))
    fp = L.lib().ants_syn_footprint
    fp.restype = C.c_int64
    N = 192 * 224 * 192
    # images (2 copies x 2 images + scratch), warped images, 8 fields (4
    # middle, 4 work), the inversion's candidate masks (2 parities x 2
    # problems x one bit per voxel), request inputs (2 images + labels) and
    # outputs (2 fields + labels)
    want = (2 * N * 4 * 2 + N * 4 + 2 * N * 4 + 8 * N * 3 * 4 + 4 * ((N + 63) // 64) * 8 + 3 * N * 4
            + 2 * N * 3 * 4 + N * 4)
    assert fp(C.byref(p), C.byref(d), C.byref(d)) == want
    pb = S.stage_params(stage, 3, precision="bf16")
    assert fp(C.byref(pb), C.byref(d), C.byref(d)) == want + 2 * N * 2
    # a 3-D update variance below 0.5 adds the two blend fields
    small = R.SyNStage.from_args("SyN[0.1,0.25,0]", "CC[f,m,1,4]", "[100x70x20,1e-6,10]", "4x2x1", "2x1x0vox")
    ps = S.stage_params(small, 3)
    assert fp(C.byref(ps), C.byref(d), C.byref(d)) == want + 2 * N * 3 * 4
    p.dim = 5
    assert fp(C.byref(p), C.byref(d), C.byref(d)) == -1


def test_batch_rejects_bad_slot_counts():
    from ants_amd import serving as S
    S._bind()
    cfg = S.BatchConfigC()
    b = C.c_void_p()
    for slots in (0, 65):
        cfg.slots = slots
        assert L.lib().ants_batch_create(C.byref(cfg), C.byref(b)) == 1
        assert b"slots" in L.lib().ants_last_error()


def test_level_geometry_with_direction_matches_oracle():
    p = L.SynParamsC()
    p.dim, p.n_levels = 3, 2
    p.shrink[0], p.shrink[1], p.iterations[0], p.iterations[1] = 4, 2, 1, 1
    p.learning_rate, p.update_variance, p.cc_radius, p.convergence_window = 0.1, 3.0, 4, 10
    q, _ = np.linalg.qr(np.random.default_rng(1).standard_normal((3, 3)))
    size, sp, org = (97, 64, 33), (1.0, 0.8, 2.5), (3.0, -7.0, 11.0)
    for lvl in range(2):
        out = L.ImageDesc()
        L.check(L.lib().ants_syn_level_geometry(C.byref(p), C.byref(L.desc(size, sp, org, q)), lvl, C.byref(out)))
        ref = o.shrink_grid(o.Grid(list(size), list(sp), list(org), q), p.shrink[lvl])
        assert [out.size[k] for k in range(3)] == ref.size
        np.testing.assert_allclose([out.origin[k] for k in range(3)], ref.origin, atol=1e-12)
        np.testing.assert_allclose(np.array([out.direction[k] for k in range(9)]).reshape(3, 3), q)


def test_observer_argument_converts():
    """Without an observer the wrappers pass a null callback that the
    ctypes signature accepts; with one, a callable that returns 1 to cancel."""
    from ants_amd.registration import SyNRegistration, IterationRecord
    null_cb = SyNRegistration._callback(None)
    L.ITER_CB.from_param(null_cb)
    assert not bool(null_cb)
    kin = (C.c_int * 4)(1, 2, 3, 4)
    got = []
    cb = SyNRegistration._callback(lambda r: got.append(r) or r.iteration == 2)
    L.ITER_CB.from_param(cb)
    assert cb(None, 0, 1, -0.5, 0.1, kin) == 0 and cb(None, 0, 2, -0.6, 0.1, kin) == 1
    assert got[0] == IterationRecord(0, 1, -0.5, 0.1, (1, 2, 3, 4))


def test_field_stage_grammar():
    """-t GaussianDisplacementField / Exponential / BSplineDisplacementField /
    BSplineExponential as antsRegistrationTemplateHeader.h:692-743, 871-935
    reads them: variances, integration steps, meshes (AxBxC or a knot
    spacing with the fixed geometry), spline order positions."""
    from ants_amd import fieldmodels as FM
    st = FM.FieldStage.from_args("GaussianDisplacementField[0.25,3,0.5]", "CC[f,m,1,2]", "[20x10,1e-7,5]", "2x1",
                                 "1x0mm")
    assert (st.model, st.gradient_step, st.update_variance, st.total_variance, st.radius) == (0, 0.25, 3.0, 0.5, 2)
    assert list(st.iterations) == [20, 10] and st.threshold == 1e-7 and st.window == 5 and st.sigma_in_mm
    st = FM.FieldStage.from_args("Exponential[0.5,3,0.5,7]", "CC[f,m]", "[5]", "1", "0")
    assert (st.model, st.integration_steps, st.radius, st.expo) == (1, 7, 4, True)
    st = FM.FieldStage.from_args("Exponential[0.5,3,0.5]", "CC[f,m,1,4]", "[5]", "1", "0")
    assert st.integration_steps == 0
    st = FM.FieldStage.from_args("BSplineDisplacementField[0.3,4x5x6,2x2x2,2]", "CC[f,m,1,4]", "[5]", "1", "0")
    assert (st.model, tuple(st.update_mesh), tuple(st.total_mesh), st.spline_order) == (2, (4, 5, 6), (2, 2, 2), 2)
    st = FM.FieldStage.from_args("BSplineDisplacementField[0.3,4x5x6]", "CC[f,m,1,4]", "[5]", "1", "0")
    assert tuple(st.total_mesh) == (0, 0, 0) and st.spline_order == 3 and not st.expo
    # knot spacing 26 mm on 192x224x192 at 1 mm: ceil(191/26) = 8, ceil(223/26) = 9
    st = FM.FieldStage.from_args("BSplineExponential[0.3,26,0,5,2]", "CC[f,m,1,4]", "[5]", "1", "0",
                                 (192, 224, 192), (1.0, 1.0, 1.0))
    assert (tuple(st.update_mesh), tuple(st.total_mesh), st.integration_steps, st.spline_order) == \
        ((8, 9, 8), (0, 0, 0), 5, 2)
    for bad in ("SyN[0.1,3,0]", "GaussianDisplacementField[0.25,3]", "BSplineExponential[0.3]"):
        with pytest.raises(ValueError):
            FM.FieldStage.from_args(bad, "CC[f,m,1,4]", "[5]", "1", "0")
    with pytest.raises(ValueError):
        FM.FieldStage.from_args("Exponential[0.5,3,0.5]", "MI[f,m,1,32]", "[5]", "1", "0")
    assert FM.FieldStage.is_field("Exponential[1,2,3]") and not FM.FieldStage.is_field("BSplineSyN[0.1,26]")
