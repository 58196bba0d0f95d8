"""Host-side helpers (CPU only)."""
import numpy as np
import pytest

from ants_amd.analysis import jacobian_determinant


def test_jacobian_of_affine_field():
    # u = A x: det(I + A) everywhere
    A = np.array([[0.1, 0.02, 0.0], [-0.03, 0.05, 0.01], [0.0, 0.02, -0.04]])
    z, y, x = np.meshgrid(np.arange(6.0), np.arange(7.0), np.arange(8.0), indexing="ij")
    pts = np.stack([x, y, z], -1)
    u = pts @ A.T
    J = jacobian_determinant(u, (1.0, 1.0, 1.0))
    np.testing.assert_allclose(J, np.linalg.det(np.eye(3) + A), rtol=1e-12)


def _counter_csv(path, counter, rows):
    cols = ["Correlation_Id", "Dispatch_Id", "Agent_Id", "Grid_Size", "Kernel_Name", "Counter_Name", "Counter_Value"]
    with open(path, "w") as fh:
        fh.write(",".join(f'"{c}"' for c in cols) + "\n")
        for did, name, grid, val in rows:
            fh.write(f'{did},{did},1,{grid},"{name}","{counter}",{val}\n')


def test_counter_traffic_scaling(tmp_path):
    """Two counter passes combine into bytes per dispatch scaled by the copy
    kernel: a copy of B bytes reading 2x its counter gives scale 1/2 in each
    pass, so the kernel's KiB readings are halved."""
    from ants_amd.profiling import traffic
    B = 64 * 2**20
    fetch, write = tmp_path / "f", tmp_path / "w"
    fetch.mkdir()
    write.mkdir()
    copy_kib = B / 1024
    _counter_csv(fetch / "run_counter_collection.csv", "FETCH_SIZE",
                 [(1, "k_copy_calibration(float4 const*, float4*, long)", 1000, 2 * copy_kib),
                  (2, "void (anonymous namespace)::k_inv_step<3>(GridDev)", 512, 400.0),
                  (3, "void (anonymous namespace)::k_inv_step<3>(GridDev)", 512, 200.0),
                  (4, "k_cc3d<4, 64, 4>", 64, 9999.0)])
    _counter_csv(write / "run_counter_collection.csv", "WRITE_SIZE",
                 [(1, "k_copy_calibration(float4 const*, float4*, long)", 1000, copy_kib),
                  (2, "void (anonymous namespace)::k_inv_step<3>(GridDev)", 512, 100.0)])
    out = dict(ln.split(": ", 1) for ln in traffic(str(fetch), str(write), "k_inv_step", B).splitlines()
               if ln.startswith(("fetch_", "write_", "traffic_", "kernel")))
    assert out["kernel"] == "k_inv_step"
    assert float(out["fetch_bytes_per_launch"]) == 0.5 * 300.0 * 1024
    assert float(out["write_bytes_per_launch"]) == 100.0 * 1024
    assert float(out["traffic_bytes_per_launch"]) == 0.5 * 300.0 * 1024 + 100.0 * 1024


def test_compose_affines_applies_the_first_transform_first():
    from ants_amd.apply_transforms import AffineTransform, compose_affines
    rng = np.random.default_rng(3)
    ts = [AffineTransform(rng.normal(size=(3, 3)) + 2 * np.eye(3), rng.normal(size=3), rng.normal(size=3))
          for _ in range(3)]
    ts[1].invert = True

    def apply(t, p):
        A, tr, c = t.parameters()
        return A @ (p - c) + tr + c

    comp = compose_affines([ts[0], None, ts[1], ts[2]])
    for p in rng.normal(size=(6, 3)):
        np.testing.assert_allclose(apply(comp, p), apply(ts[2], apply(ts[1], apply(ts[0], p))), atol=1e-9)
    with pytest.raises(ValueError):
        compose_affines([None])


def test_tuning_table_set_get_load(tmp_path):
    """The launch-shape tuning table: set / get per knob and log2 bucket,
    a file in the tuner's format, and refusal of bad entries."""
    import ctypes as C
    from ants_amd import _lib as L
    lib = L.lib()
    saved = {}
    v = C.c_int()
    for k in range(5):
        for b in range(40):
            L.check(lib.ants_tuning_get(k, b, C.byref(v)))
            saved[(k, b)] = v.value
    try:
        L.check(lib.ants_tuning_clear())
        L.check(lib.ants_tuning_set(0, 20, 512))
        L.check(lib.ants_tuning_get(0, 20, C.byref(v)))
        assert v.value == 512
        L.check(lib.ants_tuning_get(0, 19, C.byref(v)))
        assert v.value == 0
        assert lib.ants_tuning_set(3, 20, 3) == 1      # tiles are 1 or 2
        assert lib.ants_tuning_set(7, 20, 1) == 1      # no such knob
        assert lib.ants_tuning_set(1, 40, 8) == 1      # no such bucket
        f = tmp_path / "tuning.txt"
        f.write_text("# knob bucket value\ncc_zchunk 17 12\nsmooth_tile 20 2  # narrow tiles\n\n")
        L.check(lib.ants_tuning_load(str(f).encode()))
        L.check(lib.ants_tuning_get(1, 17, C.byref(v)))
        assert v.value == 12
        L.check(lib.ants_tuning_get(4, 20, C.byref(v)))
        assert v.value == 2
        bad = tmp_path / "bad.txt"
        bad.write_text("cc_zchunk 17\n")
        assert lib.ants_tuning_load(str(bad).encode()) == 1
    finally:
        L.check(lib.ants_tuning_clear())
        for (k, b), val in saved.items():
            if val:
                L.check(lib.ants_tuning_set(k, b, val))


def test_parse_interpolation_grammar():
    """-n names (make_interpolator_snip.tmpl): case-insensitive, BSpline's
    order 0..5 defaulting to 3, GenericLabel with Linear only."""
    from ants_amd.apply_transforms import parse_interpolation
    assert parse_interpolation("Linear") == 0
    assert parse_interpolation("nearestNeighbor") == 1
    assert parse_interpolation("GenericLabel[Linear]") == 2
    assert parse_interpolation("BSpline") == 16 + 3
    assert [parse_interpolation(f"BSpline[{q}]") for q in range(6)] == [16 + q for q in range(6)]
    assert parse_interpolation("LanczosWindowedSinc") == 34
    assert [parse_interpolation(w + "WindowedSinc") for w in ("Cosine", "hamming", "Lanczos", "BLACKMAN", "Welch")] \
        == [32, 33, 34, 35, 36]
    for bad in ("BSpline[6]", "BSpline[-1]", "BSpline[2.5]", "GenericLabel[NearestNeighbor]", "Gaussian", "Sinc[3]",
                "MultiLabel"):
        with pytest.raises(ValueError):
            parse_interpolation(bad)


def test_apply_output_forms():
    """antsApplyTransforms -o grammar (Examples/antsApplyTransforms.cxx:643-770)."""
    from ants_amd.cli import parse_apply_output
    assert parse_apply_output("out.nii.gz") == ("image", "out.nii.gz", False)
    assert parse_apply_output("[out.nii.gz,0]") == ("image", "out.nii.gz", False)
    assert parse_apply_output("[warp.nii.gz,1]") == ("field", "warp.nii.gz", False)
    assert parse_apply_output("DisplacementField[warp.nii.gz]") == ("field", "warp.nii.gz", False)
    assert parse_apply_output("Linear[a.mat]") == ("linear", "a.mat", False)
    assert parse_apply_output("Linear[a.mat,1]") == ("linear", "a.mat", True)
    assert parse_apply_output("CompositeTransform[c.h5]") == ("composite", "c.h5", False)
    with pytest.raises(ValueError):
        parse_apply_output("Gaussian[x.nii.gz]")


def test_apply_collapse_linear(tmp_path):
    """-o Linear[file(,1)] writes the -t affines collapsed into one (and its
    inverse): the written transform maps points as the list does; a
    displacement field in the list is refused."""
    from ants_amd import io as aio
    from ants_amd.apply_transforms import AffineTransform, DisplacementFieldTransform, Geometry, collapse_linear
    from ants_amd.cli import cmd_apply
    rng = np.random.default_rng(3)
    ts = [AffineTransform(np.eye(3) + 0.1 * rng.standard_normal((3, 3)), rng.standard_normal(3), rng.standard_normal(3))
          for _ in range(2)]
    paths = []
    for k, t in enumerate(ts):
        paths.append(str(tmp_path / f"t{k}.mat"))
        aio.write_transform(paths[-1], t)
    pts = rng.standard_normal((20, 3)) * 10

    def apply(t, p):
        A, tr, c = t.parameters()
        return (p - c) @ A.T + tr + c

    want = apply(ts[1], apply(ts[0], pts))
    for inv in (0, 1):
        out = str(tmp_path / f"c{inv}.mat")
        assert cmd_apply(["-d", "3", "-o", f"Linear[{out},{inv}]", "-t", paths[0], "-t", paths[1]]) == 0
        got = apply(aio.read_transform(out), pts)
        if inv:
            np.testing.assert_allclose(apply(aio.read_transform(out), want), pts, atol=1e-9)
        else:
            np.testing.assert_allclose(got, want, atol=1e-9)
    g = Geometry((4, 4, 4))
    with pytest.raises(ValueError):
        collapse_linear([ts[0], DisplacementFieldTransform(np.zeros((4, 4, 4, 3), np.float32), g)], 3)
