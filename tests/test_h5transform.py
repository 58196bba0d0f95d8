"""ITK HDF5 composite transform files: round trips through the HDF5 C
library and the structure h5dump reports (group and dataset names, string
and element types), which is the layout ITK's HDF5TransformIO reads.  CPU
only; skipped where no HDF5 library is installed."""
import ctypes as C
import os
import shutil
import subprocess

import numpy as np
import pytest

from ants_amd import h5transform as H
from ants_amd.apply_transforms import AffineTransform, DisplacementFieldTransform, Geometry

try:
    H.lib()
    HAVE_H5 = True
except RuntimeError:
    HAVE_H5 = False
pytestmark = pytest.mark.skipif(not HAVE_H5, reason="no HDF5 C library")

rng = np.random.default_rng(8)


def h5dump():
    for p in (shutil.which("h5dump"), "/opt/conda/bin/h5dump"):
        if p and os.path.exists(p):
            return p
    return None


def sample(D=3):
    A = np.eye(D) + 0.05 * rng.standard_normal((D, D))
    aff = AffineTransform(A, rng.standard_normal(D), rng.standard_normal(D))
    size = (7, 5, 4)[:D]
    geo = Geometry(size, [1.5, 0.8, 2.0][:D], [3.0, -1.0, 4.0][:D], np.eye(D)[::-1] if D == 2 else
                   np.array([[0, 1, 0], [-1, 0, 0], [0, 0, 1]], float))
    fld = DisplacementFieldTransform(rng.standard_normal(geo.shape + (D,)).astype(np.float32), geo)
    return aff, fld


@pytest.mark.parametrize("D", [2, 3])
@pytest.mark.parametrize("scalar", ["float", "double"])
def test_composite_round_trip(tmp_path, D, scalar):
    aff, fld = sample(D)
    path = str(tmp_path / "c.h5")
    H.write_composite(path, [aff, fld], scalar)
    back = H.read_composite(path)
    assert isinstance(back[0], AffineTransform) and isinstance(back[1], DisplacementFieldTransform)
    tol = 1e-6 if scalar == "float" else 0
    A, t, c = aff.parameters()
    np.testing.assert_allclose(back[0].matrix, A, atol=tol)
    np.testing.assert_allclose(back[0].translation, t, atol=tol * 10)
    np.testing.assert_array_equal(back[0].center, c)
    np.testing.assert_array_equal(back[1].field, fld.field)
    g = back[1].geometry
    assert tuple(g.size) == tuple(fld.geometry.size)
    np.testing.assert_array_equal(g.spacing, fld.geometry.spacing)
    np.testing.assert_array_equal(g.origin, fld.geometry.origin)
    np.testing.assert_array_equal(np.asarray(g.direction).reshape(D, D), fld.geometry.direction)


def test_inverted_affine_is_written_as_its_inverse(tmp_path):
    aff, _ = sample(3)
    inv = AffineTransform(aff.matrix, aff.translation, aff.center, invert=True)
    path = str(tmp_path / "i.h5")
    H.write_composite(path, [inv], "double")
    (b,) = H.read_composite(path)
    p = rng.standard_normal(3)
    A, t, c = aff.parameters()
    y = A @ (p - c) + t + c
    np.testing.assert_allclose(b.matrix @ (y - b.center) + b.translation + b.center, p, atol=1e-12)


@pytest.mark.skipif(h5dump() is None, reason="no h5dump")
def test_layout_as_h5dump_reports_it(tmp_path):
    aff, fld = sample(3)
    big = DisplacementFieldTransform(np.zeros((12, 11, 10, 3), np.float32), Geometry((10, 11, 12)))
    path = str(tmp_path / "l.h5")
    H.write_composite(path, [aff, fld, big], "float")
    out = subprocess.run([h5dump(), "-H", "-p", path], capture_output=True, text=True, check=True).stdout
    for name in ("ITKVersion", "HDFVersion", "OSName", "OSVersion"):
        assert f'DATASET "{name}"' in out
    assert 'GROUP "TransformGroup"' in out
    for i in range(4):
        assert f'GROUP "{i}"' in out
    assert out.count('DATASET "TransformType"') == 4
    assert out.count('DATASET "TransformParameters"') == 3
    assert out.count('DATASET "TransformFixedParameters"') == 3
    assert "STRSIZE H5T_VARIABLE" in out
    assert "H5T_IEEE_F32LE" in out and "H5T_IEEE_F64LE" in out
    assert "SIMPLE { ( 3960 ) / ( 3960 ) }" in out      # 10 x 11 x 12 x 3 parameters
    assert "DEFLATE" in out                              # the large dataset is compressed
    full = subprocess.run([h5dump(), "-d", "/TransformGroup/0/TransformType", path], capture_output=True,
                          text=True, check=True).stdout
    assert '"CompositeTransform_float_3_3"' in full
    full = subprocess.run([h5dump(), "-d", "/TransformGroup/2/TransformType", path], capture_output=True,
                          text=True, check=True).stdout
    assert '"DisplacementFieldTransform_float_3_3"' in full
    fixed = subprocess.run([h5dump(), "-d", "/TransformGroup/2/TransformFixedParameters", path],
                           capture_output=True, text=True, check=True).stdout
    # size, origin, spacing, direction (row-major)
    nums = [float(v) for v in fixed.split("DATA {")[1].split("}")[0].replace("(0):", "").replace("\n", " ")
            .replace("(", " ").replace(")", " ").replace(":", " ").split(",")]
    assert nums[:9] == [7, 5, 4, 3.0, -1.0, 4.0, 1.5, 0.8, 2.0]
    assert nums[9:] == [0, 1, 0, -1, 0, 0, 0, 0, 1]


def test_reads_fixed_length_strings(tmp_path):
    """Files whose TransformType is a fixed-length string read the same."""
    aff, _ = sample(2)
    path = str(tmp_path / "f.h5")
    L = H.lib()
    with H._File(path, write=True) as f:
        f.group("/TransformGroup")
        f.group("/TransformGroup/0")
        s = "AffineTransform_double_2_2"
        t = L.H5Tcopy(L.T_C_S1)
        L.H5Tset_size(t, len(s))
        dims = (C.c_uint64 * 1)(1)
        sp = L.H5Screate_simple(1, dims, None)
        d = L.H5Dcreate2(f.id, b"/TransformGroup/0/TransformType", t, sp, 0, 0, 0)
        L.H5Dwrite(d, t, 0, 0, 0, C.cast(C.c_char_p(s.encode()), C.c_void_p))
        L.H5Dclose(d)
        L.H5Sclose(sp)
        L.H5Tclose(t)
        A, tr, c = aff.parameters()
        f.array("/TransformGroup/0/TransformFixedParameters", c)
        f.array("/TransformGroup/0/TransformParameters", np.concatenate([A.ravel(), tr]))
    (b,) = H.read_composite(path)
    np.testing.assert_array_equal(b.matrix, A)


def test_syn_state_round_trip(tmp_path):
    geo = Geometry((6, 5, 4), [1.0, 1.2, 0.9], [0.5, 0.0, -2.0])
    state = [rng.standard_normal(geo.shape + (3,)).astype(np.float32) for _ in range(4)]
    aff, _ = sample(3)
    path = str(tmp_path / "s.h5")
    H.write_syn_state(path, state, geo, [aff])
    lin, fields, g = H.read_syn_state(path)
    assert len(lin) == 1 and isinstance(lin[0], AffineTransform)
    for a, b in zip(fields, state):
        np.testing.assert_array_equal(a, b)
    assert tuple(g.size) == geo.size
    with pytest.raises(ValueError):
        H.write_composite(str(tmp_path / "x.h5"), [aff], "half")
    H.write_composite(str(tmp_path / "y.h5"), [aff], "float")
    with pytest.raises(ValueError):
        H.read_syn_state(str(tmp_path / "y.h5"))
