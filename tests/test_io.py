"""File formats (CPU only): NIfTI-1 round trips of images and displacement
fields with ITK's LPS geometry, the header fields ITK-written files carry,
qform-only headers, and affine transform files (text as the reference's
scripts write them, MATLAB v4 through scipy.io)."""
import struct

import numpy as np
import pytest

from ants_amd import io as aio
from ants_amd.apply_transforms import AffineTransform, Geometry

rng = np.random.default_rng(3)


def rot(a, b, c):
    rx = np.array([[1, 0, 0], [0, np.cos(a), -np.sin(a)], [0, np.sin(a), np.cos(a)]])
    ry = np.array([[np.cos(b), 0, np.sin(b)], [0, 1, 0], [-np.sin(b), 0, np.cos(b)]])
    rz = np.array([[np.cos(c), -np.sin(c), 0], [np.sin(c), np.cos(c), 0], [0, 0, 1]])
    return rz @ ry @ rx


@pytest.mark.parametrize("ext", [".nii", ".nii.gz"])
@pytest.mark.parametrize("improper", [False, True])
def test_scalar_round_trip(tmp_path, ext, improper):
    R = rot(0.3, -0.2, 0.1)
    if improper:
        R = R @ np.diag([1, 1, -1])
    g = Geometry((9, 8, 7), (1.2, 0.8, 2.5), (10.0, -20.0, 30.0), R)
    img = rng.standard_normal(g.shape).astype(np.float32)
    p = str(tmp_path / ("img" + ext))
    aio.write_nifti(p, img, g)
    back, g2 = aio.read_nifti(p)
    np.testing.assert_array_equal(back, img)
    assert g2.size == g.size
    np.testing.assert_allclose(g2.spacing, g.spacing, rtol=1e-6)
    np.testing.assert_allclose(g2.origin, g.origin, atol=1e-5)
    np.testing.assert_allclose(g2.direction, R, atol=1e-6)


def test_displacement_field_layout_and_header(tmp_path):
    g = Geometry((6, 5, 4), (1.0, 1.0, 1.0), (1.0, 2.0, 3.0), np.eye(3))
    fld = rng.standard_normal(g.shape + (3,)).astype(np.float32)
    p = str(tmp_path / "warp.nii.gz")
    aio.write_nifti(p, fld, g, vector=True)
    back, g2 = aio.read_nifti(p)
    np.testing.assert_array_equal(back, fld)
    import gzip
    raw = gzip.open(p).read()
    assert struct.unpack_from("<i", raw, 0)[0] == 348 and raw[344:348] == b"n+1\0"
    assert struct.unpack_from("<8h", raw, 40)[:6] == (5, 6, 5, 4, 1, 3)
    assert struct.unpack_from("<h", raw, 68)[0] == aio.NIFTI_INTENT_VECTOR
    assert struct.unpack_from("<f", raw, 108)[0] == 352.0
    # LPS origin (1, 2, 3) is RAS (-1, -2, 3): qoffset and the sform translation
    assert struct.unpack_from("<3f", raw, 268) == (-1.0, -2.0, 3.0)
    assert [struct.unpack_from("<4f", raw, 280 + 16 * r)[3] for r in range(3)] == [-1.0, -2.0, 3.0]
    # components planar: the first voxel block of the data is the x component
    data = np.frombuffer(raw, "<f4", offset=352)
    np.testing.assert_array_equal(data[:fld[..., 0].size], fld[..., 0].ravel())


def test_two_dimensional_images_and_fields(tmp_path):
    g = Geometry((7, 5), (0.5, 2.0), (1.0, -1.0), np.array([[0.0, 1.0], [1.0, 0.0]]))
    img = rng.integers(0, 9, g.shape).astype(np.int16)
    aio.write_nifti(str(tmp_path / "a.nii"), img, g)
    back, g2 = aio.read_nifti(str(tmp_path / "a.nii"))
    assert back.dtype == np.int16 and g2.dim == 2
    np.testing.assert_array_equal(back, img)
    np.testing.assert_allclose(g2.direction, g.direction, atol=1e-7)
    fld = rng.standard_normal(g.shape + (2,)).astype(np.float32)
    aio.write_nifti(str(tmp_path / "f.nii"), fld, g, vector=True)
    fb, gf = aio.read_nifti(str(tmp_path / "f.nii"))
    np.testing.assert_array_equal(fb, fld)
    assert gf.dim == 2


def test_qform_only_header_quaternion(tmp_path):
    """A header with sform_code 0: the geometry comes from the quaternion
    (a 90-degree rotation about z: b = c = 0, d = sin 45)."""
    g = Geometry((4, 4, 4), (2.0, 2.0, 2.0), (0.0, 0.0, 0.0), np.eye(3))
    p = str(tmp_path / "q.nii")
    aio.write_nifti(p, np.zeros(g.shape, np.float32), g)
    raw = bytearray(open(p, "rb").read())
    struct.pack_into("<h", raw, 254, 0)
    struct.pack_into("<3f", raw, 256, 0.0, 0.0, np.sin(np.pi / 4))
    open(p, "wb").write(bytes(raw))
    _, g2 = aio.read_nifti(p)
    rz = np.array([[0.0, -1.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 1.0]])   # RAS
    np.testing.assert_allclose(g2.direction, np.diag([-1, -1, 1]) @ rz, atol=1e-6)


def test_quaternion_round_trip_of_random_rotations():
    for _ in range(20):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        b, c, d, qfac = aio._rotation_to_quatern(q)
        R = aio._quatern_to_rotation(b, c, d)
        R[:, 2] *= qfac
        np.testing.assert_allclose(R, q, atol=1e-12)


def test_text_transform_as_the_reference_scripts_write_it(tmp_path):
    p = tmp_path / "avg.txt"
    p.write_text("#Insight Transform File V1.0 \n# Transform 0 \nTransform: MatrixOffsetTransformBase_double_3_3  \n"
                 "Parameters:  1 0.1 0 0 1 0 0 0 1  2 3 4  \nFixedParameters: 5 6 7 \n")
    t = aio.read_transform(str(p))
    np.testing.assert_array_equal(t.matrix, [[1, 0.1, 0], [0, 1, 0], [0, 0, 1]])
    np.testing.assert_array_equal(t.translation, [2, 3, 4])
    np.testing.assert_array_equal(t.center, [5, 6, 7])


@pytest.mark.parametrize("ext", [".txt", ".mat"])
def test_transform_round_trip(tmp_path, ext):
    A = np.eye(3) + 0.05 * rng.standard_normal((3, 3))
    t = AffineTransform(A, [1.5, -2.0, 0.25], [10.0, 20.0, 30.0])
    p = str(tmp_path / ("t" + ext))
    aio.write_transform(p, t)
    back = aio.read_transform(p)
    np.testing.assert_array_equal(back.matrix, A)
    np.testing.assert_array_equal(back.translation, t.translation)
    np.testing.assert_array_equal(back.center, t.center)
    if ext == ".mat":
        from scipy.io import loadmat
        m = loadmat(p)
        assert m["AffineTransform_double_3_3"].shape == (12, 1) and m["fixed"].shape == (3, 1)


def test_nifti_series_round_trip(tmp_path):
    """4-D NIfTI: the spatial header, pixdim[4] (time spacing) and toffset
    survive a round trip; a 3-D image reads as a series of one."""
    from ants_amd.io import read_nifti_series, write_nifti_series
    rng = np.random.default_rng(2)
    th = 0.3
    R = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    g = Geometry((7, 6, 5), (1.5, 1.0, 2.0), (3.0, -4.0, 5.0), R)
    data = rng.standard_normal((4,) + g.shape).astype(np.float32)
    p = str(tmp_path / "ts.nii.gz")
    write_nifti_series(p, data, g, 2.5, 1.25)
    a, g2, tsp, tor = read_nifti_series(p)
    np.testing.assert_array_equal(a, data)
    assert (tsp, tor) == (2.5, 1.25)
    assert g2.size == g.size
    np.testing.assert_allclose(g2.spacing, g.spacing, atol=1e-6)
    np.testing.assert_allclose(g2.origin, g.origin, atol=1e-5)
    np.testing.assert_allclose(g2.direction, R, atol=1e-6)
    from ants_amd.io import write_nifti
    p3 = str(tmp_path / "v.nii.gz")
    write_nifti(p3, data[0], g)
    a3, _, tsp3, _ = read_nifti_series(p3)
    assert a3.shape == (1,) + g.shape and tsp3 == 1.0
