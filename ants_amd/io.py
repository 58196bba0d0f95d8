"""File formats around the SyN stage, host side: NIfTI-1 images and
displacement fields, and ITK affine transform files.

Conventions follow what ANTs writes through ITK (Examples/
antsRegistrationTemplateHeader.h:1405-1476, Utilities/
itkantsReadWriteTransform.h:130-270):

* Geometry is ITK's, in LPS physical space; NIfTI stores RAS, so the first
  two rows of the direction matrix and of the origin change sign on the way
  in and out.
* Displacement fields are 5-D NIfTI vector images (dim = [5, nx, ny, nz, 1,
  D], intent NIFTI_INTENT_VECTOR = 1007), components stored planar (the
  component axis is the slowest), values in LPS millimetres as ITK keeps
  them.  In memory they are (z, y, x, D) interleaved, the layout the
  library's kernels use.
* Affine transforms are MatrixOffsetTransformBase / AffineTransform files:
  text ("#Insight Transform File V1.0", Parameters = row-major matrix then
  translation, FixedParameters = centre; as Scripts/ANTSAverage3DAffine.sh
  writes them) or MATLAB v4 ".mat" (a parameter vector named after the
  transform class and a "fixed" vector).
"""
from __future__ import annotations

import gzip
import os
import re
import struct
from typing import Optional, Tuple

import numpy as np

from .apply_transforms import AffineTransform, Geometry

NIFTI_INTENT_VECTOR = 1007
_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8, 512: np.uint16,
           768: np.uint32, 1024: np.int64, 1280: np.uint64}
_CODES = {np.dtype(v): k for k, v in _DTYPES.items()}
_LPS = np.diag([-1.0, -1.0, 1.0])


def _open(path, mode):
    return gzip.open(path, mode) if str(path).endswith(".gz") else open(path, mode)


# ---------------------------------------------------------------- quaternions
def _quatern_to_rotation(b: float, c: float, d: float) -> np.ndarray:
    """NIfTI-1 qform rotation from the quaternion (b, c, d), a = sqrt(1 - b^2 - c^2 - d^2)."""
    a = 1.0 - (b * b + c * c + d * d)
    a = np.sqrt(a) if a > 1e-7 else 0.0
    if a == 0.0:  # 180-degree rotation: normalise (b, c, d)
        n = np.sqrt(b * b + c * c + d * d)
        b, c, d = b / n, c / n, d / n
    return np.array([[a * a + b * b - c * c - d * d, 2 * (b * c - a * d), 2 * (b * d + a * c)],
                     [2 * (b * c + a * d), a * a + c * c - b * b - d * d, 2 * (c * d - a * b)],
                     [2 * (b * d - a * c), 2 * (c * d + a * b), a * a + d * d - c * c - b * b]])


def _rotation_to_quatern(R: np.ndarray) -> Tuple[float, float, float, float]:
    """(b, c, d, qfac) of an orthonormal matrix: a proper rotation R', with R =
    R' diag(1, 1, qfac).  Shepperd's method on the largest diagonal term."""
    R = np.asarray(R, dtype=np.float64).copy()
    qfac = 1.0
    if np.linalg.det(R) < 0:
        qfac = -1.0
        R[:, 2] = -R[:, 2]
    t = np.trace(R)
    if t > 0:
        s = 2.0 * np.sqrt(1.0 + t)
        a, b, c, d = 0.25 * s, (R[2, 1] - R[1, 2]) / s, (R[0, 2] - R[2, 0]) / s, (R[1, 0] - R[0, 1]) / s
    elif R[0, 0] > R[1, 1] and R[0, 0] > R[2, 2]:
        s = 2.0 * np.sqrt(1.0 + R[0, 0] - R[1, 1] - R[2, 2])
        a, b, c, d = (R[2, 1] - R[1, 2]) / s, 0.25 * s, (R[0, 1] + R[1, 0]) / s, (R[0, 2] + R[2, 0]) / s
    elif R[1, 1] > R[2, 2]:
        s = 2.0 * np.sqrt(1.0 + R[1, 1] - R[0, 0] - R[2, 2])
        a, b, c, d = (R[0, 2] - R[2, 0]) / s, (R[0, 1] + R[1, 0]) / s, 0.25 * s, (R[1, 2] + R[2, 1]) / s
    else:
        s = 2.0 * np.sqrt(1.0 + R[2, 2] - R[0, 0] - R[1, 1])
        a, b, c, d = (R[1, 0] - R[0, 1]) / s, (R[0, 2] + R[2, 0]) / s, (R[1, 2] + R[2, 1]) / s, 0.25 * s
    if a < 0:
        b, c, d = -b, -c, -d
    return float(b), float(c), float(d), qfac


# ---------------------------------------------------------------------- NIfTI
def _geometry_to_ras(geom: Geometry) -> Tuple[np.ndarray, np.ndarray]:
    """3x3 RAS direction and RAS origin (3-vectors, a 2-D image padded to 3-D)."""
    D = geom.dim
    dirm = np.eye(3)
    if geom.direction is not None and np.any(np.asarray(geom.direction) != 0):
        dirm[:D, :D] = np.asarray(geom.direction, dtype=np.float64).reshape(D, D)
    org = np.zeros(3)
    if geom.origin is not None:
        org[:D] = np.asarray(geom.origin, dtype=np.float64)
    return _LPS @ dirm, _LPS @ org


def write_nifti(path: str, data: np.ndarray, geom: Geometry, vector: bool = False) -> None:
    """Write a scalar image (numpy (z,y,x) / (y,x)) or, with vector=True, a
    vector image ((z,y,x,K) interleaved: a displacement field has K = D) as
    a NIfTI-1 single file."""
    D = geom.dim
    a = np.asarray(data)
    if vector:
        if a.ndim != D + 1 or a.shape[:-1] != geom.shape:
            raise ValueError(f"vector image shape {a.shape} does not match {geom.shape} + (components,)")
        planar = np.moveaxis(a, -1, 0)  # (K, z, y, x): component slowest
    else:
        if a.shape != geom.shape:
            raise ValueError(f"image shape {a.shape} does not match {geom.shape}")
        planar = a
    if planar.dtype not in _CODES:
        planar = planar.astype(np.float32)
    code = _CODES[planar.dtype]
    size = list(geom.size) + [1] * (3 - D)
    sp = list(geom.spacing if geom.spacing is not None else [1.0] * D) + [1.0] * (3 - D)
    dims = [5, size[0], size[1], size[2], 1, a.shape[-1] if vector else 1, 1, 1] if vector else \
        [D, size[0], size[1], size[2] if D == 3 else 1,
                                                                       1, 1, 1, 1]
    R, org = _geometry_to_ras(geom)
    b, c, d, qfac = _rotation_to_quatern(R)
    affine = R * np.asarray(sp, dtype=np.float64)  # columns scaled by spacing
    hdr = bytearray(352)
    struct.pack_into("<i", hdr, 0, 348)
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<h", hdr, 68, NIFTI_INTENT_VECTOR if vector else 0)
    struct.pack_into("<hh", hdr, 70, code, planar.dtype.itemsize * 8)
    struct.pack_into("<8f", hdr, 76, qfac, sp[0], sp[1], sp[2], 1.0, 1.0, 1.0, 1.0)
    struct.pack_into("<f", hdr, 108, 352.0)
    struct.pack_into("<f", hdr, 112, 1.0)  # scl_slope
    hdr[123] = 2 | 8  # mm, seconds
    struct.pack_into("<hh", hdr, 252, 1, 1)  # qform_code, sform_code: scanner anatomical
    struct.pack_into("<3f", hdr, 256, b, c, d)
    struct.pack_into("<3f", hdr, 268, *org)
    for r in range(3):
        struct.pack_into("<4f", hdr, 280 + 16 * r, affine[r, 0], affine[r, 1], affine[r, 2], org[r])
    if vector:
        hdr[328:328 + 6] = b"vector"
    hdr[344:348] = b"n+1\0"
    with _open(path, "wb") as fh:
        fh.write(bytes(hdr))
        fh.write(np.ascontiguousarray(planar).astype(planar.dtype.newbyteorder("<"), copy=False).tobytes())


def read_nifti(path: str) -> Tuple[np.ndarray, Geometry]:
    """Read a NIfTI-1 image.  Scalars come back as numpy (z,y,x) / (y,x) in
    the file's data type (scl_slope/inter applied as float32 when set);
    5-D vector images as (z,y,x,K) interleaved.  Geometry in LPS."""
    with _open(path, "rb") as fh:
        raw = fh.read()
    return read_nifti_from_bytes(raw)


def read_nifti_from_bytes(raw: bytes) -> Tuple[np.ndarray, Geometry]:
    """read_nifti of a file's bytes."""
    if len(raw) < 348:
        raise ValueError("not a NIfTI-1 file")
    end = "<" if struct.unpack_from("<i", raw, 0)[0] == 348 else ">"
    if struct.unpack_from(end + "i", raw, 0)[0] != 348 or raw[344:347] not in (b"n+1", b"ni1"):
        raise ValueError("not a single-file NIfTI-1 image")
    dims = struct.unpack_from(end + "8h", raw, 40)
    intent = struct.unpack_from(end + "h", raw, 68)[0]
    code = struct.unpack_from(end + "h", raw, 70)[0]
    pixdim = struct.unpack_from(end + "8f", raw, 76)
    vox_offset = int(struct.unpack_from(end + "f", raw, 108)[0])
    slope, inter = struct.unpack_from(end + "2f", raw, 112)
    qcode, scode = struct.unpack_from(end + "2h", raw, 252)
    if code not in _DTYPES:
        raise ValueError(f"unsupported NIfTI datatype {code}")
    nd = dims[0]
    vector = nd == 5 and dims[5] > 1
    if vector and intent not in (0, NIFTI_INTENT_VECTOR):
        raise ValueError(f"unsupported 5-D intent {intent}")
    n = [max(int(v), 1) for v in dims[1:4]]
    # a single-slice vector image is 2-D when it holds 2 (a 2-D field) or 4
    # (a 2 x 2 matrix) components
    D = 2 if (n[2] == 1 and (nd == 2 or (vector and dims[5] in (2, 4)))) else 3
    ncomp = dims[5] if vector else 1
    dt = np.dtype(_DTYPES[code]).newbyteorder(end)
    count = n[0] * n[1] * n[2] * ncomp
    a = np.frombuffer(raw, dtype=dt, count=count, offset=vox_offset).astype(dt.newbyteorder("="))
    if slope not in (0.0, 1.0) or inter != 0.0:
        a = a.astype(np.float32) * np.float32(slope if slope != 0.0 else 1.0) + np.float32(inter)
    # geometry: sform when present, else qform, else pixdim only (RAS), then to LPS
    sp = np.array([abs(pixdim[1]) or 1.0, abs(pixdim[2]) or 1.0, abs(pixdim[3]) or 1.0])
    if scode > 0:
        srow = np.array([struct.unpack_from(end + "4f", raw, 280 + 16 * r) for r in range(3)], dtype=np.float64)
        M, org_ras = srow[:, :3], srow[:, 3]
        sp = np.linalg.norm(M, axis=0)
        R = M / sp
    elif qcode > 0:
        b, c, d = struct.unpack_from(end + "3f", raw, 256)
        org_ras = np.array(struct.unpack_from(end + "3f", raw, 268), dtype=np.float64)
        R = _quatern_to_rotation(b, c, d)
        if pixdim[0] < 0:
            R[:, 2] = -R[:, 2]
    else:
        R, org_ras = np.eye(3), np.zeros(3)
    R_lps, org_lps = _LPS @ R, _LPS @ org_ras
    if vector:
        a = np.moveaxis(a.reshape((ncomp, n[2], n[1], n[0])), 0, -1)
        if D == 2:
            a = a[0]
    else:
        a = a.reshape((n[2], n[1], n[0]))
        if D == 2:
            a = a[0]
    geom = Geometry(tuple(n[:D]), tuple(float(v) for v in sp[:D]), tuple(float(v) for v in org_lps[:D]),
                    R_lps[:D, :D].copy())
    return a, geom


def write_nifti_series(path: str, data: np.ndarray, geom: Geometry, t_spacing: float = 1.0,
                       t_origin: float = 0.0) -> None:
    """A 3-D time series (numpy (t, z, y, x)) as a 4-D NIfTI-1 file: the
    spatial header from `geom`, pixdim[4] = t_spacing, toffset = t_origin
    (ITK's fourth origin component)."""
    a = np.asarray(data)
    if geom.dim != 3 or a.ndim != 4 or a.shape[1:] != geom.shape:
        raise ValueError(f"series shape {a.shape} does not match (t,) + {geom.shape}")
    write_nifti(path, a[0], geom)   # header template (one volume)
    with _open(path, "rb") as fh:
        raw = bytearray(fh.read())
    vox = int(struct.unpack_from("<f", raw, 108)[0])
    hdr = raw[:vox]
    dims = list(struct.unpack_from("<8h", hdr, 40))
    dims[0], dims[4] = 4, a.shape[0]
    struct.pack_into("<8h", hdr, 40, *dims)
    struct.pack_into("<f", hdr, 76 + 4 * 4, float(t_spacing))
    struct.pack_into("<f", hdr, 136, float(t_origin))
    planar = a if a.dtype in _CODES else a.astype(np.float32)
    if _CODES[planar.dtype] != struct.unpack_from("<h", hdr, 70)[0]:
        raise ValueError("series data type differs from its first volume's")
    with _open(path, "wb") as fh:
        fh.write(bytes(hdr))
        fh.write(np.ascontiguousarray(planar).astype(planar.dtype.newbyteorder("<"), copy=False).tobytes())


def read_nifti_series(path: str):
    """A 4-D NIfTI-1 scalar image as (numpy (t, z, y, x), spatial Geometry,
    t_spacing, t_origin); a 3-D image reads as a series of one."""
    with _open(path, "rb") as fh:
        raw = fh.read()
    end = "<" if struct.unpack_from("<i", raw, 0)[0] == 348 else ">"
    dims = list(struct.unpack_from(end + "8h", raw, 40))
    nt = int(dims[4]) if dims[0] >= 4 else 1
    t_sp = float(struct.unpack_from(end + "f", raw, 76 + 16)[0]) if dims[0] >= 4 else 1.0
    t_org = float(struct.unpack_from(end + "f", raw, 136)[0]) if dims[0] >= 4 else 0.0
    if dims[0] >= 5 and dims[5] > 1:
        raise ValueError("a vector image is not a time series")
    # read the spatial part through the 3-D reader with the time axis folded away
    hdr = bytearray(raw)
    d3 = list(dims)
    d3[0], d3[4] = 3, 1
    struct.pack_into(end + "8h", hdr, 40, *d3)
    vol, geom = read_nifti_from_bytes(bytes(hdr))
    if geom.dim != 3:
        raise ValueError("time series are 3-D in space")
    code = struct.unpack_from(end + "h", raw, 70)[0]
    dt = np.dtype(_DTYPES[code]).newbyteorder(end)
    vox_offset = int(struct.unpack_from(end + "f", raw, 108)[0])
    n = geom.shape
    a = np.frombuffer(raw, dtype=dt, count=nt * int(np.prod(n)), offset=vox_offset).astype(dt.newbyteorder("="))
    slope, inter = struct.unpack_from(end + "2f", raw, 112)
    if slope not in (0.0, 1.0) or inter != 0.0:
        a = a.astype(np.float32) * np.float32(slope if slope != 0.0 else 1.0) + np.float32(inter)
    return a.reshape((nt,) + n), geom, (t_sp if t_sp > 0 else 1.0), t_org


# ------------------------------------------------------------ transform files
def _affine_parts(params, fixed, D):
    p = np.asarray(params, dtype=np.float64).ravel()
    if p.size != D * D + D:
        raise ValueError(f"an affine transform of dimension {D} has {D * D + D} parameters, got {p.size}")
    c = np.zeros(D) if fixed is None or len(fixed) == 0 else np.asarray(fixed, dtype=np.float64).ravel()[:D]
    return AffineTransform(p[:D * D].reshape(D, D), p[D * D:], c)


def read_transform(path: str) -> AffineTransform:
    """Read an affine transform file (text or MATLAB v4 .mat)."""
    if str(path).endswith(".mat"):
        from scipy.io import loadmat
        m = loadmat(path)
        names = [k for k in m if not k.startswith("__")]
        pname = next((k for k in names if re.search(r"_(\d)_\1$", k)), None)
        if pname is None:
            raise ValueError(f"{path}: no transform parameters (found {names})")
        D = int(pname[-1])
        return _affine_parts(m[pname], m.get("fixed"), D)
    text = open(path).read()
    mt = re.search(r"Transform:\s*(\w+?)_(?:double|float)_(\d)_\2", text)
    mp = re.search(r"^Parameters:\s*(.*)$", text, re.M)
    mf = re.search(r"^FixedParameters:\s*(.*)$", text, re.M)
    if not (mt and mp):
        raise ValueError(f"{path}: not an ITK affine transform file")
    if mt.group(1) not in ("MatrixOffsetTransformBase", "AffineTransform"):
        raise ValueError(f"{path}: transform {mt.group(1)} is not supported (affine only)")
    D = int(mt.group(2))
    params = [float(v) for v in mp.group(1).split()]
    fixed = [float(v) for v in mf.group(1).split()] if mf else []
    return _affine_parts(params, fixed, D)


def write_transform(path: str, t: AffineTransform) -> None:
    """Write an AffineTransform as text (.txt/.tfm) or MATLAB v4 (.mat)."""
    A, tr, c = t.parameters()
    D = A.shape[0]
    name = f"AffineTransform_double_{D}_{D}"
    params = np.concatenate([A.ravel(), tr])
    if str(path).endswith(".mat"):
        from scipy.io import savemat
        savemat(path, {name: params.reshape(-1, 1), "fixed": np.asarray(c, dtype=np.float64).reshape(-1, 1)},
                format="4")
        return
    with open(path, "w") as fh:
        fh.write("#Insight Transform File V1.0\n#Transform 0\n")
        fh.write(f"Transform: {name}\n")
        fh.write("Parameters: " + " ".join(repr(float(v)) for v in params) + "\n")
        fh.write("FixedParameters: " + " ".join(repr(float(v)) for v in c) + "\n")


def read_points_csv(path: str):
    """A point file as antsApplyTransformsToPoints reads it
    (Examples/antsApplyTransformsToPoints.cxx:57-85): a header row of
    column names (x,y,z,t,label,...), then comma-separated rows; cells that
    are not numbers read as NaN.  Returns (names, float64 array)."""
    import csv
    with open(path, newline="") as fh:
        rows = [r for r in csv.reader(fh, delimiter=",", quotechar='"') if r]
    if not rows:
        raise ValueError(f"{path}: empty point file")
    names = [c.strip() for c in rows[0]]

    def num(v):
        try:
            return float(v)
        except ValueError:
            return float("nan")

    data = np.array([[num(v) for v in r] + [float("nan")] * (len(names) - len(r)) for r in rows[1:]],
                    dtype=np.float64).reshape(-1, len(names))
    return names, data


def write_points_csv(path: str, names, data) -> None:
    """The header row, then one row per point (numbers in full precision)."""
    data = np.asarray(data, dtype=np.float64)
    with open(path, "w") as fh:
        fh.write(",".join(names) + "\n")
        for row in data:
            fh.write(",".join(repr(float(v)) for v in row) + "\n")
