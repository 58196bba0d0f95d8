"""ITK HDF5 composite transform files (``.h5``), host side, through the HDF5
C library (libhdf5 1.10, loaded with ctypes from the Python installation's
prefix or ``ANTS_HDF5_LIB``).

ANTs writes and reads composite transforms with ITK's HDF5TransformIO
(``Utilities/itkantsReadWriteTransform.h:130-270``): the composite output of
``--write-composite-transform`` and the registration state of
``--save-state`` / ``--restore-state``
(``Examples/antsRegistrationTemplateHeader.h:1308-1348, 293-327``,
``Examples/itkantsRegistrationHelper.hxx:3637-3700``).  The layout restated
here (ITK v5.4.6's HDF5TransformIO, absent from the reference tree):

* ``/TransformGroup/<i>`` per transform of the list, i = 0, 1, ...; a
  composite transform is entry 0 and holds no parameters, its components
  follow in queue order (``GetNthTransform(0)`` first: the transform added
  first, applied last);
* in each group ``TransformType`` (a variable-length string dataset of one
  element, ``<Class>_<float|double>_<D>_<D>``), ``TransformFixedParameters``
  (rank-1 double) and ``TransformParameters`` (rank-1, the transform's scalar
  type; chunked and deflated when large);
* string datasets ``/ITKVersion``, ``/HDFVersion``, ``/OSName``,
  ``/OSVersion`` at the root.

Parameters: AffineTransform / MatrixOffsetTransformBase: the row-major
matrix then the translation, fixed parameters = centre.
DisplacementFieldTransform: the field, x fastest, components interleaved;
fixed parameters = size, origin, spacing, direction (row-major).
"""
from __future__ import annotations

import ctypes as C
import os
import platform
import sys
from typing import List, Sequence, Tuple

import numpy as np

from .apply_transforms import AffineTransform, DisplacementFieldTransform, Geometry

_lib = None
H5P_DEFAULT = 0
H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5T_VARIABLE = C.c_size_t(-1).value
H5S_ALL = 0


def _candidates():
    env = os.environ.get("ANTS_HDF5_LIB")
    if env:
        yield env
    for prefix in (sys.prefix, sys.base_prefix, "/opt/conda", "/usr"):
        for sub in ("lib", "lib/x86_64-linux-gnu"):
            for name in ("libhdf5.so", "libhdf5_serial.so"):
                yield os.path.join(prefix, sub, name)


def lib():
    """The HDF5 C library; RuntimeError when none is installed."""
    global _lib
    if _lib is None:
        for path in _candidates():
            if os.path.exists(path):
                L = C.CDLL(path)
                break
        else:
            raise RuntimeError("HDF5 C library not found (set ANTS_HDF5_LIB to libhdf5.so)")
        if L.H5open() < 0:
            raise RuntimeError("H5open failed")
        hid, herr, sz = C.c_int64, C.c_int, C.c_size_t
        p = C.POINTER
        sig = {
            "H5Fcreate": ([C.c_char_p, C.c_uint, hid, hid], hid),
            "H5Fopen": ([C.c_char_p, C.c_uint, hid], hid),
            "H5Fclose": ([hid], herr),
            "H5Gcreate2": ([hid, C.c_char_p, hid, hid, hid], hid),
            "H5Gclose": ([hid], herr),
            "H5Screate_simple": ([C.c_int, p(C.c_uint64), p(C.c_uint64)], hid),
            "H5Sget_simple_extent_npoints": ([hid], C.c_int64),
            "H5Sclose": ([hid], herr),
            "H5Tcopy": ([hid], hid),
            "H5Tset_size": ([hid, sz], herr),
            "H5Tget_size": ([hid], sz),
            "H5Tis_variable_str": ([hid], C.c_int),
            "H5Tget_class": ([hid], C.c_int),
            "H5Tclose": ([hid], herr),
            "H5Pcreate": ([hid], hid),
            "H5Pset_chunk": ([hid, C.c_int, p(C.c_uint64)], herr),
            "H5Pset_deflate": ([hid, C.c_uint], herr),
            "H5Pclose": ([hid], herr),
            "H5Dcreate2": ([hid, C.c_char_p, hid, hid, hid, hid, hid], hid),
            "H5Dopen2": ([hid, C.c_char_p, hid], hid),
            "H5Dwrite": ([hid, hid, hid, hid, hid, C.c_void_p], herr),
            "H5Dread": ([hid, hid, hid, hid, hid, C.c_void_p], herr),
            "H5Dget_space": ([hid], hid),
            "H5Dget_type": ([hid], hid),
            "H5Dvlen_reclaim": ([hid, hid, hid, C.c_void_p], herr),
            "H5Dclose": ([hid], herr),
            "H5Lexists": ([hid, C.c_char_p, hid], C.c_int),
            "H5Eset_auto2": ([hid, C.c_void_p, C.c_void_p], herr),
            "H5get_libversion": ([p(C.c_uint), p(C.c_uint), p(C.c_uint)], herr),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        g = lambda n: C.c_int64.in_dll(L, n).value  # noqa: E731
        L.T_C_S1 = g("H5T_C_S1_g")
        L.T_DOUBLE = g("H5T_NATIVE_DOUBLE_g")
        L.T_FLOAT = g("H5T_NATIVE_FLOAT_g")
        L.P_DATASET_CREATE = g("H5P_CLS_DATASET_CREATE_ID_g")
        L.H5Eset_auto2(0, None, None)   # errors come back as return codes
        _lib = L
    return _lib


def _check(v, what):
    if v < 0:
        raise OSError(f"HDF5: {what} failed")
    return v


class _File:
    def __init__(self, path: str, write: bool):
        L = lib()
        self.L = L
        if write:
            self.id = _check(L.H5Fcreate(path.encode(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), f"create {path}")
        else:
            self.id = _check(L.H5Fopen(path.encode(), H5F_ACC_RDONLY, H5P_DEFAULT), f"open {path}")

    def close(self):
        self.L.H5Fclose(self.id)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- writing
    def group(self, name: str):
        self.L.H5Gclose(_check(self.L.H5Gcreate2(self.id, name.encode(), H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT),
                               f"group {name}"))

    def string(self, name: str, value: str):
        L = self.L
        t = _check(L.H5Tcopy(L.T_C_S1), "H5Tcopy")
        L.H5Tset_size(t, H5T_VARIABLE)
        dims = (C.c_uint64 * 1)(1)
        s = L.H5Screate_simple(1, dims, None)
        d = _check(L.H5Dcreate2(self.id, name.encode(), t, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT), f"dataset {name}")
        buf = (C.c_char_p * 1)(value.encode())
        _check(L.H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(buf, C.c_void_p)), f"write {name}")
        L.H5Dclose(d)
        L.H5Sclose(s)
        L.H5Tclose(t)

    def array(self, name: str, values: np.ndarray, compress: bool = True):
        L = self.L
        a = np.ascontiguousarray(values).ravel()
        mem = L.T_DOUBLE if a.dtype == np.float64 else L.T_FLOAT
        dims = (C.c_uint64 * 1)(a.size)
        s = L.H5Screate_simple(1, dims, None)
        plist = H5P_DEFAULT
        if compress and a.size > 1024:
            plist = L.H5Pcreate(L.P_DATASET_CREATE)
            chunk = (C.c_uint64 * 1)(min(a.size, 1 << 20))
            L.H5Pset_chunk(plist, 1, chunk)
            L.H5Pset_deflate(plist, 6)
        d = _check(L.H5Dcreate2(self.id, name.encode(), mem, s, H5P_DEFAULT, plist, H5P_DEFAULT), f"dataset {name}")
        _check(L.H5Dwrite(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), f"write {name}")
        L.H5Dclose(d)
        L.H5Sclose(s)
        if plist != H5P_DEFAULT:
            L.H5Pclose(plist)

    # ---- reading
    def exists(self, name: str) -> bool:
        # check each prefix: H5Lexists needs the intermediate groups to exist
        parts = [p for p in name.split("/") if p]
        for k in range(1, len(parts) + 1):
            if self.L.H5Lexists(self.id, ("/" + "/".join(parts[:k])).encode(), H5P_DEFAULT) <= 0:
                return False
        return True

    def read_string(self, name: str) -> str:
        L = self.L
        d = _check(L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), f"open {name}")
        ft = L.H5Dget_type(d)
        try:
            if L.H5Tis_variable_str(ft) > 0:
                t = L.H5Tcopy(L.T_C_S1)
                L.H5Tset_size(t, H5T_VARIABLE)
                buf = (C.c_char_p * 1)()
                _check(L.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(buf, C.c_void_p)), f"read {name}")
                out = buf[0].decode()
                s = L.H5Dget_space(d)
                L.H5Dvlen_reclaim(t, s, H5P_DEFAULT, C.cast(buf, C.c_void_p))
                L.H5Sclose(s)
            else:
                n = L.H5Tget_size(ft)
                t = L.H5Tcopy(L.T_C_S1)
                L.H5Tset_size(t, n)
                buf = C.create_string_buffer(n + 1)
                _check(L.H5Dread(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, C.cast(buf, C.c_void_p)), f"read {name}")
                out = buf.value.decode()
            L.H5Tclose(t)
            return out
        finally:
            L.H5Tclose(ft)
            L.H5Dclose(d)

    def read_array(self, name: str, dtype) -> np.ndarray:
        L = self.L
        d = _check(L.H5Dopen2(self.id, name.encode(), H5P_DEFAULT), f"open {name}")
        s = L.H5Dget_space(d)
        n = L.H5Sget_simple_extent_npoints(s)
        a = np.empty(n, dtype=dtype)
        mem = L.T_DOUBLE if a.dtype == np.float64 else L.T_FLOAT
        _check(L.H5Dread(d, mem, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)), f"read {name}")
        L.H5Sclose(s)
        L.H5Dclose(d)
        return a


def _type_name(cls: str, scalar: str, D: int) -> str:
    return f"{cls}_{scalar}_{D}_{D}"


def _dim_of(t) -> int:
    if isinstance(t, DisplacementFieldTransform):
        return t.geometry.dim
    return np.asarray(t.matrix).shape[0]


def write_composite(path: str, transforms: Sequence, scalar: str = "float") -> None:
    """A CompositeTransform file whose components are `transforms` in queue
    order (the first is applied last).  `scalar`: "float" (--float 1) or
    "double"."""
    if scalar not in ("float", "double"):
        raise ValueError("scalar must be 'float' or 'double'")
    if not transforms:
        raise ValueError("a composite needs at least one transform")
    D = _dim_of(transforms[0])
    dt = np.float32 if scalar == "float" else np.float64
    maj, mi, rel = C.c_uint(), C.c_uint(), C.c_uint()
    lib().H5get_libversion(C.byref(maj), C.byref(mi), C.byref(rel))
    with _File(path, write=True) as f:
        f.string("/ITKVersion", "5.4.6")
        f.string("/HDFVersion", f"{maj.value}.{mi.value}.{rel.value}")
        f.string("/OSName", platform.system())
        f.string("/OSVersion", platform.release())
        f.group("/TransformGroup")
        f.group("/TransformGroup/0")
        f.string("/TransformGroup/0/TransformType", _type_name("CompositeTransform", scalar, D))
        for i, t in enumerate(transforms, start=1):
            g = f"/TransformGroup/{i}"
            f.group(g)
            if _dim_of(t) != D:
                raise ValueError("all transforms of a composite have one dimension")
            if isinstance(t, DisplacementFieldTransform):
                geo = t.geometry
                fld = t.field if isinstance(t.field, np.ndarray) else t.field.to_host()
                fld = np.asarray(fld)
                if fld.shape != geo.shape + (D,):
                    raise ValueError(f"field shape {fld.shape} does not match its geometry {geo.shape + (D,)}")
                sp = np.ones(D) if geo.spacing is None else np.asarray(geo.spacing, float)
                org = np.zeros(D) if geo.origin is None else np.asarray(geo.origin, float)
                dirm = np.eye(D) if geo.direction is None else np.asarray(geo.direction, float).reshape(D, D)
                fixed = np.concatenate([np.asarray(geo.size, float), org, sp, dirm.ravel()])
                f.string(g + "/TransformType", _type_name("DisplacementFieldTransform", scalar, D))
                f.array(g + "/TransformFixedParameters", fixed.astype(np.float64), compress=False)
                f.array(g + "/TransformParameters", fld.astype(dt).ravel())
            elif isinstance(t, AffineTransform):
                A, tr, c = t.parameters()
                f.string(g + "/TransformType", _type_name("AffineTransform", scalar, D))
                f.array(g + "/TransformFixedParameters", np.asarray(c, np.float64), compress=False)
                f.array(g + "/TransformParameters", np.concatenate([A.ravel(), tr]).astype(dt), compress=False)
            else:
                raise TypeError(f"cannot write {type(t).__name__}")


_AFFINE_CLASSES = ("AffineTransform", "MatrixOffsetTransformBase", "Euler3DTransform", "Euler2DTransform",
                   "Rigid2DTransform", "Similarity3DTransform", "Similarity2DTransform", "VersorRigid3DTransform",
                   "ScaleVersor3DTransform", "CenteredAffineTransform")


def read_composite(path: str) -> List:
    """The components of a composite (queue order) -- or the single transform
    of a non-composite file -- as AffineTransform / DisplacementFieldTransform.
    Matrix-offset classes whose parameters are a matrix and a translation
    (AffineTransform, MatrixOffsetTransformBase) are read directly; other
    linear parameterisations are refused."""
    out = []
    with _File(path, write=False) as f:
        i = 0
        while f.exists(f"/TransformGroup/{i}"):
            g = f"/TransformGroup/{i}"
            ttype = f.read_string(g + "/TransformType")
            parts = ttype.split("_")
            cls = parts[0]
            if cls == "CompositeTransform":
                if i != 0:
                    raise ValueError("a composite transform can only be the first entry of a file")
                i += 1
                continue
            D = int(parts[-1])
            scalar = np.float32 if "float" in parts[1:-2] else np.float64
            fixed = f.read_array(g + "/TransformFixedParameters", np.float64)
            params = f.read_array(g + "/TransformParameters", scalar)
            if cls == "DisplacementFieldTransform":
                size = tuple(int(round(v)) for v in fixed[:D])
                geo = Geometry(size, list(fixed[2 * D:3 * D]), list(fixed[D:2 * D]),
                               fixed[3 * D:3 * D + D * D].reshape(D, D))
                fld = params.astype(np.float32).reshape(geo.shape + (D,))
                out.append(DisplacementFieldTransform(fld, geo))
            elif cls in ("AffineTransform", "MatrixOffsetTransformBase"):
                A = params[:D * D].astype(np.float64).reshape(D, D)
                out.append(AffineTransform(A, params[D * D:D * D + D].astype(np.float64), fixed[:D].copy()))
            else:
                raise ValueError(f"{path}: transform type {ttype} is not supported")
            i += 1
    if not out:
        raise ValueError(f"{path}: no transforms")
    return out


def write_syn_state(path: str, state: Sequence[np.ndarray], geometry: Geometry, linear: Sequence = (),
                    scalar: str = "float") -> None:
    """--save-state: the composite [linear transforms..., FixedToMiddle
    forward, FixedToMiddle inverse, MovingToMiddle forward, MovingToMiddle
    inverse] (Examples/antsRegistrationTemplateHeader.h:1308-1348); `state`
    is the four fields in that order on `geometry`."""
    if len(state) != 4:
        raise ValueError("a SyN state is four fields")
    fields = [DisplacementFieldTransform(np.asarray(s, np.float32), geometry) for s in state]
    write_composite(path, list(linear) + fields, scalar)


def read_syn_state(path: str) -> Tuple[List, List[np.ndarray], Geometry]:
    """--restore-state (Examples/itkantsRegistrationHelper.hxx:3637-3700):
    the last four components, when all are displacement fields, are the SyN
    state (F2M fwd, F2M inv, M2M fwd, M2M inv); the components before them
    are returned as the other transforms of the state."""
    ts = read_composite(path)
    last = ts[-4:]
    if len(ts) < 4 or not all(isinstance(t, DisplacementFieldTransform) for t in last):
        raise ValueError(f"{path}: the last four transforms are not a SyN state")
    geo = last[0].geometry
    for t in last[1:]:
        if tuple(t.geometry.size) != tuple(geo.size):
            raise ValueError(f"{path}: the SyN state fields do not share one grid")
    return ts[:-4], [np.asarray(t.field, np.float32) for t in last], geo
