"""ctypes binding of ``libants_syn.so`` (the gfx950 C-ABI in include/ants_syn.h).

The library is required: importing the bindings on a machine without the
built library raises, and there is no CPU fallback anywhere in the package.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libants_syn.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "ants_syn.h")

ANTS_OK = 0
ANTS_ERR_CANCELLED = 5
ERRORS = {1: "INVALID", 2: "HIP", 3: "OOM", 4: "UNSUPPORTED", 5: "CANCELLED"}
MAX_LEVELS = 8
MODEL_SYN = 0
MODEL_BSPLINE_SYN = 1
KC_NAMES = ["warp", "cc", "smooth", "compose", "inv"]


class AntsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"ants_syn error {ERRORS.get(code, code)}: {msg}")
        self.code = code


class ImageDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("size", C.c_int64 * 3), ("spacing", C.c_double * 3),
                ("origin", C.c_double * 3), ("direction", C.c_double * 9)]


class SynParamsC(C.Structure):
    _fields_ = [
        ("dim", C.c_int),
        ("n_levels", C.c_int),
        ("shrink", C.c_int * MAX_LEVELS),
        ("sigma", C.c_double * MAX_LEVELS),
        ("sigma_in_mm", C.c_int),
        ("iterations", C.c_int * MAX_LEVELS),
        ("convergence_threshold", C.c_double),
        ("convergence_window", C.c_int),
        ("learning_rate", C.c_double),
        ("update_variance", C.c_double),
        ("total_variance", C.c_double),
        ("cc_radius", C.c_int),
        ("precision", C.c_int),
        ("preprocess", C.c_int),
        ("force_iterations", C.c_int),
        ("model", C.c_int),
        ("update_mesh", C.c_int * 3),
        ("total_mesh", C.c_int * 3),
        ("spline_order", C.c_int),
    ]


class SynStatsC(C.Structure):
    _fields_ = [
        ("levels_run", C.c_int),
        ("iterations", C.c_int * MAX_LEVELS),
        ("level_voxels", C.c_int64 * MAX_LEVELS),
        ("inv_steps", C.c_int64 * MAX_LEVELS),
        ("voxel_iterations", C.c_double),
        ("algorithmic_bytes", C.c_double),
        ("graph_launches", C.c_int64),
        ("eager_iterations", C.c_int64),
        ("inv_tails", C.c_int64),
    ]


class KernelStatsC(C.Structure):
    _fields_ = [("ms", C.c_double * 5), ("launches", C.c_int64 * 5), ("bytes", C.c_double * 5)]


class TransformC(C.Structure):
    _fields_ = [("kind", C.c_int), ("matrix", C.c_double * 9), ("translation", C.c_double * 3),
                ("center", C.c_double * 3), ("field", ImageDesc), ("d_field", C.c_void_p)]


class IterRecordC(C.Structure):
    _fields_ = [
        ("level", C.c_int),
        ("iteration", C.c_int),
        ("energy", C.c_double),
        ("convergence", C.c_double),
        ("inv_iters", C.c_int * 4),
        ("full_scale_cc", C.c_double),
    ]


ITER_CB = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_int))

_lib = None


def header_functions():
    """Names of every function the public header declares."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ants_[a-z0-9_]+)\s*\(", text)) - {"ants_syn_iteration_cb"})


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m ants_amd.build` "
                           "(the package has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, sz = C.c_void_p, C.c_int, C.c_double, C.c_size_t
    pdesc = C.POINTER(ImageDesc)
    L.ants_last_error.restype = C.c_char_p
    L.ants_version.restype = i
    L.ants_device_malloc.argtypes = [C.POINTER(vp), sz]
    L.ants_device_free.argtypes = [vp]
    L.ants_memcpy_h2d.argtypes = [vp, vp, sz, vp]
    L.ants_memcpy_d2h.argtypes = [vp, vp, sz, vp]
    L.ants_memset_device.argtypes = [vp, i, sz, vp]
    L.ants_copy_calibration.argtypes = [vp, vp, sz, vp]
    L.ants_device_memory_info.argtypes = [C.POINTER(sz), C.POINTER(sz)]
    L.ants_stream_create.argtypes = [C.POINTER(vp)]
    L.ants_stream_destroy.argtypes = [vp]
    L.ants_stream_synchronize.argtypes = [vp]
    L.ants_event_create.argtypes = [C.POINTER(vp)]
    L.ants_event_destroy.argtypes = [vp]
    L.ants_event_record.argtypes = [vp, vp]
    L.ants_event_elapsed_ms.argtypes = [C.POINTER(C.c_float), vp, vp]
    L.ants_syn_level_geometry.argtypes = [C.POINTER(SynParamsC), pdesc, i, pdesc]
    L.ants_syn_create.argtypes = [C.POINTER(SynParamsC), C.POINTER(vp)]
    L.ants_syn_destroy.argtypes = [vp]
    L.ants_syn_destroy.restype = None
    L.ants_syn_set_stream.argtypes = [vp, vp]
    L.ants_syn_register.argtypes = [vp, pdesc, vp, pdesc, vp, vp, vp, ITER_CB, vp]
    L.ants_syn_register_device.argtypes = [vp, pdesc, vp, pdesc, vp, vp, vp, ITER_CB, vp]
    L.ants_syn_get_state_device.argtypes = [vp, vp, vp, vp, vp]
    L.ants_syn_get_stats.argtypes = [vp, C.POINTER(SynStatsC)]
    L.ants_syn_convergence_value.argtypes = [C.POINTER(d), i, i, C.POINTER(d)]
    L.ants_syn_get_trace.argtypes = [vp, C.POINTER(IterRecordC), C.c_int64, C.POINTER(C.c_int64)]
    L.ants_syn_set_profiling.argtypes = [vp, C.c_uint]
    L.ants_syn_set_profiling_stride.argtypes = [vp, i]
    L.ants_syn_set_full_scale_cc_interval.argtypes = [vp, i]
    L.ants_syn_last_full_scale_cc.argtypes = [vp, C.POINTER(d)]
    L.ants_syn_get_kernel_stats.argtypes = [vp, C.POINTER(KernelStatsC)]
    L.ants_syn_get_kernel_stats_level.argtypes = [vp, i, C.POINTER(KernelStatsC)]
    L.ants_op_smooth_image.argtypes = [pdesc, vp, vp, d, i, vp]
    L.ants_op_smooth_field.argtypes = [pdesc, vp, vp, d, vp]
    L.ants_op_bspline_field.argtypes = [pdesc, vp, vp, C.POINTER(i), i, C.POINTER(d), vp]
    L.ants_image_center_of_mass.argtypes = [pdesc, vp, C.POINTER(d), C.POINTER(d), vp]
    L.ants_op_warp_image.argtypes = [pdesc, vp, pdesc, vp, vp, vp]
    L.ants_op_cc.argtypes = [pdesc, vp, vp, i, vp, vp, C.POINTER(d), vp]
    L.ants_op_warp_pair.argtypes = [pdesc, pdesc, vp, vp, vp, pdesc, vp, vp, vp, i, vp]
    L.ants_op_compose.argtypes = [pdesc, vp, vp, vp, vp]
    L.ants_op_inverse_consistency.argtypes = [pdesc, vp, vp, i, C.POINTER(d), C.POINTER(d), C.POINTER(C.c_int64), vp]
    L.ants_op_scale_update.argtypes = [pdesc, vp, d, C.POINTER(d), vp]
    L.ants_op_invert.argtypes = [pdesc, vp, vp, vp, C.POINTER(i), C.POINTER(d), C.POINTER(d), vp]
    L.ants_op_invert_noalloc.argtypes = [pdesc, vp, vp, vp, C.POINTER(i), C.POINTER(d), C.POINTER(d), vp]
    L.ants_op_resample_field.argtypes = [pdesc, vp, pdesc, vp, vp]
    L.ants_op_preprocess.argtypes = [C.c_int64, vp, vp, vp]
    L.ants_apply_transform_labels_nn.argtypes = [pdesc, vp, pdesc, vp, vp, vp]
    L.ants_apply_transforms.argtypes = [pdesc, pdesc, vp, C.POINTER(TransformC), i, i, d, vp, vp]
    L.ants_compose_transforms_to_field.argtypes = [pdesc, C.POINTER(TransformC), i, vp, vp]
    L.ants_apply_transforms_to_points.argtypes = [i, vp, C.c_int64, C.POINTER(TransformC), i, vp, vp]
    L.ants_jacobian_determinant.argtypes = [pdesc, vp, i, i, vp, vp, vp]
    L.ants_syn_set_initial_transforms.argtypes = [vp, C.POINTER(TransformC), i, C.POINTER(TransformC), i]
    L.ants_syn_set_initial_state_device.argtypes = [vp, pdesc, vp, vp, vp, vp]
    L.ants_tuning_set.argtypes = [i, i, i]
    L.ants_tuning_get.argtypes = [i, i, C.POINTER(i)]
    L.ants_tuning_load.argtypes = [C.c_char_p]
    _lib = L
    return L


def check(rc: int):
    if rc != ANTS_OK:
        raise AntsError(rc, lib().ants_last_error().decode())


def desc(size, spacing=None, origin=None, direction=None) -> ImageDesc:
    """size in (x, y[, z]) order; direction row-major dim x dim (identity by default)."""
    D = len(size)
    g = ImageDesc()
    g.dim = D
    for k in range(3):
        g.size[k] = int(size[k]) if k < D else 1
        g.spacing[k] = float(spacing[k]) if spacing is not None and k < D else 1.0
        g.origin[k] = float(origin[k]) if origin is not None and k < D else 0.0
    dm = np.eye(D) if direction is None else np.asarray(direction, dtype=np.float64).reshape(D, D)
    for k, v in enumerate(dm.ravel()):
        g.direction[k] = float(v)
    return g


class DeviceArray:
    """A device buffer owned by the library's allocator, shaped like a numpy array."""

    def __init__(self, shape, dtype=np.float32):
        self.shape = tuple(int(s) for s in shape)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().ants_device_malloc(C.byref(p), max(self.nbytes, 4)))
        self.ptr = p

    @classmethod
    def from_host(cls, a: np.ndarray, dtype=None):
        a = np.ascontiguousarray(a, dtype=dtype or a.dtype)
        d = cls(a.shape, a.dtype)
        check(lib().ants_memcpy_h2d(d.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, None))
        return d

    def copy_from_host(self, a: np.ndarray):
        """Overwrite the buffer with a host array of the same size."""
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError(f"host array of {a.nbytes} bytes for a {self.nbytes}-byte buffer")
        check(lib().ants_memcpy_h2d(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, None))
        check(lib().ants_stream_synchronize(None))
        return self

    def to_host(self) -> np.ndarray:
        out = np.empty(self.shape, self.dtype)
        check(lib().ants_memcpy_d2h(out.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes, None))
        return out

    def zero(self):
        check(lib().ants_memset_device(self.ptr, 0, self.nbytes, None))
        check(lib().ants_stream_synchronize(None))
        return self

    def free(self):
        if self.ptr:
            lib().ants_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        check(lib().ants_event_create(C.byref(p)))
        self.ptr = p

    def record(self, stream=None):
        check(lib().ants_event_record(self.ptr, stream))

    def elapsed_ms(self, stop: "Event") -> float:
        ms = C.c_float()
        check(lib().ants_event_elapsed_ms(C.byref(ms), self.ptr, stop.ptr))
        return ms.value

    def __del__(self):
        try:
            lib().ants_event_destroy(self.ptr)
        except Exception:
            pass


def device_free_bytes() -> int:
    """Free device memory (hipMemGetInfo) of the current device."""
    f = C.c_size_t()
    check(lib().ants_device_memory_info(C.byref(f), None))
    return int(f.value)
