"""ants_op_invert_noalloc (fixed per-device scratch, no device allocation)
against ants_op_invert on the same inputs: the same kernels, so the same
bits.  Two calls in a row, the second on a smaller grid, so that a stale
step slot, partial or mask word left by the first call would show."""
import ctypes as C

import numpy as np
import pytest
from scipy import ndimage

from ants_amd import _lib as L

rng = np.random.default_rng(11)


def smooth_field(shape, amp):
    D = len(shape)
    f = np.stack([ndimage.gaussian_filter(rng.standard_normal(shape), 3.0) for _ in range(D)], -1)
    return np.ascontiguousarray(amp * f / np.abs(f).max(), np.float32)


def run(fn, fld, init):
    g = L.desc(tuple(reversed(fld.shape[:-1])))
    a = L.DeviceArray.from_host(fld)
    b = L.DeviceArray.from_host(init)
    out = L.DeviceArray(fld.shape)
    k, mx, mn = C.c_int(), C.c_double(), C.c_double()
    L.check(fn(C.byref(g), a.ptr, b.ptr, out.ptr, C.byref(k), C.byref(mx), C.byref(mn), None))
    return out.to_host(), k.value, mx.value, mn.value


@pytest.mark.gpu
def test_invert_noalloc_matches_invert_twice():
    lib = L.lib()
    # first call: a grid above the speculation threshold (2^21 voxels), so
    # the candidate masks in the scratch are used; second call: smaller
    cases = [((130, 128, 128), 3.0), ((28, 36, 40), 2.5)]
    for n, (shape, amp) in enumerate(cases):
        print(f"noalloc inversion call {n}: grid {shape}", flush=True)
        f = smooth_field(shape, amp)
        init = smooth_field(shape, 0.3 * amp)
        ref = run(lib.ants_op_invert, f, init)
        got = run(lib.ants_op_invert_noalloc, f, init)
        assert got[1] == ref[1] and ref[1] > 1
        np.testing.assert_array_equal(got[0], ref[0])
        assert got[2] == ref[2] and got[3] == ref[3]


@pytest.mark.gpu
def test_invert_noalloc_rejects_oversized_grid():
    lib = L.lib()
    g = L.desc((512, 512, 512))
    small = L.DeviceArray((4,))
    k, mx, mn = C.c_int(), C.c_double(), C.c_double()
    rc = lib.ants_op_invert_noalloc(C.byref(g), small.ptr, None, small.ptr, C.byref(k), C.byref(mx), C.byref(mn),
                                    None)
    assert rc == 1
    assert b"scratch" in lib.ants_last_error()
