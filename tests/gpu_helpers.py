"""Thin wrappers that run one library building block on device buffers."""
import ctypes as C

import numpy as np

from ants_amd import _lib as L


def dev(a, dtype=np.float32):
    return L.DeviceArray.from_host(np.ascontiguousarray(a, dtype=dtype))


def desc_of(shape_np, spacing=None, origin=None):
    return L.desc(tuple(reversed(shape_np)), spacing, origin)


def smooth_image(img, sigma, in_mm=False, spacing=None):
    d_in = dev(img)
    d_out = L.DeviceArray(img.shape)
    g = desc_of(img.shape, spacing)
    L.check(L.lib().ants_op_smooth_image(C.byref(g), d_in.ptr, d_out.ptr, sigma, int(in_mm), None))
    return d_out.to_host()


def smooth_field(fld, variance, spacing=None):
    d_in = dev(fld)
    d_out = L.DeviceArray(fld.shape)
    g = desc_of(fld.shape[:-1], spacing)
    L.check(L.lib().ants_op_smooth_field(C.byref(g), d_in.ptr, d_out.ptr, variance, None))
    return d_out.to_host()


def bspline_field(fld, mesh, order=3, spacing=None):
    """BSplineSyN's B-spline fit of a field; returns (fit, max |u/spacing|)."""
    d_in = dev(fld)
    d_out = L.DeviceArray(fld.shape)
    g = desc_of(fld.shape[:-1], spacing)
    m = (C.c_int * 3)(*[int(v) for v in list(mesh) + [1] * (3 - len(mesh))])
    mx = C.c_double()
    L.check(L.lib().ants_op_bspline_field(C.byref(g), d_in.ptr, d_out.ptr, m, int(order), C.byref(mx), None))
    return d_out.to_host(), mx.value


def warp_image(img, out_shape, u=None, gi=None, go=None):
    gi = gi or desc_of(img.shape)
    go = go or desc_of(out_shape)
    d_img = dev(img)
    d_u = dev(u) if u is not None else None
    d_out = L.DeviceArray(out_shape)
    L.check(L.lib().ants_op_warp_image(C.byref(gi), d_img.ptr, C.byref(go), d_u.ptr if d_u else None, d_out.ptr,
                                       None))
    return d_out.to_host()


def cc(F, M, radius, spacing=None):
    D = F.ndim
    g = desc_of(F.shape, spacing)
    dF, dM = dev(F), dev(M)
    g1 = L.DeviceArray(F.shape + (D,))
    g2 = L.DeviceArray(F.shape + (D,))
    v = C.c_double()
    L.check(L.lib().ants_op_cc(C.byref(g), dF.ptr, dM.ptr, radius, g1.ptr, g2.ptr, C.byref(v), None))
    return v.value, g1.to_host(), g2.to_host()


def compose(disp, warp):
    g = desc_of(disp.shape[:-1])
    a, b = dev(disp), dev(warp)
    out = L.DeviceArray(disp.shape)
    L.check(L.lib().ants_op_compose(C.byref(g), a.ptr, b.ptr, out.ptr, None))
    return out.to_host()


def scale_update(u, rate, spacing=None):
    g = desc_of(u.shape[:-1], spacing)
    a = dev(u)
    mx = C.c_double()
    L.check(L.lib().ants_op_scale_update(C.byref(g), a.ptr, rate, C.byref(mx), None))
    return a.to_host(), mx.value


def invert(fld, init=None, spacing=None):
    g = desc_of(fld.shape[:-1], spacing)
    a = dev(fld)
    b = dev(init) if init is not None else None
    out = L.DeviceArray(fld.shape)
    k = C.c_int()
    mx = C.c_double()
    mn = C.c_double()
    L.check(L.lib().ants_op_invert(C.byref(g), a.ptr, b.ptr if b else None, out.ptr, C.byref(k), C.byref(mx),
                                   C.byref(mn), None))
    return out.to_host(), k.value, mx.value, mn.value


def resample_field(fld, gi, go, out_shape):
    a = dev(fld)
    out = L.DeviceArray(out_shape)
    L.check(L.lib().ants_op_resample_field(C.byref(gi), a.ptr, C.byref(go), out.ptr, None))
    return out.to_host()


def preprocess(img):
    a = dev(img)
    out = L.DeviceArray(img.shape)
    L.check(L.lib().ants_op_preprocess(img.size, a.ptr, out.ptr, None))
    return out.to_host()


def labels_nn(u, labels, go, gl, out_shape):
    du = dev(u) if u is not None else None
    dl = dev(labels, np.int32)
    out = L.DeviceArray(out_shape, np.int32)
    L.check(L.lib().ants_apply_transform_labels_nn(C.byref(go), du.ptr if du else None, C.byref(gl), dl.ptr,
                                                   out.ptr, None))
    return out.to_host()


def warp_pair(gl, gF, F, uF, gM, M, uM, out_shape, bf16=False):
    dF, dM, duF, duM = dev(F), dev(M), dev(uF), dev(uM)
    oF, oM = L.DeviceArray(out_shape), L.DeviceArray(out_shape)
    L.check(L.lib().ants_op_warp_pair(C.byref(gl), C.byref(gF), dF.ptr, duF.ptr, oF.ptr, C.byref(gM), dM.ptr, duM.ptr,
                                      oM.ptr, int(bf16), None))
    return oF.to_host(), oM.to_host()


def round_bf16(a):
    """float32 -> nearest BF16 (ties to even), returned as float32."""
    b = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    r = ((b + 0x7FFF + ((b >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def inverse_consistency(fwd, inv, margin, spacing=None):
    g = desc_of(fwd.shape[:-1], spacing)
    a, b = dev(fwd), dev(inv)
    mx, mn, n = C.c_double(), C.c_double(), C.c_int64()
    L.check(L.lib().ants_op_inverse_consistency(C.byref(g), a.ptr, b.ptr, margin, C.byref(mx), C.byref(mn),
                                                C.byref(n), None))
    return mx.value, mn.value, n.value
