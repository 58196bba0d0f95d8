"""The convergence monitor (WindowConvergenceMonitoringFunction, described at
Examples/antsRegistration.cxx:397-404) of the library and of the oracle,
against an independent multilevel B-spline approximation.

The reference fits a B-spline of order 1 with 2 control points, one level,
by BSplineScatteredDataPointSetToImageFilter to the last W energies at
t_n = n/(W-1), each divided by the running total of |E| since the level
began, and reports minus the derivative at t = 1.  The approximation is the
MBA rule of Lee, Wolberg and Shin ("Scattered data interpolation with
multilevel B-splines", IEEE TVCG 3(3), 1997): every data point c proposes
phi_kc = w_kc z_c / sum_l w_lc^2 to each control point k it touches, and
phi_k = sum_c w_kc^2 phi_kc / sum_c w_kc^2.  mba() below builds that from
scipy's B-spline basis for any order and control-point count (uniform
unclamped knots, as ITK's parametric domain [0, 1] has) and is then used with
order 1 and 2 control points.  A least-squares line, or normalising by the
window's own |E| total, gives different values on these sequences, so the
check tells them apart."""
import ctypes as C

import numpy as np
import pytest
from scipy.interpolate import BSpline

import oracle as o
from ants_amd import _lib as L

W = 10


def basis(order, ncp, t):
    spans = ncp - order
    knots = (np.arange(ncp + order + 1) - order) / spans
    B = np.empty((len(t), ncp))
    for k in range(ncp):
        c = np.zeros(ncp)
        c[k] = 1.0
        B[:, k] = BSpline(knots, c, order, extrapolate=False)(t)
    return knots, B


def mba(z, order=1, ncp=2):
    """Control points of the one-level MBA fit to z at t_n = n/(len(z)-1)."""
    t = np.arange(len(z)) / (len(z) - 1)
    knots, B = basis(order, ncp, t)
    w2 = (B ** 2).sum(1)
    phi_c = B * (z / w2)[:, None]                   # proposals phi_kc
    num = (B ** 2 * phi_c).sum(0)
    den = (B ** 2).sum(0)
    phi = np.where(den > 0, num / np.where(den > 0, den, 1), 0.0)
    return BSpline(knots, phi, order)


def reference_value(energies, window=W):
    e = np.asarray(energies, float)
    if len(e) < window:
        return np.finfo(float).max
    total = np.abs(e).sum()                         # running total over the level
    spline = mba(e[-window:] / total)
    return -float(spline.derivative()(1.0))


def library_value(energies, window=W):
    e = np.ascontiguousarray(energies, dtype=np.float64)
    out = C.c_double()
    L.check(L.lib().ants_syn_convergence_value(e.ctypes.data_as(C.POINTER(C.c_double)), len(e), window,
                                               C.byref(out)))
    return out.value


rng = np.random.default_rng(5)
SEQUENCES = {
    "linear": -0.5 - 0.01 * np.arange(W),
    "constant": np.full(W, -0.62),
    "alternating": -0.5 + 0.01 * (-1.0) ** np.arange(W),
    "decaying": -0.8 + 0.3 * np.exp(-np.arange(W) / 3.0),
    "noisy": -0.7 - 0.002 * np.arange(W) + 1e-3 * rng.standard_normal(W),
    # 15 values: the running |E| total (all 15) differs from the window's (last 10)
    "running_total": np.concatenate([np.full(5, -0.95), -0.5 - 0.02 * np.arange(W)]),
    "rising": -0.9 + 0.015 * np.arange(12),
}


def test_mba_reduces_to_hat_functions():
    """Order 1 with 2 control points: the basis is (1 - t, t)."""
    t = np.linspace(0, 1, 7)
    _, B = basis(1, 2, t)
    np.testing.assert_allclose(B, np.stack([1 - t, t], 1), atol=1e-15)
    # with two points each touches one control point alone: the fit passes through both
    s = mba(np.array([0.0, 1.0]))
    assert s(0.0) == pytest.approx(0.0) and s(1.0) == pytest.approx(1.0)
    # three points on the line z = t: the middle one proposes 0.5 to both
    # control points with weight 1/4, so phi = (0.125 / 1.25, 1.125 / 1.25)
    # = (0.1, 0.9) by hand -- slope 0.8, where least squares gives 1
    s = mba(np.array([0.0, 0.5, 1.0]))
    np.testing.assert_allclose(s.c[:2], [0.1, 0.9], atol=1e-15)


@pytest.mark.parametrize("name", sorted(SEQUENCES))
def test_library_monitor_matches_mba(name):
    e = SEQUENCES[name]
    ref = reference_value(e)
    got = library_value(e)
    assert got == pytest.approx(ref, rel=1e-12, abs=1e-15)
    assert o.conv_value(e, W) == pytest.approx(ref, rel=1e-12, abs=1e-15)


def test_monitor_before_window_is_max():
    assert library_value(SEQUENCES["linear"][:W - 1]) == np.finfo(float).max
    assert o.conv_value(SEQUENCES["linear"][:W - 1], W) == np.finfo(float).max


def test_alternatives_are_told_apart():
    """A least-squares line or a window-only |E| total would fail the checks above."""
    for name in ("linear", "decaying", "noisy", "running_total"):
        e = SEQUENCES[name]
        ref = reference_value(e)
        z = e[-W:] / np.abs(e).sum()
        t = np.arange(W) / (W - 1)
        ls = -np.polyfit(t, z, 1)[0]
        assert abs(ls - ref) > 1e-3 * abs(ref), name
    e = SEQUENCES["running_total"]
    win_only = -float(mba(e[-W:] / np.abs(e[-W:]).sum()).derivative()(1.0))
    assert abs(win_only - reference_value(e)) > 0.1 * abs(reference_value(e))


def test_monitor_sign_and_scale():
    """Falling energy gives a positive value, rising a negative one; a
    constant window gives 0 (the stop rule is value < threshold)."""
    assert library_value(SEQUENCES["linear"]) > 0
    assert library_value(SEQUENCES["rising"]) < 0
    assert library_value(SEQUENCES["constant"]) == pytest.approx(0.0, abs=1e-15)
