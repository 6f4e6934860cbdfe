"""CPU checks of the C ABI: the library loads, exports every entry point
include/gwl.h declares, and its host-side routines (spline coefficient
solve and evaluation) match scipy's RectBivariateSpline (the reference's
BoundedRectBivariateSpline, bilby/core/utils/calculus.py:221-262)."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "gwl.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(gwl_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from bilby_amd import _lib
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared_functions()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name
    assert set(declared) == set(_lib.EXPORTED_SYMBOLS)


def test_spline_coefficients_match_scipy():
    from scipy.interpolate import RectBivariateSpline
    from bilby_amd._lib import LIB, check, dptr
    rng = np.random.default_rng(0)
    for x in (np.logspace(-5, 10, 80), np.hstack((-np.logspace(3, -3, 40), np.logspace(-3, 10, 40)))):
        y = np.logspace(-5, 10, 40)
        X, Y = np.meshgrid(x, y, indexing="ij")
        z = np.ascontiguousarray(np.tanh(X / 50) * 30 - 0.01 * np.sqrt(Y) + 0.1 * np.sin(np.arange(X.size).reshape(X.shape)))
        tx, ty, c = np.zeros(len(x) + 4), np.zeros(len(y) + 4), np.zeros_like(z)
        check(LIB.gwl_spline2d_coefficients(len(x), len(y), dptr(x), dptr(y), dptr(z), dptr(tx),
                                            dptr(ty), dptr(c)))
        sp = RectBivariateSpline(x, y, z, kx=3, ky=3, s=0)
        kx, ky = sp.get_knots()
        assert np.array_equal(kx, tx) and np.array_equal(ky, ty)
        ref_c = sp.get_coeffs().reshape(len(x), len(y))
        assert np.max(np.abs(ref_c - c)) < 1e-10 * np.max(np.abs(ref_c))
        m = 500
        xs = rng.uniform(x.min(), x.max(), m)
        ys = 10 ** rng.uniform(-5, 10, m)
        out = np.zeros(m)
        check(LIB.gwl_spline2d_eval_host(len(x), len(y), dptr(tx), dptr(ty), dptr(c), m, dptr(xs),
                                         dptr(ys), dptr(out)))
        ref = sp(xs, ys, grid=False)
        assert np.max(np.abs(out - ref)) < 1e-9 * max(1.0, np.max(np.abs(ref)))


def test_errors_are_codes_not_exceptions():
    from bilby_amd._lib import LIB, dptr
    x = np.array([0.0, 1.0, 1.0, 2.0, 3.0])  # not strictly increasing
    y = np.arange(5.0)
    z = np.zeros(25)
    t = np.zeros(9)
    rc = LIB.gwl_spline2d_coefficients(5, 5, dptr(x), dptr(y), dptr(z), dptr(t), dptr(t), dptr(z))
    assert rc == -1
    assert b"increasing" in LIB.gwl_last_error()


def test_new_entry_points_reject_null_handles_without_a_device():
    """The reconstruction and series entry points validate their arguments
    before any device call: a null handle or output is GWL_E_INVALID with a
    message naming the entry point."""
    from bilby_amd._lib import LIB, NPARAM
    calls = {
        "gwl_roq_inner_products": lambda: LIB.gwl_roq_inner_products(None, None, 1, NPARAM, 0, 0,
                                                                     None, None, None),
        "gwl_relbin_time_series": lambda: LIB.gwl_relbin_time_series(None, None, 1, NPARAM, 0,
                                                                     None, None, None),
        "gwl_mb_time_series": lambda: LIB.gwl_mb_time_series(None, None, 1, NPARAM, 0, None, None,
                                                             None),
        "gwl_relbin_set_upsampled": lambda: LIB.gwl_relbin_set_upsampled(None, None, None, None,
                                                                         1 << 16),
        "gwl_relbin_upsampled_series": lambda: LIB.gwl_relbin_upsampled_series(
            None, None, 1, NPARAM, 0, 8, None, None, None),
        "gwl_weighted_inner_products": lambda: LIB.gwl_weighted_inner_products(None, None, None,
                                                                               None),
        "gwl_window_time_series": lambda: LIB.gwl_window_time_series(None, None, 1 << 16, 0, 8,
                                                                     None, None),
        "gwl_mb_set_fft": lambda: LIB.gwl_mb_set_fft(None, 0, None),
    }
    for name, call in calls.items():
        assert call() == -1, name
        msg = LIB.gwl_last_error()
        assert name.encode() in msg or b"invalid" in msg, (name, msg)
