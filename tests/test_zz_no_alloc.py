"""The no-allocation entry point fds_window_sq_diff_static (partial sums in the library's
static device scratch) against fds_window_sq_diff (which allocates its partials per call):
the same numbers bit for bit, on two calls in a row, the second on a smaller window, and both
against numpy on the downloaded states."""
import ctypes as ct

import numpy as np
import pytest

import pyfds_amd
import scenes
from pyfds_amd import _native as nat

pytestmark = pytest.mark.gpu


def _field(name, steps):
    scene = scenes.SCENES[name]()
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    fld.simulate(steps)
    return fld


def _call(fn, a, b, window):
    num, den = ct.c_double(), ct.c_double()
    wx, wy, ax0, ay0, bx0, by0 = window
    rc = fn(a._runner().ctx, 0, b._runner().ctx, 0, wx, wy, ax0, ay0, bx0, by0,
            ct.byref(num), ct.byref(den))
    nat.check(rc)
    return num.value, den.value


def test_static_scratch_window_reduction_matches_allocating_one():
    lib = nat.load()
    fn_static = lib.fds_window_sq_diff_static
    fn_static.restype = ct.c_int
    fn_static.argtypes = lib.fds_window_sq_diff.argtypes
    a, b = _field('a2d_main', 60), _field('a2d_main', 45)
    nx = a.x.samples
    pa = np.asarray(a.pressure.values).reshape(-1, nx)
    pb = np.asarray(b.pressure.values).reshape(-1, nx)
    for window in [(40, 30, 3, 4, 5, 2), (9, 7, 20, 11, 1, 30)]:     # the second one smaller
        want = _call(lib.fds_window_sq_diff, a, b, window)
        got = _call(fn_static, a, b, window)
        assert got == want, (window, got, want)
        wx, wy, ax0, ay0, bx0, by0 = window
        wa, wb = pa[ay0:ay0 + wy, ax0:ax0 + wx], pb[by0:by0 + wy, bx0:bx0 + wx]
        np.testing.assert_allclose(got, (np.sum((wa - wb) ** 2), np.sum(wb ** 2)), rtol=1e-12)
        assert got[0] > 0 and got[1] > 0


def test_static_scratch_rejects_windows_outside_the_grid():
    lib = nat.load()
    fn_static = lib.fds_window_sq_diff_static
    fn_static.restype = ct.c_int
    fn_static.argtypes = lib.fds_window_sq_diff.argtypes
    a = _field('a2d_main', 5)
    with pytest.raises(IndexError):
        _call(fn_static, a, a, (a.x.samples + 1, 2, 0, 0, 0, 0))
