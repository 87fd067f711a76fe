"""The correctly rounded float64 power of IdealGas1D's nonlinear equation of state
(pyfds/nonlinear_acoustics.py:82-87, ``((rho0 + rho')/rho0)**gamma``).

* The oracle's ``pow_cr`` is pinned against Python's ``decimal`` (exp(y ln x) at 80 digits,
  rounded to the nearest double) on inputs from the equation of state's range and far
  outside it; a planted mistake (the long double power without the decision near rounding
  midpoints) fails the same comparison.
* The product's power (``pyfds_amd/csrc/fds_crmath.h``, the source the kernels include) is
  compiled for the host with g++ and compared with ``pow_cr`` on 1.7 M inputs, bit for bit,
  and with glibc's pow on the IEEE special cases (C99 Annex F).
* What the choice of power does to the reference itself: the ``ig1d_cnl`` fixtures written
  with numpy's AVX-512 power (SVML) and with glibc's pow differ by 2.7e-12 relative L2; the
  oracle with ``pow_cr`` reproduces the glibc-path fixture bit for bit.
"""
import ctypes
import math
import os
import subprocess
from decimal import Decimal, localcontext

import numpy as np
import pytest

import oracle
import scenes

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, 'golden')


def decimal_pow(x, y):
    out = np.empty(len(x))
    with localcontext() as ctx:
        ctx.prec = 80
        for i, (a, b) in enumerate(zip(x, y)):
            out[i] = float((Decimal(float(b)) * Decimal(float(a)).ln()).exp())
    return out


def sample(rng, n):
    """(x, y) pairs: the equation of state's range (x near 1, gamma-like y) and wide ranges."""
    xs, ys = [], []
    for g in (1.4, 5.0 / 3.0, 520.0 / 312.0, 1.3):
        for s in (1e-1, 1e-3, 1e-6, 1e-9):
            xs.append(1.0 + rng.standard_normal(n) * s)
            ys.append(np.full(n, g))
    xs.append(np.exp(rng.uniform(-700.0, 700.0, n)))
    ys.append(rng.uniform(-1.0, 1.0, n))
    xs.append(np.exp(rng.uniform(-5.0, 5.0, n)))
    ys.append(rng.uniform(-100.0, 100.0, n))
    xs.append(rng.uniform(0.0, 4.0, n) + 1e-300)
    ys.append(rng.uniform(-3.0, 3.0, n))
    return np.concatenate(xs), np.concatenate(ys)


def test_oracle_pow_cr_matches_decimal():
    x, y = sample(np.random.default_rng(11), 250)
    ref = decimal_pow(x, y)
    np.testing.assert_array_equal(oracle.pow_cr(x, y), ref)


def test_planted_mistake_in_oracle_power_is_caught():
    # the long double power alone, without the decision near midpoints, misrounds some inputs
    x, y = sample(np.random.default_rng(12), 250)
    planted = np.power(x.astype(np.longdouble), y.astype(np.longdouble)).astype(np.float64)
    assert np.count_nonzero(planted != decimal_pow(x, y)) > 0


@pytest.fixture(scope='module')
def crm(tmp_path_factory):
    src = os.path.join(HERE, 'crmath_host.cpp')
    lib = str(tmp_path_factory.mktemp('crm') / 'libcrm_host.so')
    subprocess.run(['g++', '-O2', '-std=c++17', '-ffp-contract=off', '-fPIC', '-shared', src,
                    '-o', lib], check=True)
    so = ctypes.CDLL(lib)
    dp = ctypes.POINTER(ctypes.c_double)

    def pw(x, y):
        x = np.ascontiguousarray(x, np.float64)
        y = np.ascontiguousarray(np.broadcast_to(y, x.shape), np.float64)
        out = np.empty_like(x)
        so.crm_pow(x.ctypes.data_as(dp), y.ctypes.data_as(dp), out.ctypes.data_as(dp),
                   ctypes.c_long(x.size))
        return out

    def ex(z):
        z = np.ascontiguousarray(z, np.float64)
        out = np.empty_like(z)
        so.crm_exp(z.ctypes.data_as(dp), out.ctypes.data_as(dp), ctypes.c_long(z.size))
        return out
    pw.exp = ex
    return pw


def test_product_power_is_correctly_rounded(crm):
    x, y = sample(np.random.default_rng(13), 60000)
    got, ref = crm(x, y), oracle.pow_cr(x, y)
    bad = np.flatnonzero(got != ref)
    assert bad.size == 0, [(x[i], y[i], got[i], ref[i]) for i in bad[:5]]


def test_product_power_agrees_with_glibc_where_glibc_is_exact(crm):
    # glibc's pow is within 0.52 ulp: it may differ from the correctly rounded value only
    # where that is ambiguous at glibc's accuracy, and then by one ulp
    x, y = sample(np.random.default_rng(14), 5000)
    got = crm(x, y)
    lib = np.array([math.pow(a, b) for a, b in zip(x, y)])
    diff = np.flatnonzero(got != lib)
    assert diff.size <= x.size // 200
    assert np.all(np.abs(got[diff] - lib[diff]) <= np.spacing(np.abs(got[diff])))


SPECIAL = [0.0, -0.0, 1.0, -1.0, 2.0, -2.0, 0.5, -0.5, 3.0, -3.0, 1e-310, -1e-310, 1e300,
           float('inf'), float('-inf'), float('nan')]


def test_product_power_special_cases(crm):
    libm = ctypes.CDLL('libm.so.6')
    libm.pow.restype = ctypes.c_double
    libm.pow.argtypes = [ctypes.c_double, ctypes.c_double]
    x = np.array([a for a in SPECIAL for b in SPECIAL])
    y = np.array([b for a in SPECIAL for b in SPECIAL])
    got = crm(x, y)
    for a, b, g in zip(x, y, got):
        want = libm.pow(a, b)
        if math.isnan(want):
            assert math.isnan(g), (a, b, g)
        else:
            assert g == want and math.copysign(1, g) == math.copysign(1, want), (a, b, g, want)


def test_reference_power_paths_and_cr_oracle():
    scene = scenes.SCENES['ig1d_cnl']()
    svml = np.load(os.path.join(GOLDEN, 'ig1d_cnl.npz'))
    libm = np.load(os.path.join(GOLDEN, 'ig1d_cnl_libm.npz'))
    cr = scenes.run_oracle(scene, power='cr')
    for k in cr:
        np.testing.assert_array_equal(cr[k], libm[k], err_msg=k)
    spread = max(np.linalg.norm(svml[k] - libm[k]) / np.linalg.norm(libm[k])
                 for k in cr if k != 'step')
    assert 1e-12 < spread < 5e-12, spread


def test_product_exp_is_correctly_rounded(crm):
    # the exponential material coupling's exp (b * q with b ~ -0.01 and q a temperature) and a
    # wide range; the oracle's pow_cr(e, z) is e**z correctly rounded only up to the rounding
    # of e itself, so decimal is the reference here
    rng = np.random.default_rng(15)
    z = np.concatenate([rng.uniform(-1.0, 1.0, 3000), rng.uniform(-700.0, 700.0, 3000),
                        -0.01 * rng.uniform(0.0, 60.0, 3000)])
    with localcontext() as ctx:
        ctx.prec = 80
        ref = np.array([float(Decimal(float(v)).exp()) for v in z])
    np.testing.assert_array_equal(crm.exp(z), ref)
    special = np.array([0.0, -0.0, float('inf'), float('-inf'), 710.0, -746.0])
    np.testing.assert_array_equal(crm.exp(special), np.exp(special))
    assert np.isnan(crm.exp(np.array([float('nan')])))[0]
