"""Pins the CPU oracle (oracle/plain.py) before anything is compared with it.

* golden fixtures written by running the reference itself (bit-exact);
* the reference's own known-answer operator matrices
  (test/test_fields.py:57-107, test/test_acoustics.py:19-66,
  test/test_coupling.py:47-84), restated here as data;
* scipy's ``dia_matrix`` product and sum (the library routine the reference
  multiplies with);
* closed forms: telescoping sums of column-indexed differences, the 1D
  magic time step, linearity, dense brute force;
* planted mistakes in a copy of the oracle must fail these checks.
"""
import os

import numpy as np
import pytest
import scipy.sparse as sp

import oracle
import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


# ------------------------------------------------------------------ golden fixtures
@pytest.mark.parametrize('name', sorted(scenes.SCENES))
def test_oracle_matches_reference_golden(name):
    scene = scenes.SCENES[name]()
    # IdealGas1D's nonlinear state: the fixture written with this host's numpy power
    suffix = '_libm' if name in scenes.POW_SCENES and oracle.numpy_power_is_libm() else ''
    gold = np.load(os.path.join(GOLDEN, name + suffix + '.npz'))
    assert str(gold['fingerprint']) == scenes.fingerprint(scene), 'scene changed: regenerate'
    res = scenes.run_oracle(scene)
    keys = [k for k in gold.files if k != 'fingerprint']
    assert sorted(keys) == sorted(res)
    for k in keys:
        np.testing.assert_array_equal(res[k], gold[k], err_msg='%s %s' % (name, k))


# ------------------------------------------------------------------ reference KATs
def _mats(params, n):
    return [(list(range(n)), params)]


def _acoustic_params(c, rho, **kw):
    p = dict(sound_velocity=c, density=rho, shear_viscosity=0, bulk_viscosity=0,
             thermal_conductivity=0, isobaric_heat_cap=1, isochoric_heat_cap=1)
    p.update(kw)
    p['absorption_coef'] = oracle.absorption_coef(p)
    return p


def test_absorption_coef_kat():
    # test/test_acoustics.py:5-16
    p = _acoustic_params(1500, 1000, bulk_viscosity=1e-3, shear_viscosity=1e-3)
    assert np.isclose(p['absorption_coef'], 7e-3 / 3)
    assert np.isclose(oracle.absorption_coef(dict(absorption_coef=2e-3)), 2e-3)


def test_acoustic1d_matrices_kat():
    # test/test_acoustics.py:19-26
    f = oracle.PlainAcoustic1D(3, 1, 1, 1, _mats(_acoustic_params(700, 0.01, bulk_viscosity=1), 3))
    f.assemble()
    assert np.allclose(f.a_p_v.toarray(), [[-4900, 4900, 0], [0, -4900, 4900], [0, 0, -4900]])
    assert np.allclose(f.a_v_p.toarray(), [[100, 0, 0], [-100, 100, 0], [0, -100, 100]])
    assert np.allclose(f.a_v_v.toarray(), [[-200, 100, 0], [100, -200, 100], [0, 100, -200]])


def test_acoustic2d_matrices_kat():
    # test/test_acoustics.py:29-46 (2x2: pins the x-stencil wrap across rows)
    f = oracle.PlainAcoustic2D(2, 1, 1, 1, _mats(_acoustic_params(700, 0.01, bulk_viscosity=1), 4),
                               ny=2, dy=1)
    f.assemble()
    assert np.allclose(f.a_p_vx.toarray(), [[-4900, 4900, 0, 0], [0, -4900, 4900, 0],
                                            [0, 0, -4900, 4900], [0, 0, 0, -4900]])
    assert np.allclose(f.a_p_vy.toarray(), [[-4900, 0, 4900, 0], [0, -4900, 0, 4900],
                                            [0, 0, -4900, 0], [0, 0, 0, -4900]])
    assert np.allclose(f.a_vx_p.toarray(), [[100, 0, 0, 0], [-100, 100, 0, 0],
                                            [0, -100, 100, 0], [0, 0, -100, 100]])
    assert np.allclose(f.a_vy_p.toarray(), [[100, 0, 0, 0], [0, 100, 0, 0],
                                            [-100, 0, 100, 0], [0, -100, 0, 100]])
    v = [[-400, 100, 100, 0], [100, -400, 100, 100], [100, 100, -400, 100], [0, 100, 100, -400]]
    assert np.allclose(f.a_vx_vx.toarray(), v)
    assert np.allclose(f.a_vy_vy.toarray(), v)


def test_acoustic_axi_matrices_kat():
    # test/test_acoustics.py:49-66 (pins column-indexed 1/r: a_p_vx[0,1] = 2/3)
    f = oracle.PlainAcoustic3DAxi(2, 1, 1, 1, _mats(_acoustic_params(1, 1, bulk_viscosity=1), 4),
                                  ny=2, dy=1)
    f.assemble()
    assert np.allclose(f.a_p_vx.toarray(), [[-2, 2 / 3, 0, 0], [0, -2 / 3, 2, 0],
                                            [0, 0, -2, 2 / 3], [0, 0, 0, -2 / 3]])
    assert np.allclose(f.a_p_vy.toarray(), [[-1, 0, 1, 0], [0, -1, 0, 1],
                                            [0, 0, -1, 0], [0, 0, 0, -1]])
    assert np.allclose(f.a_vx_p.toarray(), [[1, 0, 0, 0], [-1, 1, 0, 0], [0, -1, 1, 0],
                                            [0, 0, -1, 1]])
    assert np.allclose(f.a_vy_p.toarray(), [[1, 0, 0, 0], [0, 1, 0, 0], [-1, 0, 1, 0],
                                            [0, -1, 0, 1]])
    v = [[-4, 4 / 3, 1, 0], [0, -4, 2, 1], [1, 2 / 3, -4, 4 / 3], [0, 1, 0, -4]]
    assert np.allclose(f.a_vx_vx.toarray(), v)
    assert np.allclose(f.a_vy_vy.toarray(), v)


def test_thermal1d_matrix_kat():
    # test/test_coupling.py:113-118 and the density-scaled variants :126, :133
    for rho, scale in ((1, 1), (2, 2), (5, 5)):
        mats = _mats(dict(heat_capacity=1, density=rho, thermal_conductivity_x=1,
                          thermal_conductivity_y=1), 3)
        f = oracle.PlainThermal1D(3, 1, 1, 1, mats)
        f.assemble()
        assert np.allclose(f.a_t_q.toarray(),
                           np.array([[-1, 1, 0], [0, -1, 1], [0, 0, -1]]) / scale)


def test_difference_operator_kats():
    # test/test_fields.py:57-107
    e3 = np.array(1).repeat(3)
    assert np.allclose(oracle.d_first(3, 1).toarray(), [[-1, 1, 0], [0, -1, 1], [0, 0, -1]])
    assert np.allclose(oracle.d_first(3, 1, variant='backward').toarray(),
                       [[1, 0, 0], [-1, 1, 0], [0, -1, 1]])
    assert np.allclose(oracle.d_first(3, 1, variant='central').toarray(),
                       [[0, 0.5, 0], [-0.5, 0, 0.5], [0, -0.5, 0]])
    assert np.allclose(oracle.d_second(3, 1, e3).toarray(), [[-2, 1, 0], [1, -2, 1], [0, 1, -2]])
    assert np.allclose(oracle.d_first(4, 1).toarray(), [[-1, 1, 0, 0], [0, -1, 1, 0],
                                                        [0, 0, -1, 1], [0, 0, 0, -1]])
    assert np.allclose(oracle.d_first(4, 1, variant='central').toarray(),
                       [[0, 0.5, 0, 0], [-0.5, 0, 0.5, 0], [0, -0.5, 0, 0.5], [0, 0, -0.5, 0]])
    assert np.allclose(oracle.d_first(4, 2).toarray(), [[-1, 0, 1, 0], [0, -1, 0, 1],
                                                        [0, 0, -1, 0], [0, 0, 0, -1]])
    assert np.allclose(oracle.d_first(4, 2, variant='backward').toarray(),
                       [[1, 0, 0, 0], [0, 1, 0, 0], [-1, 0, 1, 0], [0, -1, 0, 1]])
    assert np.allclose(oracle.d_first(4, 2, variant='central').toarray(),
                       [[0, 0, 0.5, 0], [0, 0, 0, 0.5], [-0.5, 0, 0, 0], [0, -0.5, 0, 0]])
    assert np.allclose(oracle.d_second(4, 2).toarray(), [[-2, 0, 1, 0], [0, -2, 0, 1],
                                                         [1, 0, -2, 0], [0, 1, 0, -2]])
    with pytest.raises(ValueError):
        oracle.d_first(3, 1, variant='sideways')


def test_material_vector_kat():
    # test/test_fields.py:110-122 style: later regions overwrite, KeyError if missing
    mats = [(list(range(4)), {'real': 5}), ([1, 2], {'real': 23})]
    assert np.allclose(oracle.material_vector(4, mats, 'real'), [5, 23, 23, 5])
    with pytest.raises(KeyError):
        oracle.material_vector(4, mats, 'imag')


def test_boundary_kats():
    # test/test_fields.py:12-38
    v = np.random.default_rng(0).random(100)
    oracle.apply_boundaries(v, [dict(indices=[5, 6, 7], value=23, additive=False)], 0)
    assert np.allclose(v[[5, 6, 7]], 23)
    v = np.ones(100)
    sig = [np.arange(0, 3) * 23, np.arange(0, 3) * 42, np.arange(0, 3) * 23]
    oracle.apply_boundaries(v, [dict(indices=[5, 6, 7], value=sig, additive=True)], 1)
    assert np.allclose(v[[5, 6, 7]], [24, 43, 24])
    v = np.ones(100)
    oracle.apply_boundaries(v, [dict(indices=[5, 6, 7], value=np.arange(0, 3) * 23,
                                     additive=True)], 2)
    assert np.allclose(v[[5, 6, 7]], [47, 47, 47])


def test_output_kat():
    # test/test_fields.py:41-47, test/test_regions.py:5-8
    out = [dict(indices=[0, 1, 2], signals=[])]
    v = np.zeros(100)
    oracle.record_outputs(v, out)
    oracle.record_outputs(v, out)
    assert np.allclose(out[0]['signals'], [[0, 0], [0, 0], [0, 0]])


# ------------------------------------------------------------------ library routine
@pytest.mark.parametrize('seed', range(6))
def test_dia_matvec_matches_scipy(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(2, 60))
    stride = int(rng.integers(1, n))
    offs = sorted(set(int(o) for o in rng.choice([-stride, -1, 0, 1, stride], 3)))
    data = [rng.standard_normal(n) for _ in offs]
    x = rng.standard_normal(n)
    mine = oracle.Dia(n, list(zip(offs, data)))
    ref = sp.dia_matrix((np.array(data), offs), shape=(n, n))
    np.testing.assert_array_equal(mine.matvec(x), ref.dot(x))
    np.testing.assert_array_equal(mine.toarray(), ref.toarray())


@pytest.mark.parametrize('stride', [1, 2, 5])
def test_dia_sum_matches_scipy(stride):
    rng = np.random.default_rng(stride)
    n = 12
    fx, fy, fc = rng.random(n), rng.random(n), rng.random(n)
    mine = (oracle.d_second(n, 1, fx) + oracle.d_second(n, stride, fy)
            + oracle.d_first(n, 1, fc, 'central'))
    ref = (sp.dia_matrix((np.array([fx, -2 * fx, fx]), [-1, 0, 1]), shape=(n, n))
           + sp.dia_matrix((np.array([fy, -2 * fy, fy]), [-stride, 0, stride]), shape=(n, n))
           + sp.dia_matrix((np.array([-fc / 2, fc / 2]), [-1, 1]), shape=(n, n))).todia()
    assert mine.offsets == list(ref.offsets)
    x = rng.standard_normal(n)
    np.testing.assert_array_equal(mine.matvec(x), ref.dot(x))


# ------------------------------------------------------------------ closed forms
def _thermal2d(nx, ny, seed):
    rng = np.random.default_rng(seed)
    n = nx * ny
    mats = [(list(range(n)), dict(heat_capacity=4000.0, density=1000.0,
                                  thermal_conductivity_x=0.6, thermal_conductivity_y=0.6)),
            (list(rng.choice(n, n // 3, replace=False)),
             dict(heat_capacity=2000.0, density=3000.0, thermal_conductivity_x=50.0,
                  thermal_conductivity_y=20.0))]
    f = oracle.PlainThermal2D(nx, 1e-3, 10, 0.005, mats, ny=ny, dy=1e-3)
    f.components['temperature'] = rng.random(n)
    f.assemble()
    return f


def test_thermal2d_telescoping_sum():
    """Column-indexed forward differences telescope: per step
    sum(T_new - T) = G_x[0] q_x[0] + sum_{j<nx} G_y[j] q_y[j]."""
    f = _thermal2d(9, 7, 0)
    t0 = f.components['temperature'].copy()
    f.sim_step()
    qx, qy = f.components['heat_flux_x'], f.components['heat_flux_y']
    gx = f.a_t_qx.diags[1][1]
    gy = f.a_t_qy.diags[1][1]
    expected = gx[0] * qx[0] + np.sum(gy[:f.nx] * qy[:f.nx])
    got = np.sum(f.components['temperature'] - t0)
    assert np.isclose(got, expected, rtol=1e-9, atol=1e-12)


def test_acoustic1d_magic_step():
    """Uniform lossless medium with c dt = dx: a right-going pulse
    p_i = f(i), v_i = f(i) / (rho c) (v_i sits at x_{i-1/2}, half a step back in
    time) moves exactly one cell per step."""
    n, dx, c, rho = 200, 1e-3, 1500.0, 1000.0
    dt = dx / c
    f = oracle.PlainAcoustic1D(n, dx, 1, dt, [(list(range(n)), dict(sound_velocity=c,
                                                                       density=rho,
                                                                       absorption_coef=0))])
    i = np.arange(n)
    pulse = lambda x: np.exp(-0.5 * ((x - 60) / 5.0) ** 2)
    f.components['pressure'] = pulse(i)
    f.components['velocity'] = pulse(i) / (rho * c)
    f.assemble()
    steps = 40
    for _ in range(steps):
        f.sim_step()
        f.step += 1
    np.testing.assert_allclose(f.components['pressure'][20:180], pulse(i - steps)[20:180],
                               rtol=0, atol=1e-12)
    assert np.argmax(f.components['pressure']) == 60 + steps


def test_acoustic2d_linearity():
    """The run from a sum of initial states equals the sum of the runs."""
    base = scenes.scene_a2d_wrap()
    rng = np.random.default_rng(3)
    n = 7 * 5
    a, b = rng.standard_normal(n), rng.standard_normal(n)
    runs = []
    for p0 in (a, b, a + b):
        s = dict(base, init={'pressure': p0})
        runs.append(scenes.run_oracle(s)['values/pressure'])
    scale = np.linalg.norm(runs[2])
    assert np.linalg.norm(runs[0] + runs[1] - runs[2]) / scale < 1e-13


def test_acoustic2d_dense_brute_force():
    """Iterate the dense block operator built from the textbook staggered
    scheme (x neighbour = flat i+-1, y neighbour = flat i+-nx, zero ghosts only
    at the ends of the flat array) and compare."""
    s = scenes.scene_a2d_wrap()
    nx, ny = 7, 5
    n = nx * ny
    f = scenes.build_oracle(s)
    f.assemble()
    kx, ky, bx, by = f.Kx, f.Ky, f.Bx, f.By
    dxf = np.zeros((n, n))
    dyf = np.zeros((n, n))
    dxb = np.zeros((n, n))
    dyb = np.zeros((n, n))
    for i in range(n):
        dxf[i, i] -= kx[i]
        dyf[i, i] -= ky[i]
        dxb[i, i] += bx[i]
        dyb[i, i] += by[i]
        if i + 1 < n:
            dxf[i, i + 1] += kx[i + 1]
        if i + nx < n:
            dyf[i, i + nx] += ky[i + nx]
        if i - 1 >= 0:
            dxb[i, i - 1] -= bx[i - 1]
        if i - nx >= 0:
            dyb[i, i - nx] -= by[i - nx]
    p = s['init']['pressure'].copy()
    vx = s['init']['velocity_x'].copy()
    vy = s['init']['velocity_y'].copy()
    for _ in range(s['kwargs']['t_samples']):
        vx = vx - dxb @ p
        vy = vy - dyb @ p
        p = p - (dxf @ vx + dyf @ vy)
    res = scenes.run_oracle(s)
    np.testing.assert_allclose(res['values/pressure'], p, rtol=1e-10, atol=1e-12)
    np.testing.assert_allclose(res['values/velocity_x'], vx, rtol=1e-10, atol=1e-16)


# ------------------------------------------------------------------ planted mistakes
def test_planted_mistakes_are_caught(monkeypatch):
    """A row-indexed factor (the docstring's reading of ``d_x``) or a dropped
    row wrap must fail the reference's known-answer matrices."""
    good = oracle.d_first

    def row_indexed(n, stride, factors=None, variant='forward'):
        op = good(n, stride, factors, variant)
        shifted = []
        for off, data in op.diags:
            shifted.append((off, np.roll(data, off)))  # factor taken at the row
        return oracle.Dia(n, shifted)

    monkeypatch.setattr(oracle.plain, 'd_first', row_indexed)
    with pytest.raises(AssertionError):
        test_acoustic_axi_matrices_kat()
    monkeypatch.setattr(oracle.plain, 'd_first', good)

    def no_wrap(n, stride, factors=None, variant='forward'):
        op = good(n, stride, factors, variant)
        if stride != 1:
            return op
        nx = 2
        out = []
        for off, data in op.diags:
            d = data.copy()
            if off == 1:
                d[::nx] = 0   # drop coupling from the last cell of a row to the next row
            if off == -1:
                d[nx - 1::nx] = 0
            out.append((off, d))
        return oracle.Dia(n, out)

    monkeypatch.setattr(oracle.plain, 'd_first', no_wrap)
    with pytest.raises(AssertionError):
        test_acoustic2d_matrices_kat()


def test_scipy_timed_path_equals_numpy_restatement(monkeypatch):
    # the CPU baseline times the oracle with scipy's own dia_matrix products: same bits
    scene = scenes.SCENES['a2d_main']()
    ref = scenes.run_oracle(scene)
    monkeypatch.setattr(oracle.Dia, 'use_scipy', True)
    got = scenes.run_oracle(scene)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
