"""The benchmark configurations at their full sizes (Acoustic2D 2048^2, Thermal2D 4096^2):
short runs against the oracle bit for bit on every engine the autotuner picks from, and
size-independent properties over longer runs (superposition; FP32 within the FP64 tolerance)."""
import numpy as np
import pytest

import pyfds_amd
import scenes

pytestmark = pytest.mark.gpu

WATER = dict(sound_velocity=1500.0, density=1000.0)
STEEL = dict(sound_velocity=5900.0, density=7850.0)


def headline_scene(steps, sources=None, seed=1, n=2048):
    """Acoustic2D 2048^2 as the benchmark builds it: a steel inclusion, a 1 MHz burst at the
    centre (or ``sources``: (index, amplitude) pairs), 64 recorded points."""
    dx = 1e-4
    dt = 0.5 * dx / 5900.0
    t = np.arange(steps + 8) * dt
    rng = np.random.default_rng(seed)
    w, h = (int(v) for v in rng.integers(128, 513, size=2))
    x0, y0 = int(rng.integers(0, n - w)), int(rng.integers(0, n - h))
    outs = [int(v) for v in rng.choice(n * n, 64, replace=False)]
    burst = np.sin(2 * np.pi * 1e6 * t) * (t < 5e-6)
    if sources is None:
        sources = [(n // 2 + n * (n // 2), 1.0)]
    return dict(
        name='headline', model='Acoustic2D',
        kwargs=dict(x_samples=n, x_delta=dx, y_samples=n, y_delta=dx, t_samples=steps + 8,
                    t_delta=dt),
        material=('AcousticMaterial', WATER),
        regions=[(scenes.rect_indices(n, x0, y0, w, h), ('AcousticMaterial', STEEL))],
        init={},
        boundaries={'pressure': [dict(indices=[i], value=a * burst, additive=True)
                                 for i, a in sources]},
        outputs={'pressure': [outs]},
        plan=[('simulate', steps)])


def _stream(slot):
    return [i for i, e in enumerate(pyfds_amd._native.tile_engines()) if e == 'stream'][slot]


def _run(scene, form='fp64', tile=None, strip=0):
    f = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    f.form = form
    if tile is not None:
        f.tile, f.strip_height = tile, strip
    scenes.run_plan(f, scene)
    res = scenes.collect(f, scene)
    f._runner().close()
    return res


def lossy_scene(steps, seed=4):
    """The benchmark's lossy configuration (every material viscous, as bench.make_lossy sets
    it) on the headline grid, from random fields so that every cell's viscous terms count."""
    scene = headline_scene(steps)
    visc = dict(shear_viscosity=1e-3, bulk_viscosity=2.4e-3)
    scene['name'] = 'headline_lossy'
    scene['material'] = ('AcousticMaterial', dict(WATER, **visc))
    scene['regions'] = [(idx, ('AcousticMaterial', dict(STEEL, **visc)))
                        for idx, _ in scene['regions']]
    rng = np.random.default_rng(seed)
    n = 2048 * 2048
    scene['init'] = {'pressure': rng.standard_normal(n),
                     'velocity_x': 1e-3 * rng.standard_normal(n),
                     'velocity_y': 1e-3 * rng.standard_normal(n)}
    return scene


@pytest.fixture(scope='module')
def lossy_oracle():
    scene = dict(lossy_scene(9), plan=[('simulate', 5), ('simulate', 4)])
    return scene, scenes.run_oracle(scene)


@pytest.mark.parametrize('engine', ['stream64x12', 'ring64x12', 'tile'])
def test_acoustic2d_2048_lossy_matches_oracle_bitwise(lossy_oracle, engine):
    """Lossy Acoustic2D at the benchmark's size: the register and LDS-ring streaming shapes
    the autotuner picks from, and the LDS-tile engine, against the oracle bit for bit."""
    scene, ref = lossy_oracle
    tile, strip = {'stream64x12': (_stream(7), 12), 'ring64x12': (_stream(14), 12),
                   'tile': (1, 0)}[engine]
    print('lossy 2048^2', engine, flush=True)
    got = _run(scene, tile=tile, strip=strip)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (engine, k))


@pytest.fixture(scope='module')
def headline_oracle():
    scene = dict(headline_scene(24), plan=[('simulate', 13), ('simulate', 11)])
    return scene, scenes.run_oracle(scene)


@pytest.mark.parametrize('engine', ['ring64x12', 'ring64x8', 'stream64x24', 'tile64x8'])
def test_acoustic2d_2048_matches_oracle_bitwise(headline_oracle, engine):
    """24 steps in two runs of 13 and 11 (fused launches and their remainders) on the
    LDS-ring shape, the register-path three-step shape and the LDS-tile engine."""
    scene, ref = headline_oracle
    tile, strip = {'ring64x12': (_stream(14), 12), 'ring64x8': (_stream(14), 8),
                   'stream64x24': (_stream(9), 24),
                   'tile64x8': (1, 0)}[engine]
    print('headline 2048^2', engine, flush=True)
    got = _run(scene, tile=tile, strip=strip)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (engine, k))


def thermal_scene(steps, n=4096, seed=3):
    """Thermal2D 4096^2 as the benchmark builds it (two materials in 16x16 blocks, Fo 0.2)."""
    dx = 1e-3
    rho, cp, kmax = 1000.0, 4000.0, 50.0
    rng = np.random.default_rng(seed)
    nb = n // 16
    hot = rng.random((nb, nb)) < 0.5
    blk = np.kron(hot, np.ones((16, 16), bool)).ravel()
    iy, ix = np.divmod(np.arange(n * n), n)
    return dict(
        name='thermal4096', model='Thermal2D',
        kwargs=dict(x_samples=n, x_delta=dx, y_samples=n, y_delta=dx, t_samples=steps,
                    t_delta=0.2 * dx ** 2 * rho * cp / kmax),
        material=('ThermalMaterial', dict(heat_capacity=cp, density=rho,
                                          thermal_conductivity=0.6)),
        regions=[(np.flatnonzero(blk), ('ThermalMaterial', dict(
            heat_capacity=cp, density=rho, thermal_conductivity=kmax)))],
        init={'temperature': np.exp(-0.5 * ((ix - n / 2) ** 2 + (iy - n / 2) ** 2) / 64.0 ** 2)},
        boundaries={},
        outputs={'temperature': [[int(v) for v in rng.choice(n * n, 4, replace=False)]]},
        plan=[('simulate', 5), ('simulate', 4)])


@pytest.mark.parametrize('engine', ['ring64x20', 'ring64x16pf2', 'stream256x24nt'])
def test_thermal2d_4096_matches_oracle_bitwise(engine):
    scene = thermal_scene(9)
    ref = scenes.run_oracle(scene)
    tile, strip = {'ring64x20': (_stream(14), 20), 'ring64x16pf2': (_stream(12), 16),
                   'stream256x24nt': (_stream(6), 24)}[engine]
    print('thermal 4096^2', engine, flush=True)
    got = _run(scene, tile=tile, strip=strip)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (engine, k))


def test_acoustic2d_2048_superposition():
    """The step is linear: 300 steps with two sources 40 cells apart (their waves overlap)
    equal the sum of the runs with each source alone, to rounding (relative L2 <= 1e-12), on
    the default engine; probes the waves have not reached read exactly zero in all three."""
    n = 2048
    a = (n // 2 + n * (n // 2), 1.0)
    b = (n // 2 + 40 + n * (n // 2 + 20), -0.7)
    runs = [_run(headline_scene(300, [a])), _run(headline_scene(300, [b])),
            _run(headline_scene(300, [a, b]))]
    for k in runs[0]:
        if k == 'step':
            continue
        both, s = runs[2][k], runs[0][k] + runs[1][k]
        den = np.linalg.norm(both)
        if den == 0:
            assert not np.any(s), k
            continue
        err = np.linalg.norm(both - s) / den
        print('superposition', k, '%.3e' % err)
        assert err <= 1e-12, (k, err)


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
def test_acoustic2d_2048_fp32_forms_within_tolerance(form):
    """300 steps: the FP32 forms stay within 1e-5 relative L2 of the FP64 (oracle-exact) run;
    probes the wave has not reached read exactly zero in both."""
    scene = headline_scene(300)
    ref = _run(scene)
    got = _run(scene, form=form)
    for k in ref:
        if k == 'step':
            continue
        den = np.linalg.norm(ref[k])
        if den == 0:
            assert not np.any(got[k]), (form, k)
            continue
        err = np.linalg.norm(got[k] - ref[k]) / den
        print(form, k, '%.3e' % err)
        assert err <= 1e-5, (form, k, err)
