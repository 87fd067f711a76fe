"""Resident 1D engine (one workgroup keeps an Acoustic1D / Thermal1D / IdealGas1D /
Acoustic2ndOrder1D line in registers for up to 2048 steps per launch).

GPU: FP64 bit-identical to the oracle for lines that exercise every event path (cached
boundary cell, further cells of one thread, overlapping boundaries on one cell, several
recorded points per thread, runs that cross the 2048-step launch limit, ``sim_step`` and
``reset`` between runs), for every cells-per-thread variant; the FP32 and FP32-Kahan forms
equal the LDS-tile engine's bits (both evaluate the same expressions in the same order); the
engine choice follows the line length and the tile setting.  The nonlinear models: FP64
bit-identical to the oracle (linear and second-order equations of state), and every form equal
to the nonlinear tile kernel's bits.  The ideal-gas state's power is the correctly rounded one
(``fds_crmath.h``): its FP64 runs equal the oracle with ``pow_cr`` bit for bit, on one
workgroup and on segments.  The reference's own 1D scenes (golden fixtures) run on this engine
by default in ``test_device_parity.py``.
"""
import numpy as np
import pytest

import pyfds_amd
from pyfds_amd import _native as nat
import scenes

pytestmark = pytest.mark.gpu

TILE_ROW = 5          # the LDS-tile engine's one-row shape
CPT = (1, 2, 4, 8)
MAX_THREADS = {1: 1024, 2: 1024, 4: 512, 8: 512}


def _signal(rng, steps):
    return rng.standard_normal(steps + 5)


def line_acoustic(n, steps, lossy, seed):
    rng = np.random.default_rng(seed)
    dx = 1e-4
    dt = 0.4 * dx / 2500.0
    i = np.arange(n)
    a = n // 3
    mat = dict(scenes.WATER)
    if lossy:
        mat.update(shear_viscosity=1e-3, bulk_viscosity=2e-3)
    return dict(
        name='line_a1d', model='Acoustic1D',
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt),
        material=('AcousticMaterial', mat),
        regions=[(list(range(n // 2, n // 2 + n // 5)),
                  ('AcousticMaterial', dict(sound_velocity=2500.0, density=1800.0,
                                            absorption_coef=5e-3 if lossy else 0)))],
        init={'pressure': np.exp(-0.5 * ((i - n / 4) / 9.0) ** 2),
              'velocity': 1e-6 * rng.standard_normal(n)},
        boundaries={
            'pressure': [
                dict(indices=[a], value=_signal(rng, steps), additive=True),
                # three neighbouring cells: one thread's block holds several of them
                dict(indices=[a + 5, a + 6, a + 7],
                     value=[_signal(rng, steps) for _ in range(3)], additive=True),
                # two boundaries on one cell: a two-entry program
                dict(indices=[a + 6, 2 * n // 3], value=0.25, additive=True),
                dict(indices=[n - 1], value=0, additive=False)],
            'velocity': [dict(indices=[0], value=0, additive=False),
                         dict(indices=[n - 1], value=0, additive=False),
                         dict(indices=[n // 2 + 1], value=_signal(rng, steps) * 1e-4,
                              additive=True)]},
        outputs={'pressure': [[5, a + 5, a + 6, a + 7, n - 1], [a + 6, a + 6]],
                 'velocity': [[0, n // 2 + 1, n // 2 + 2]]},
        plan=[('simulate', 2100), ('sim_step',), ('simulate', steps - 2101 - 40),
              ('reset',), ('simulate', 40)])


def line_thermal(n, steps, seed):
    rng = np.random.default_rng(seed)
    dx = 1e-3
    dt = 0.2 * dx ** 2 * 1000.0 * 4000.0 / 50.0
    i = np.arange(n)
    return dict(
        name='line_t1d', model='Thermal1D',
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt),
        material=('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                          thermal_conductivity=0.6)),
        regions=[(list(range(n // 3, n // 3 + n // 4)),
                  ('ThermalMaterial', dict(heat_capacity=3000.0, density=2000.0,
                                           thermal_conductivity=50.0)))],
        init={'temperature': np.exp(-0.5 * ((i - n / 2) / 20.0) ** 2)},
        boundaries={
            'temperature': [dict(indices=[0], value=1.0, additive=False),
                            dict(indices=[n // 4, n // 4 + 1], value=0.01 * _signal(rng, steps),
                                 additive=True)],
            'heat_flux': [dict(indices=[n - 1], value=0.0, additive=False),
                          dict(indices=[n // 2], value=_signal(rng, steps), additive=True)]},
        outputs={'temperature': [[3, n // 4, n // 4 + 1]], 'heat_flux': [[n // 2, n - 1]]},
        plan=[('simulate', 2100), ('sim_step',), ('simulate', steps - 2101)])


def line_nonlinear(model, n, steps, seed, convective=True, nl_state=False):
    rng = np.random.default_rng(seed)
    dx = 1e-3
    t = np.arange(steps + 5)
    burst = np.sin(2 * np.pi * t / 25.0) * (t < 100)
    if model == 'IdealGas1D':
        c = 323.0
        material = ('AcousticMaterial', dict(sound_velocity=c, density=1.66,
                                             shear_viscosity=2.2e-5, isobaric_heat_cap=520.0,
                                             isochoric_heat_cap=312.0))
        amp, extra = 200.0, dict(model_kwargs=dict(convective=convective, nl_state=nl_state))
    else:
        c = 1500.0
        material = ('AcousticMaterial2ndOrder', dict(sound_velocity=c, density=1000.0,
                                                     shear_viscosity=1e-3, d_rho_p=c ** 2,
                                                     d_rho2_p=2.0e7))
        amp, extra = 1e5, {}
    a = n // 6
    return dict(
        name='line_nl', model=model, **extra,
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=steps, t_delta=0.4 * dx / c),
        material=material, regions=[], init={},
        boundaries={
            'pressure': [dict(indices=[a], value=amp * burst, additive=True),
                         dict(indices=[a + 3, a + 4, a + 5],
                              value=[1e-3 * amp * rng.standard_normal(steps + 5)
                                     for _ in range(3)], additive=True),
                         dict(indices=[a + 4, n // 2], value=1e-3 * amp, additive=True)],
            'velocity': [dict(indices=[0], value=0, additive=False),
                         dict(indices=[n - 1], value=0, additive=False)],
            'density': [dict(indices=[2 * n // 3], value=1e-9 * rng.standard_normal(steps + 5),
                             additive=True)]},
        outputs={'pressure': [[a + 3, a + 4, a + 5, n // 2]], 'velocity': [[n // 3, 1]],
                 'density': [[2 * n // 3, n - 1]]},
        plan=[('simulate', 2100), ('sim_step',), ('simulate', steps - 2101)])


NL_SCENES = {
    'ig1d_clin_700': lambda: line_nonlinear('IdealGas1D', 700, 2300, 7, True, False),
    'ig1d_nlin_2048': lambda: line_nonlinear('IdealGas1D', 2048, 2200, 8, False, False),
    'a2o1d_1000': lambda: line_nonlinear('Acoustic2ndOrder1D', 1000, 2200, 9),
    # the nonlinear equation of state (correctly rounded power)
    'ig1d_cnl_900': lambda: line_nonlinear('IdealGas1D', 900, 2200, 10, True, True),
}
NL_MAX_THREADS = {1: 512, 2: 512, 4: 512, 8: 256}


SCENES = {
    'a1d_900': lambda: line_acoustic(901, 2500, False, 1),
    'a1d_lossy_900': lambda: line_acoustic(901, 2500, True, 2),
    'a1d_2000': lambda: line_acoustic(2000, 2300, False, 3),
    'a1d_lossy_3000': lambda: line_acoustic(3000, 2300, True, 4),
    't1d_900': lambda: line_thermal(900, 2400, 5),
    't1d_4096': lambda: line_thermal(4096, 2200, 6),
}


def _n(scene):
    return scene['kwargs']['x_samples']


def _fits(n, cpt):
    return -(-n // cpt) <= MAX_THREADS[cpt]


def run(scene, form='fp64', **attrs):
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    fld.form = form
    for k, v in attrs.items():
        setattr(fld, k, v)
    scenes.run_plan(fld, scene)
    return scenes.collect(fld, scene), nat.engine_of(fld._runner().ctx)


CASES = [(name, c) for name, make in SCENES.items() for c in (0,) + CPT
         if c == 0 or _fits(_n(make()), c)]


@pytest.mark.parametrize('name,cpt', CASES)
def test_line_fp64_matches_oracle_bitwise(name, cpt):
    scene = SCENES[name]()
    print('line scene', name, 'cells per thread', cpt, flush=True)
    ref = scenes.run_oracle(scene)
    res, eng = run(scene, strip_height=cpt)
    assert eng == 'line'
    assert sorted(res) == sorted(ref)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg='%s %s' % (name, k))


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('name', sorted(SCENES))
def test_line_fp32_forms_equal_tile_engine(name, form):
    scene = SCENES[name]()
    print('line scene', name, form, flush=True)
    line, eng = run(scene, form=form)
    tile, eng_t = run(scene, form=form, tile=TILE_ROW)
    assert (eng, eng_t) == ('line', 'tile')
    for k in tile:
        np.testing.assert_array_equal(line[k], tile[k], err_msg='%s %s %s' % (name, form, k))


def test_engine_choice():
    mat = pyfds_amd.AcousticMaterial(1500.0, 1000.0)

    def field(n, **attrs):
        f = pyfds_amd.Acoustic1D(x_samples=n, x_delta=1e-4, t_samples=4, t_delta=1e-8,
                                 material=mat)
        for k, v in attrs.items():
            setattr(f, k, v)
        f.simulate(2)
        return nat.engine_of(f._runner().ctx)

    assert field(4096) == 'line'
    assert field(100000) == 'line'           # segments
    assert field((1 << 24) + 1) == 'tile'     # past the resident engine's table limit
    assert field(100, tile=TILE_ROW) == 'tile'
    with pytest.raises(ValueError):
        field(3000, strip_height=2)       # 1500 threads: beyond the 2-cell variant's 1024
    with pytest.raises(ValueError):
        field(100, strip_height=3)


NL_CASES = [(name, c) for name, make in NL_SCENES.items()
            for c in (0,) + CPT if c == 0 or -(-_n(make()) // c) <= NL_MAX_THREADS[c]]


@pytest.mark.parametrize('name,cpt', NL_CASES)
def test_nonlinear_line_fp64_matches_oracle_bitwise(name, cpt):
    scene = NL_SCENES[name]()
    print('nonlinear line scene', name, 'cells per thread', cpt, flush=True)
    ref = scenes.run_oracle(scene, power='cr')
    res, eng = run(scene, strip_height=cpt)
    assert eng == 'line'
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg='%s %s' % (name, k))


@pytest.mark.parametrize('form', ['fp64', 'fp32', 'fp32_kahan'])
@pytest.mark.parametrize('name', sorted(NL_SCENES))
def test_nonlinear_line_equals_tile_engine(name, form):
    scene = NL_SCENES[name]()
    print('nonlinear line scene', name, form, flush=True)
    line, eng = run(scene, form=form)
    tile, eng_t = run(scene, form=form, tile=0)
    assert (eng, eng_t) == ('line', 'nonlinear')
    for k in tile:
        np.testing.assert_array_equal(line[k], tile[k], err_msg='%s %s %s' % (name, form, k))


def test_nonlinear_engine_choice():
    mat = pyfds_amd.AcousticMaterial(323.0, 1.66)

    def field(n):
        f = pyfds_amd.IdealGas1D(x_samples=n, x_delta=1e-3, t_samples=4, t_delta=1e-6,
                                 material=mat)
        f.simulate(2)
        return nat.engine_of(f._runner().ctx)

    assert field(2048) == 'line'
    assert field(50000) == 'line'


# Segmented lines: several workgroups, each stepping its owned cells plus `ghost` cells of each
# neighbour for at most `ghost` steps per launch.  (cells per thread, threads, ghost cells)
SEG_CFGS = [(1, 128, 16), (2, 256, 64), (4, 128, 100), (8, 64, 7), (2, 64, 48), (4, 64, 64)]
SEG_SCENES = {
    'a1d_2000': SCENES['a1d_2000'],
    'a1d_lossy_900': SCENES['a1d_lossy_900'],
    't1d_900': SCENES['t1d_900'],
    'a1d_lossy_9000': lambda: line_acoustic(9000, 2200, True, 11),
    'ig1d_clin_700': NL_SCENES['ig1d_clin_700'],
    'a2o1d_1000': NL_SCENES['a2o1d_1000'],
    'ig1d_cnl_900': NL_SCENES['ig1d_cnl_900'],
}


@pytest.mark.parametrize('cfg', SEG_CFGS)
@pytest.mark.parametrize('name', sorted(SEG_SCENES))
def test_segmented_line_fp64_matches_oracle_bitwise(name, cfg):
    scene = SEG_SCENES[name]()
    print('segmented line scene', name, cfg, flush=True)
    ref = scenes.run_oracle(scene, power='cr')
    res, eng = run(scene, line_shape=cfg)
    assert eng == 'line'
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg='%s %s %s' % (name, cfg, k))


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('name', ['a1d_lossy_900', 't1d_900', 'ig1d_cnl_900'])
def test_segmented_line_fp32_forms_equal_one_segment(name, form):
    make = SEG_SCENES.get(name) or NL_SCENES[name]
    one, _ = run(make(), form=form)
    for shape in ((2, 128, 40), (4, 64, 40)):          # LDS exchange, one-wave DPP exchange
        seg, eng = run(make(), form=form, line_shape=shape)
        assert eng == 'line'
        for k in one:
            np.testing.assert_array_equal(seg[k], one[k],
                                          err_msg='%s %s %s %s' % (name, form, shape, k))


@pytest.mark.parametrize('engine', ['line', 'nonlinear'])
def test_ideal_gas_long_run_matches_cr_oracle_bitwise(engine):
    # IdealGas1D, nonlinear state, 2048 cells x 2000 steps: every one of the ~4 M powers must
    # be the correctly rounded one, or the run drifts (a one-ulp power error grows to ~1e-10
    # relative L2 over this run)
    scene = scenes.scene_ig1d_long()
    print('ideal gas long run', engine, flush=True)
    ref = scenes.run_oracle(scene, power='cr')
    res, eng = run(scene, **({} if engine == 'line' else dict(tile=0)))
    assert eng == engine
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg='%s %s' % (engine, k))
