"""Device parity: pyfds_amd (HIP kernels through the C-ABI) against the reference's golden
fixtures and the CPU oracle.

* FP64 form: bit-identical to the reference's float64 results (golden fixtures written by
  running the reference) for every scene.  IdealGas1D's nonlinear equation of state
  p = rho0 c^2/gamma (x**gamma - 1), x = (rho0 + rho')/rho0, depends on the power the host's
  numpy uses (AVX-512 SVML or glibc pow; the two reference runs differ by 2.7e-12 relative
  L2).  The device's power is correctly rounded, and glibc's is on every call of that scene,
  so the device equals the glibc-path fixture (``<scene>_libm.npz``) bit for bit.
* FP32 and FP32-Kahan forms: relative L2 error against the FP64 oracle within the per-scene
  tolerances below (stated per result; FP32 rounding of coefficients and states).
"""
import os

import numpy as np
import pytest

import pyfds_amd
import scenes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')

pytestmark = pytest.mark.gpu


def run_device(scene, form='fp64', **attrs):
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    fld.form = form
    for k, v in attrs.items():
        setattr(fld, k, v)
    scenes.run_plan(fld, scene)
    return scenes.collect(fld, scene)


def rel_l2(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    den = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (den if den > 0 else 1.0)


@pytest.mark.parametrize('name', sorted(scenes.SCENES))
def test_fp64_matches_reference_golden(name):
    scene = scenes.SCENES[name]()
    suffix = '_libm' if name in scenes.POW_SCENES else ''
    gold = np.load(os.path.join(GOLDEN, name + suffix + '.npz'))
    print('scene', name, flush=True)
    res = run_device(scene)
    keys = [k for k in gold.files if k != 'fingerprint']
    assert sorted(keys) == sorted(res)
    for k in keys:
        np.testing.assert_array_equal(res[k], gold[k], err_msg='%s %s' % (name, k))


# FP32 forms: relative L2 bound against the FP64 oracle (values and signals)
FP32_TOL = {name: 1e-5 for name in scenes.SCENES}


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('name', sorted(scenes.SCENES))
def test_fp32_forms_against_oracle(name, form):
    scene = scenes.SCENES[name]()
    ref = scenes.run_oracle(scene)
    print('scene', name, form, flush=True)
    res = run_device(scene, form=form)
    for k in ref:
        if k == 'step':
            assert int(res[k]) == int(ref[k])
            continue
        err = rel_l2(res[k], ref[k])
        print('  %-28s rel L2 %.3e' % (k, err))
        assert err <= FP32_TOL[name], (name, form, k, err)


@pytest.mark.parametrize('tile', range(5))
def test_fp64_tile_shapes_bit_identical(tile):
    """Every 2D tile shape gives the same bits (tiling must not change the arithmetic)."""
    scene = scenes.scene_a2d_main()
    gold = np.load(os.path.join(GOLDEN, 'a2d_main.npz'))
    res = run_device(scene, tile=tile)
    for k in gold.files:
        if k != 'fingerprint':
            np.testing.assert_array_equal(res[k], gold[k], err_msg=k)


# Row-streaming engine: shape indices after the LDS-tile table (lossless 2D models only; lossy
# fields asked for a streaming shape run on the default tile shape).  Slots index the streaming
# table: 0 (256 x 16, one step per launch), 1 (256 x 16, two fused steps), 2 (128 x 32, two),
# 3 (256 x 32, three), 4-6 the same with non-temporal state traffic, 7-9 one-wave 64-wide
# shapes with one, two and three steps, 10-11 64-wide non-temporal with one and two steps
# (tests/test_host_api.py pins these traits), 12-14 loading their rows straight into LDS rings:
# 64 x 32 three steps two rows ahead, 64 x 16 two steps two rows ahead (lossless single fields
# only), 64 x 32 three steps one row ahead (lossy fields run the lossy kernel on it); a single
# field whose rows do not start on 16-byte boundaries (nx odd in FP64, not a multiple of 4 in
# FP32) runs the register-path kernel of the same tiling.
NSTREAM = 15


def _stream_indices():
    return [i for i, e in enumerate(pyfds_amd._native.tile_engines()) if e == 'stream']


LOSSLESS_2D = ['a2d_main', 'a2d_wrap', 'axi_lossless', 'bounds_semantics', 't2d_aniso',
               't2d_blocks', 'taxi']


@pytest.mark.parametrize('slot', range(NSTREAM))
@pytest.mark.parametrize('name', LOSSLESS_2D + ['a2d_lossy', 'axi_lossy'])
def test_fp64_stream_shapes_match_reference_golden(name, slot):
    idx = _stream_indices()
    if slot >= len(idx):
        pytest.skip('fewer streaming shapes compiled')
    scene = scenes.SCENES[name]()
    gold = np.load(os.path.join(GOLDEN, name + '.npz'))
    print('scene', name, 'shape', idx[slot], flush=True)
    info = pyfds_amd._native.tile_table()[idx[slot]]
    if 'lossy' in name and info['single_only']:
        with pytest.raises(ValueError):    # prefetch shapes serve lossless fields only
            run_device(scene, tile=idx[slot])
        return
    res = run_device(scene, tile=idx[slot])
    for k in gold.files:
        if k != 'fingerprint':
            np.testing.assert_array_equal(res[k], gold[k], err_msg='%s %s' % (name, k))


@pytest.mark.parametrize('form,slot,strip', [('fp32', 8, 12), ('fp32', 8, 20), ('fp32', 9, 24),
                                             ('fp32_kahan', 9, 20), ('fp32_kahan', 9, 12),
                                             ('fp32_kahan', 8, 24)])
def test_fp32_stream_strip_heights_equal_tile_engine(form, slot, strip):
    """FP32 forms on strip heights other than the shape's own: the tile engine's bits."""
    scene = _big_acoustic(210, 97, 120, model='Acoustic3DAxi' if slot == 9 else 'Acoustic2D')
    a = run_device(scene, form=form, tile=1)
    b = run_device(scene, form=form, tile=_stream_indices()[slot], strip_height=strip)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg='%s %d %d %s' % (form, slot, strip, k))


@pytest.mark.parametrize('name,slot,strip', [('t2d_blocks', 9, 20), ('t2d_aniso', 6, 24),
                                             ('taxi', 9, 12), ('t2d_blocks', 14, 16),
                                             ('taxi', 12, 8)])
def test_kahan_thermal_stream_strip_heights_equal_tile_engine(name, slot, strip):
    scene = scenes.SCENES[name]()
    a = run_device(scene, form='fp32_kahan', tile=1)
    b = run_device(scene, form='fp32_kahan', tile=_stream_indices()[slot], strip_height=strip)
    for k in a:
        np.testing.assert_array_equal(a[k], b[k], err_msg='%s %d %d %s' % (name, slot, strip, k))


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('name', LOSSLESS_2D)
def test_fp32_stream_engine_equals_tile_engine(name, form):
    """Both engines evaluate the same expressions in the same order, so the FP32 forms agree
    bit for bit as well."""
    scene = scenes.SCENES[name]()
    print('scene', name, form, flush=True)
    a = run_device(scene, form=form, tile=1)
    for idx in [_stream_indices()[i] for i in (0, 3, 5, 8, 9)]:
        b = run_device(scene, form=form, tile=idx)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %s %d' % (name, k, idx))


@pytest.mark.parametrize('model,nx,ny,slot', [('Acoustic2D', 256, 200, 0),
                                              ('Acoustic2D', 300, 75, 2),
                                              ('Acoustic3DAxi', 300, 130, 1),
                                              ('Acoustic3DAxi', 513, 171, 3),
                                              ('Acoustic2D', 130, 101, 4),
                                              ('Acoustic2D', 700, 90, 5),
                                              ('Acoustic3DAxi', 260, 133, 6),
                                              ('Acoustic2D', 509, 97, 6),
                                              ('Acoustic3DAxi', 251, 64, 7),
                                              ('Acoustic2D', 190, 77, 8),
                                              ('Acoustic3DAxi', 333, 101, 9),
                                              ('Acoustic2D', 61, 45, 9),
                                              ('Acoustic3DAxi', 140, 50, 10),
                                              ('Acoustic2D', 200, 120, 11),
                                              ('Acoustic2D', 334, 101, 12),
                                              ('Acoustic3DAxi', 190, 77, 13),
                                              ('Acoustic2D', 510, 97, 14),
                                              ('Acoustic3DAxi', 252, 100, 14)])
def test_fp64_stream_multi_strip_grids_match_oracle_bitwise(model, nx, ny, slot):
    """Partial column blocks at the right edge and partial strips at the bottom."""
    idx = _stream_indices()[slot]
    scene = _big_acoustic(nx, ny, 120, model=model)
    ref = scenes.run_oracle(scene)
    print('grid', model, nx, ny, 'shape', idx, flush=True)
    res = run_device(scene, tile=idx)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


# Lossy acoustics on the streaming engine: one step per launch on every shape, one ghost
# column per side, the viscous term's old velocities of the rows above and below.
@pytest.mark.parametrize('model,nx,ny,slot', [('Acoustic2D', 256, 200, 0),
                                              ('Acoustic2D', 300, 75, 1),
                                              ('Acoustic3DAxi', 300, 130, 2),
                                              ('Acoustic3DAxi', 513, 171, 3),
                                              ('Acoustic2D', 130, 101, 4),
                                              ('Acoustic3DAxi', 260, 133, 6),
                                              ('Acoustic2D', 509, 97, 7),
                                              ('Acoustic3DAxi', 251, 64, 7),
                                              ('Acoustic2D', 190, 77, 8),
                                              ('Acoustic3DAxi', 333, 101, 9),
                                              ('Acoustic2D', 61, 45, 10),
                                              ('Acoustic3DAxi', 200, 120, 11),
                                              ('Acoustic2D', 256, 200, 14),
                                              ('Acoustic3DAxi', 260, 133, 14),
                                              ('Acoustic2D', 509, 97, 14)])
def test_fp64_lossy_stream_grids_match_oracle_bitwise(model, nx, ny, slot):
    idx = _stream_indices()[slot]
    scene = _big_acoustic(nx, ny, 120, model=model, mu=3e-3)
    ref = scenes.run_oracle(scene)
    print('lossy grid', model, nx, ny, 'shape', idx, flush=True)
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    fld.tile = idx
    scenes.run_plan(fld, scene)
    assert pyfds_amd._native.engine_of(fld._runner().ctx) == 'stream'
    res = scenes.collect(fld, scene)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


@pytest.mark.parametrize('model,nx,ny,slot,strip', [('Acoustic2D', 190, 77, 7, 12),
                                                    ('Acoustic3DAxi', 200, 97, 9, 20),
                                                    ('Acoustic2D', 150, 71, 8, 6),
                                                    ('Acoustic3DAxi', 170, 60, 1, 24),
                                                    ('Acoustic2D', 260, 90, 3, 12),
                                                    ('Acoustic2D', 190, 77, 14, 12),
                                                    ('Acoustic3DAxi', 200, 97, 14, 6)])
def test_fp64_lossy_stream_strip_heights_match_oracle_bitwise(model, nx, ny, slot, strip):
    idx = _stream_indices()[slot]
    scene = _big_acoustic(nx, ny, 120, model=model, mu=3e-3)
    ref = scenes.run_oracle(scene)
    print('lossy grid', model, nx, ny, 'shape', idx, 'strip', strip, flush=True)
    res = run_device(scene, tile=idx, strip_height=strip)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('name', ['a2d_lossy', 'axi_lossy'])
def test_fp32_lossy_stream_equals_tile_engine(name, form):
    scene = scenes.SCENES[name]()
    print('scene', name, form, flush=True)
    a = run_device(scene, form=form, tile=1)
    for idx in [_stream_indices()[i] for i in (0, 2, 7, 10, 14)]:
        b = run_device(scene, form=form, tile=idx)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %s %d' % (name, k, idx))


@pytest.mark.parametrize('slot', [None, 1, 3, 8])
@pytest.mark.parametrize('chunk', [0, 1, 7, 50])
def test_graph_chunks_bit_identical(chunk, slot):
    """HIP-graph replayed step chunks (and plain launches) give identical results, also with
    several steps fused per launch (chunks rounded to whole launches, remainders run as
    single steps)."""
    scene = scenes.scene_bounds_semantics()
    gold = np.load(os.path.join(GOLDEN, 'bounds_semantics.npz'))
    attrs = {} if slot is None else dict(tile=_stream_indices()[slot])
    res = run_device(scene, graph_chunk=chunk, **attrs)
    for k in gold.files:
        if k != 'fingerprint':
            np.testing.assert_array_equal(res[k], gold[k], err_msg=k)


def _big_acoustic(nx, ny, steps, form='fp64', seed=1, model='Acoustic2D', mu=0.0):
    dx = 1e-4
    dt = 0.5 * dx / 5900.0
    t = np.arange(steps) * dt
    rng = np.random.default_rng(seed)
    w, h = rng.integers(nx // 8, nx // 4), rng.integers(ny // 8, ny // 4)
    x0, y0 = rng.integers(0, nx - w), rng.integers(0, ny - h)
    incl = scenes.rect_indices(nx, int(x0), int(y0), int(w), int(h))
    return dict(
        name='big', model=model,
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx, t_samples=steps,
                    t_delta=dt),
        material=('AcousticMaterial', dict(scenes.WATER, absorption_coef=mu)),
        regions=[(incl, ('AcousticMaterial', dict(scenes.STEEL)))],
        init={},
        boundaries={'pressure': [dict(indices=[nx // 2 + nx * (ny // 2)],
                                      value=np.sin(2 * np.pi * 1e6 * t) * (t < 5e-6),
                                      additive=True)]},
        outputs={'pressure': [[int(v) for v in rng.choice(nx * ny, 16, replace=False)]]},
        plan=[('simulate', None)])


# Strip heights other than the shape's own (the autotuner's candidates): one-, two- and
# three-step shapes, partial strips at the bottom, ghost rows crossing strip boundaries.
@pytest.mark.parametrize('model,nx,ny,slot,strip', [('Acoustic2D', 300, 130, 8, 6),
                                                    ('Acoustic3DAxi', 250, 101, 8, 24),
                                                    ('Acoustic2D', 190, 77, 9, 12),
                                                    ('Acoustic2D', 333, 101, 9, 24),
                                                    ('Acoustic3DAxi', 200, 97, 9, 20),
                                                    ('Acoustic2D', 260, 90, 3, 12),
                                                    ('Acoustic3DAxi', 270, 83, 6, 24),
                                                    ('Acoustic2D', 130, 61, 2, 20),
                                                    ('Acoustic2D', 150, 71, 7, 20),
                                                    ('Acoustic3DAxi', 180, 95, 1, 6),
                                                    ('Acoustic2D', 210, 85, 11, 12),
                                                    ('Acoustic2D', 334, 101, 12, 24),
                                                    ('Acoustic3DAxi', 200, 97, 12, 20),
                                                    ('Acoustic2D', 150, 71, 13, 12),
                                                    ('Acoustic2D', 270, 83, 14, 24),
                                                    ('Acoustic2D', 262, 93, 14, 8),
                                                    ('Acoustic3DAxi', 200, 97, 14, 16),
                                                    ('Acoustic2D', 190, 77, 12, 8),
                                                    ('Acoustic3DAxi', 150, 71, 12, 16)])
def test_fp64_stream_strip_heights_match_oracle_bitwise(model, nx, ny, slot, strip):
    idx = _stream_indices()[slot]
    scene = _big_acoustic(nx, ny, 120, model=model)
    ref = scenes.run_oracle(scene)
    print('grid', model, nx, ny, 'shape', idx, 'strip', strip, flush=True)
    res = run_device(scene, tile=idx, strip_height=strip)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


@pytest.mark.parametrize('model,mu,nx,ny', [('Acoustic2D', 0.0, 256, 200),
                                            ('Acoustic3DAxi', 0.0, 300, 130),
                                            ('Acoustic3DAxi', 3e-3, 201, 77),
                                            ('Acoustic2D', 2e-3, 130, 101)])
def test_fp64_multi_tile_grids_match_oracle_bitwise(model, mu, nx, ny):
    """Grids spanning many tiles, with partial tiles at the right and bottom edges."""
    scene = _big_acoustic(nx, ny, 120, model=model, mu=mu)
    ref = scenes.run_oracle(scene)
    res = run_device(scene)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


@pytest.mark.parametrize('tile', [None, 'stream0', 'stream2', 'stream3', 'stream6', 'stream8',
                                  'stream9', 'stream11', 'stream5/12', 'stream9/24',
                                  'stream3/20', 'stream6/24', 'stream9/20', 'stream12',
                                  'stream13/12', 'stream14/20', 'stream12/6', 'stream14/8',
                                  'stream12/16', 'stream14/16'])
@pytest.mark.parametrize('model', ['Thermal2D', 'Thermal3DAxi'])
def test_fp64_thermal_multi_tile_matches_oracle_bitwise(model, tile):
    nx, ny, dx = 190, 150, 1e-3
    rng = np.random.default_rng(5)
    scene = dict(
        name='tbig', model=model,
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx, t_samples=150,
                    t_delta=0.2 * dx ** 2 * 4e6 / 50.0),
        material=('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                          thermal_conductivity=0.6)),
        regions=[(scenes._thermal_blocks(nx, ny, 16, 7),
                  ('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                           thermal_conductivity=(50.0, 30.0))))],
        init={'temperature': rng.random(nx * ny)},
        boundaries={'temperature': [dict(indices=list(range(0, nx * ny, nx)), value=0.0,
                                         additive=False)]},
        outputs={'temperature': [[int(v) for v in rng.choice(nx * ny, 8, replace=False)]],
                 'heat_flux_y': [[nx * 70 + 3]]},
        plan=[('simulate', 77), ('simulate', 73)])
    ref = scenes.run_oracle(scene)
    attrs = {}
    if tile is not None:
        slot, _, strip = tile[6:].partition('/')
        attrs = dict(tile=_stream_indices()[int(slot)])
        if strip:
            attrs['strip_height'] = int(strip)
    res = run_device(scene, **attrs)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg=k)


@pytest.mark.parametrize('model', ['Acoustic2D', 'Acoustic3DAxi'])
@pytest.mark.parametrize('slot', [3, 8, 9, 12, 13, 14])
def test_fp64_fused_remainder_launches_match_oracle_bitwise(model, slot):
    # runs whose length is not a multiple of the fused stage count end in one launch of the
    # remaining stages (two-stage launches on three-step shapes)
    idx = _stream_indices()[slot]
    scene = dict(_big_acoustic(200, 97, 130, model=model),
                 plan=[('simulate', 62), ('simulate', 59), ('simulate', 5)])
    print('remainder', model, 'shape', idx, flush=True)
    ref = scenes.run_oracle(scene)
    res = run_device(scene, tile=idx, graph_chunk=9)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg='%s %s %d' % (model, k, idx))


@pytest.mark.parametrize('slot', [3, 9, 12, 14])
def test_fp64_thermal_fused_remainder_launches_match_reference(slot):
    idx = _stream_indices()[slot]
    scene = dict(scenes.SCENES['t2d_blocks'](), plan=[('simulate', 62), ('simulate', 59)])
    ref = scenes.run_oracle(scene)
    res = run_device(scene, tile=idx, graph_chunk=9)
    for k in ref:
        np.testing.assert_array_equal(res[k], ref[k], err_msg='%s %d' % (k, idx))


# LDS-ring shapes (12-14): rows reach LDS by direct 16-byte loads, so rows must start on
# 16-byte boundaries; FP32 forms (4-byte elements) and tiles reading outside the arrays.
@pytest.mark.parametrize('form,nx', [('fp64', 201), ('fp32', 202), ('fp32_kahan', 206)])
def test_ring_shapes_with_unaligned_rows_run_their_register_twin(form, nx):
    """Rows that do not start on 16-byte boundaries: the ring shape runs the register-path
    kernel of its tiling (slot 9 for 12 and 14, slot 8 for 13), bit for bit."""
    scene = dict(_big_acoustic(nx, 40, 12), plan=[('simulate', 7), ('simulate', 5)])
    for slot, twin in ((12, 9), (13, 8), (14, 9)):
        a = run_device(scene, form=form, tile=_stream_indices()[slot])
        b = run_device(scene, form=form, tile=_stream_indices()[twin])
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %d %s' % (form, slot, k))
    if form == 'fp64':
        ref = scenes.run_oracle(scene)
        for k in ref:
            np.testing.assert_array_equal(a[k], ref[k], err_msg=k)


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('model,nx,ny', [('Acoustic2D', 212, 97), ('Acoustic3DAxi', 132, 61),
                                         ('Acoustic2D', 64, 48)])
def test_fp32_ring_shapes_equal_tile_engine(model, nx, ny, form):
    scene = dict(_big_acoustic(nx, ny, 120, model=model),
                 plan=[('simulate', 61), ('simulate', 59)])
    a = run_device(scene, form=form, tile=1)
    for slot, strip in ((12, 0), (13, 0), (14, 0), (12, 12), (14, 20), (14, 8), (12, 16)):
        print('ring', model, nx, ny, form, slot, strip, flush=True)
        b = run_device(scene, form=form, tile=_stream_indices()[slot], strip_height=strip)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %d %d %s' % (form, slot,
                                                                                strip, k))


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
@pytest.mark.parametrize('model', ['Thermal2D', 'Thermal3DAxi'])
def test_fp32_thermal_ring_shapes_equal_tile_engine(model, form):
    nx, ny, dx = 196, 150, 1e-3
    rng = np.random.default_rng(9)
    scene = dict(
        name='tring', model=model,
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx, t_samples=150,
                    t_delta=0.2 * dx ** 2 * 4e6 / 50.0),
        material=('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                          thermal_conductivity=0.6)),
        regions=[(scenes._thermal_blocks(nx, ny, 16, 7),
                  ('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                           thermal_conductivity=(50.0, 30.0))))],
        init={'temperature': rng.random(nx * ny)},
        boundaries={'temperature': [dict(indices=list(range(0, nx * ny, nx)), value=0.0,
                                         additive=False)]},
        outputs={'temperature': [[int(v) for v in rng.choice(nx * ny, 8, replace=False)]]},
        plan=[('simulate', 77), ('simulate', 73)])
    a = run_device(scene, form=form, tile=1)
    for slot in (12, 13, 14):
        b = run_device(scene, form=form, tile=_stream_indices()[slot])
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %d %s' % (form, slot, k))


def test_values_array_held_across_runs_is_updated_in_place():
    """The reference updates ``values`` in place (``values -= ...``): an array taken from
    ``.values`` before a run is the array ``.values`` returns afterwards, holding the new state,
    and edits made to it in place are honoured by the next run, as an assignment would be."""
    scene = dict(scenes.SCENES['a2d_main'](), plan=[('simulate', 10)])
    ref = scenes.run_oracle(scene)
    f = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    v = f.pressure.values
    f.simulate(10)
    w = f.pressure.values
    assert w is v
    np.testing.assert_array_equal(v, ref['values/pressure'])
    v[7] += 0.25                                   # in-place edit of the held array
    f.simulate(10)
    g = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    g.simulate(10)
    p = np.array(g.pressure.values)
    p[7] += 0.25
    g.pressure.values = p                          # the same edit by assignment
    g.simulate(10)
    np.testing.assert_array_equal(f.pressure.values, g.pressure.values)
    np.testing.assert_array_equal(f.velocity_x.values, g.velocity_x.values)


@pytest.mark.gpu
@pytest.mark.parametrize('elem,ns', [(8, 3), (4, 1)])
def test_pattern_bench_requests_the_step_kernels_bytes(elem, ns):
    """The memory-only pattern kernel runs and requests what a three-stage ring kernel of the
    same tiling moves: every tile loads H + 2G + 2 rows of ns + 2 windows of CW elements (CW =
    68 in float64, 72 in float32) and every cell is stored once per state array."""
    import ctypes as ct
    from pyfds_amd import _native
    lib = _native.load()
    ms, b = ct.c_double(), ct.c_double()
    nx, ny, G, H = 512, 256, 2, 12
    rc = lib.fds_stream_pattern_bench(nx, ny, elem, 3, G, H, ns, 2, ct.byref(ms), ct.byref(b))
    assert rc == 0, rc
    ntiles = -(-nx // (64 - 2 * G)) * -(-ny // H)
    cw = 68 if elem == 8 else 72
    assert b.value == (ntiles * (H + 2 * G + 2) * (ns + 2) * cw + nx * ny * ns) * elem
    assert 0 < ms.value < 100
