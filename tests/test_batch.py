"""Batched runs (one launch per step over the tiles of every member) give the same bits as
running each member alone, and the FP64 batch matches the oracle bit for bit."""
import numpy as np
import pytest

import pyfds_amd
import scenes
from pyfds_amd.batch import BatchRunner

pytestmark = pytest.mark.gpu

MEMBERS = ['a2d_main', 'axi_lossless', 'axi_lossy', 'a2d_lossy', 't2d_blocks', 't2d_aniso',
           'bounds_semantics', 'taxi']
PLAN = [('simulate', 17), ('simulate', 23)]


def _members(form='fp64'):
    out = []
    for name in MEMBERS:
        scene = dict(scenes.SCENES[name](), plan=PLAN)
        fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
        fld.form = form
        out.append((scene, fld))
    return out


@pytest.mark.parametrize('chunk', [0, 5])
def test_fp64_batch_matches_oracle(chunk):
    members = _members()
    batch = BatchRunner([f for _, f in members], chunk=chunk)
    for _, n in PLAN:
        batch.simulate(n)
    for scene, fld in members:
        print('member', scene['name'], flush=True)
        ref = scenes.run_oracle(scene)
        got = scenes.collect(fld, scene)
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (scene['name'], k))


@pytest.mark.parametrize('tile', [1, 3])
def test_kahan_batch_equals_single_runs(tile):
    members = _members('fp32_kahan')
    batch = BatchRunner([f for _, f in members], tile=tile, chunk=4)
    for _, n in PLAN:
        batch.simulate(n)
    for scene, fld in members:
        single = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
        single.form = 'fp32_kahan'
        single.tile = tile
        scenes.run_plan(single, scene)
        a, b = scenes.collect(fld, scene), scenes.collect(single, scene)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %s' % (scene['name'], k))


def test_batch_rejects_mixed_forms():
    a = scenes.build_field(pyfds_amd, pyfds_amd.regions, scenes.scene_a2d_wrap())
    b = scenes.build_field(pyfds_amd, pyfds_amd.regions, scenes.scene_a2d_wrap())
    b.form = 'fp32'
    with pytest.raises(ValueError):
        BatchRunner([a, b])


LOSSLESS = ['a2d_main', 'axi_lossless', 't2d_blocks', 't2d_aniso', 'bounds_semantics', 'taxi',
            'a2d_wrap']


def _stream_index(slot):
    return [i for i, e in enumerate(pyfds_amd._native.tile_engines()) if e == 'stream'][slot]


@pytest.mark.parametrize('slot', [0, 2, 3, 6, 8, 9, 11, 14])
def test_fp64_stream_batch_matches_oracle(slot):
    """A batch of lossless members on the row-streaming engine (slot 14: rows streamed into
    LDS rings one dword per lane, members of any width, 7 columns included)."""
    members = []
    for name in LOSSLESS:
        scene = dict(scenes.SCENES[name](), plan=PLAN)
        members.append((scene, scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)))
    batch = BatchRunner([f for _, f in members], tile=_stream_index(slot), chunk=5)
    for _, n in PLAN:
        batch.simulate(n)
    for scene, fld in members:
        print('member', scene['name'], flush=True)
        ref = scenes.run_oracle(scene)
        got = scenes.collect(fld, scene)
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (scene['name'], k))


@pytest.mark.parametrize('slot,strip', [(9, 24), (8, 20), (9, 12), (5, 6), (14, 12), (14, 20),
                                        (14, 8), (14, 16)])
def test_fp64_stream_batch_strip_heights_match_oracle(slot, strip):
    """Batches on strip heights other than the shape's own (the autotuner's candidates)."""
    members = []
    for name in LOSSLESS:
        scene = dict(scenes.SCENES[name](), plan=PLAN)
        fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
        fld.strip_height = strip
        members.append((scene, fld))
    batch = BatchRunner([f for _, f in members], tile=_stream_index(slot), chunk=5)
    for _, n in PLAN:
        batch.simulate(n)
    for scene, fld in members:
        print('member', scene['name'], 'strip', strip, flush=True)
        ref = scenes.run_oracle(scene)
        got = scenes.collect(fld, scene)
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (scene['name'], k))


@pytest.mark.parametrize('slot', [0, 5, 9])
def test_kahan_stream_batch_equals_single_runs(slot):
    tile = _stream_index(slot)
    members = []
    for name in LOSSLESS:
        scene = dict(scenes.SCENES[name](), plan=PLAN)
        fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
        fld.form = 'fp32_kahan'
        members.append((scene, fld))
    batch = BatchRunner([f for _, f in members], tile=tile, chunk=4)
    for _, n in PLAN:
        batch.simulate(n)
    for scene, fld in members:
        single = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
        single.form = 'fp32_kahan'
        single.tile = tile
        scenes.run_plan(single, scene)
        a, b = scenes.collect(fld, scene), scenes.collect(single, scene)
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %s' % (scene['name'], k))


def test_batch_rejects_flow_and_learned_members():
    # the batch launches only the step kernels: a flow shift or learned boundary would be lost
    flow = pyfds_amd.AcousticFlow2D([250.0, 500.0, 1000.0], x_samples=6, x_delta=1e-3,
                                    y_samples=3, y_delta=1e-3, t_samples=4, t_delta=1e-6,
                                    material=pyfds_amd.AcousticMaterial(1500, 1000))
    other = scenes.build_field(pyfds_amd, pyfds_amd.regions, scenes.scene_a2d_wrap())
    with pytest.raises(ValueError):
        BatchRunner([other, flow]).simulate(2)


# ------------------------------------------------------------------ FP32 sweeps vs FP64
SWEEP_STEPS = 200
FP32_SWEEP_TOL = 1e-5


def sweep_scenes(seed=21, n_inst=6):
    """Uneven Acoustic2D / Acoustic3DAxi members as in the batched benchmark (1-3 inclusions,
    c 1000-6000 m/s, rho 800-8000 kg/m^3, dt at CFL 0.5 on each member's c_max, a 1 MHz burst
    at the centre, 8 recorded points); sides 128-640, the first two 512 or more wide."""
    rng = np.random.default_rng(seed)
    dx, steps = 1e-4, SWEEP_STEPS + 8
    out = []
    for i in range(n_inst):
        nx = int(rng.integers(512 if i < 2 else 128, 641))
        ny = int(rng.integers(128, 257 if i < 2 else 641))   # keeps the oracle's run short
        regions, cmax = [], 1500.0
        for _ in range(int(rng.integers(1, 4))):
            w = int(rng.integers(max(2, nx // 10), max(3, nx // 3)))
            h = int(rng.integers(max(2, ny // 10), max(3, ny // 3)))
            x0, y0 = int(rng.integers(0, nx - w)), int(rng.integers(0, ny - h))
            c, rho = float(rng.uniform(1000, 6000)), float(rng.uniform(800, 8000))
            cmax = max(cmax, c)
            regions.append((scenes.rect_indices(nx, x0, y0, w, h),
                            ('AcousticMaterial', dict(sound_velocity=c, density=rho))))
        dt = 0.5 * dx / cmax
        t = np.arange(steps) * dt
        out.append(dict(
            name='sweep%d' % i, model='Acoustic2D' if i % 2 == 0 else 'Acoustic3DAxi',
            kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx, t_samples=steps,
                        t_delta=dt),
            material=('AcousticMaterial', dict(scenes.WATER)), regions=regions, init={},
            boundaries={'pressure': [dict(indices=[nx // 2 + nx * (ny // 2)],
                                          value=np.sin(2 * np.pi * 1e6 * t) * (t < 5e-6),
                                          additive=True)]},
            outputs={'pressure': [[int(v) for v in rng.choice(nx * ny, 8, replace=False)]]},
            plan=[('simulate', SWEEP_STEPS)]))
    return out


def sweep_fields(form):
    out = []
    for scene in sweep_scenes():
        fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
        fld.form = form
        out.append((scene, fld))
    return out


@pytest.fixture(scope='module')
def sweep_reference():
    """Every member's values and signals after SWEEP_STEPS steps on the oracle."""
    return [scenes.run_oracle(scene) for scene in sweep_scenes()]


def _run_sweep(form, schedule='dispatch', plan=(SWEEP_STEPS,), tile=None, strip=0):
    from pyfds_amd import autotune
    members = sweep_fields(form)
    fields = [f for _, f in members]
    if tile is None:
        tile, _, _ = autotune.tune_batch(fields, steps=10, use_cache=False)
    else:
        for f in fields:
            f.strip_height = strip
    print('sweep', form, schedule, 'tile', tile, 'strip', fields[0].strip_height, flush=True)
    batch = BatchRunner(fields, tile=tile, chunk=50, schedule=schedule)
    for n in plan:
        batch.simulate(n)
    return [scenes.collect(f, scene) for scene, f in members]


@pytest.mark.parametrize('schedule', ['dispatch', 'queue'])
def test_fp64_sweep_batch_matches_oracle(sweep_reference, schedule):
    # the queue schedule in segments of uneven length (graph chunks, fused and remainder
    # launches, asynchronous segments)
    plan = (SWEEP_STEPS,) if schedule == 'dispatch' else (7, 101, SWEEP_STEPS - 108)
    for got, ref in zip(_run_sweep('fp64', schedule, plan), sweep_reference):
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


@pytest.mark.parametrize('schedule', ['dispatch', 'queue'])
def test_fp64_sweep_batch_on_ring_shape_matches_oracle(sweep_reference, schedule):
    """The uneven sweep on the LDS-ring shape (dword-mode rows, widths 128-640), both launch
    schedules, in segments of uneven length."""
    plan = (SWEEP_STEPS,) if schedule == 'dispatch' else (7, 101, SWEEP_STEPS - 108)
    got_all = _run_sweep('fp64', schedule, plan, tile=_stream_index(14), strip=12)
    for got, ref in zip(got_all, sweep_reference):
        for k in ref:
            np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
def test_fp32_ring_batch_equals_register_batch(form):
    """FP32 forms: the ring shape's batch kernels give the register shape's bits."""
    a = _run_sweep(form, 'dispatch', (60,), tile=_stream_index(9), strip=12)
    b = _run_sweep(form, 'dispatch', (60,), tile=_stream_index(14), strip=12)
    for x, y in zip(a, b):
        for k in x:
            np.testing.assert_array_equal(x[k], y[k], err_msg=k)


def test_kahan_queue_schedule_equals_dispatch():
    a = _run_sweep('fp32_kahan', 'dispatch')
    b = _run_sweep('fp32_kahan', 'queue', (50, 150))
    for x, y in zip(a, b):
        for k in x:
            np.testing.assert_array_equal(x[k], y[k], err_msg=k)


def test_queue_schedule_needs_streaming_shape():
    members = sweep_fields('fp64')[:2]
    batch = BatchRunner([f for _, f in members], tile=1, schedule='queue')   # LDS tiles
    with pytest.raises(ValueError):
        batch.simulate(2)


def test_prefetch_shapes_serve_single_fields_only():
    members = sweep_fields('fp64')[:2]
    idx = [d['index'] for d in pyfds_amd._native.tile_table() if d['single_only']]
    assert len(idx) == 2
    for i in idx:
        with pytest.raises(ValueError):
            BatchRunner([f for _, f in members], tile=i).simulate(2)


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
def test_fp32_sweep_batch_within_tolerance_of_oracle(form, sweep_reference):
    # uneven members, graph chunks of 50, the shape the autotuner picks for this sweep
    worst = 0.0
    for scene, got, ref in zip(sweep_scenes(), _run_sweep(form), sweep_reference):
        # relative L2; a probe the wave has not reached must read exactly zero
        errs = {k: (np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])
                    if np.linalg.norm(ref[k]) > 0 else np.linalg.norm(got[k]))
                for k in ref if k != 'step'}
        print('  %-14s %4dx%-4d %s' % (scene['model'], scene['kwargs']['x_samples'],
                                       scene['kwargs']['y_samples'],
                                       '  '.join('%s %.3e' % kv for kv in errs.items())))
        worst = max([worst] + list(errs.values()))
    assert worst <= FP32_SWEEP_TOL, worst
