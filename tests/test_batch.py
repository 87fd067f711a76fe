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


@pytest.mark.parametrize('slot', [0, 2, 3, 6, 8, 9, 11])
def test_fp64_stream_batch_matches_oracle(slot):
    """A batch of lossless members on the row-streaming engine."""
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


@pytest.mark.parametrize('slot,strip', [(9, 24), (8, 20), (9, 12), (5, 6)])
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
