"""Asynchronous runs.  ``simulate`` cuts a long run with recorded outputs into asynchronous
segments (``fds_run_async`` / ``fds_run_take``, ``fds_batch_run_async`` / ``fds_batch_take``)
so that the host builds the per-sample output lists of one segment while the device computes
the next.  The results equal the oracle and one synchronous run bit for bit, and the C-ABI
refuses the calls that would race with a run in flight."""
import ctypes as ct

import numpy as np
import pytest

import pyfds_amd
import scenes
from pyfds_amd import _native as nat
from pyfds_amd import device
from pyfds_amd.batch import BatchRunner

pytestmark = pytest.mark.gpu

PIPELINED = ['a2d_main', 'axi_lossy', 't2d_blocks', 'taxi', 'bounds_semantics', 'a1d_rigid',
             't1d', 'ig1d_cnl']


def _build(scene, form='fp64', chunk=5):
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    fld.form = form
    fld.graph_chunk = chunk
    return fld


def _scene(name):
    return scenes.SCENES[name]()


@pytest.mark.parametrize('name', PIPELINED)
def test_pipelined_run_matches_oracle(name, monkeypatch):
    scene = _scene(name)
    if scene['plan'] != [('simulate', None)]:
        scene = dict(scene, plan=[('simulate', None)])
    n = scene['kwargs']['t_samples']
    print('pipelined scene', name, 'steps', n, flush=True)
    m = sum(len(i) for ol in scene['outputs'].values() for i in ol)
    assert device.pipeline_sizes(n, n, 5, m), 'the run must take the asynchronous path'
    res = []
    for segments in (device.PIPELINE_SEGMENTS, 1):
        monkeypatch.setattr(device, 'PIPELINE_SEGMENTS', segments)
        fld = _build(scene)
        scenes.run_plan(fld, scene)
        res.append(scenes.collect(fld, scene))
        fld._runner().close()
    got, sync = res
    ref = scenes.run_oracle(scene, power='cr')
    for k in ref:
        np.testing.assert_array_equal(got[k], sync[k], err_msg='%s %s' % (name, k))
        np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (name, k))


@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
def test_pipelined_equals_synchronous(form, monkeypatch):
    scene = dict(_scene('a2d_main'), plan=[('simulate', 37), ('simulate', 113)])
    res = []
    for segments in (device.PIPELINE_SEGMENTS, 1):
        monkeypatch.setattr(device, 'PIPELINE_SEGMENTS', segments)
        fld = _build(scene, form)
        scenes.run_plan(fld, scene)
        res.append(scenes.collect(fld, scene))
        fld._runner().close()
    for k in res[0]:
        np.testing.assert_array_equal(res[0][k], res[1][k], err_msg='%s %s' % (form, k))


def test_pipelined_run_stops_where_the_signal_ends():
    """A boundary signal of 150 samples and a 200-step request: the asynchronous segments cover
    the 150 valid steps, then IndexError as the reference's signal[step] raises."""
    scene = _scene('a2d_main')
    fld = _build(scene)
    with pytest.raises(IndexError):
        fld.simulate(200)
    assert fld.step == 150
    ref = scenes.run_oracle(scene)
    got = scenes.collect(fld, scene)
    for k in ref:
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    fld._runner().close()


def test_async_abi_rules():
    """Two runs may be in flight; a third, a synchronous run, a setter and an output read are
    refused until they are taken; the taken outputs equal a synchronous run's."""
    scene = _scene('a2d_main')
    fld = _build(scene, chunk=0)
    fld.simulate(1)
    r = fld._runner()
    lib, ctx = r.lib, r.ctx
    m = [c.size for c in r.out_cells]
    assert nat.E_STATE == lib.fds_run_take(ctx, None, None, None)
    assert lib.fds_run_async(ctx, 1, 10, 0) == 0
    assert lib.fds_run_async(ctx, 11, 20, 4) == 0
    assert lib.fds_run_async(ctx, 31, 5, 0) == nat.E_STATE
    assert lib.fds_run(ctx, 31, 5, 0) == nat.E_STATE
    assert lib.fds_set_tile(ctx, 0) == nat.E_STATE
    buf = np.empty((10, m[0]))
    assert lib.fds_read_outputs(ctx, 0, 0, 10, nat.dptr(buf)) == nat.E_STATE
    taken = []
    for rows in (10, 20):
        d = [np.empty((rows, k)) for k in m]
        assert lib.fds_run_take(ctx, *[nat.dptr(x) for x in d]) == 0
        taken.append(d)
    assert lib.fds_run_take(ctx, None, None, None) == nat.E_STATE
    ms = ct.c_double()
    assert lib.fds_last_run_ms(ctx, ct.byref(ms)) == 0 and ms.value > 0
    final = np.empty(fld.num_points)
    assert lib.fds_get_state(ctx, 0, nat.dptr(final), fld.num_points) == 0
    r.close()

    # the oracle over the same 31 steps (step 0 ran synchronously above)
    ref = scenes.run_oracle(dict(scene, plan=[('simulate', 31)]))
    for comp, k in (('pressure', 0), ('velocity_x', 1), ('velocity_y', 2)):
        sig = np.concatenate([ref['signals/%s/%d' % (comp, j)]
                              for j in range(len(scene['outputs'][comp]))], axis=0).T
        np.testing.assert_array_equal(np.concatenate([taken[0][k], taken[1][k]]), sig[1:31],
                                      err_msg=comp)
    np.testing.assert_array_equal(final, ref['values/pressure'])


BATCH = ['a2d_main', 't2d_blocks', 'axi_lossless', 'taxi']


@pytest.mark.parametrize('form', ['fp64', 'fp32_kahan'])
def test_pipelined_batch_equals_synchronous(form, monkeypatch):
    res = []
    for segments in (device.PIPELINE_SEGMENTS, 1):
        monkeypatch.setattr(device, 'PIPELINE_SEGMENTS', segments)
        members = []
        for name in BATCH:
            scene = dict(_scene(name), plan=[('simulate', 100)])
            members.append((scene, _build(scene, form)))
        batch = BatchRunner([f for _, f in members], chunk=5)
        m = sum(len(i) for sc, _ in members for ol in sc['outputs'].values() for i in ol)
        if segments > 1:
            assert device.pipeline_sizes(100, 100, 5, m)
        batch.simulate(100)
        res.append([scenes.collect(f, s) for s, f in members])
        last_ms = batch.last_run_ms()
        assert last_ms > 0
        batch.close()
        for _, f in members:
            f._runner().close()
    for name, a, b in zip(BATCH, res[0], res[1]):
        for k in a:
            np.testing.assert_array_equal(a[k], b[k], err_msg='%s %s %s' % (form, name, k))
    if form == 'fp64':
        for name, got in zip(BATCH, res[0]):
            ref = scenes.run_oracle(dict(_scene(name), plan=[('simulate', 100)]))
            for k in ref:
                np.testing.assert_array_equal(got[k], ref[k], err_msg='%s %s' % (name, k))


def test_batch_async_abi_rules():
    """While a batch run is in flight its members refuse their own runs, setters and state
    reads (the batch runs on its own stream); taking it releases them."""
    fields = []
    for name in ('a2d_main', 'taxi'):
        fld = _build(_scene(name), chunk=0)
        fld.simulate(1)
        fields.append(fld)
    batch = BatchRunner(fields, chunk=0)
    batch.simulate(1)
    lib, handle = batch.lib, batch.handle
    ctx0 = fields[0]._runner().ctx
    assert lib.fds_batch_take(handle, None) == nat.E_STATE
    assert lib.fds_batch_run_async(handle, 2, 10, 0) == 0
    buf = np.empty(fields[0].num_points)
    assert lib.fds_get_state(ctx0, 0, nat.dptr(buf), buf.size) == nat.E_STATE
    assert lib.fds_run_async(ctx0, 12, 5, 0) == nat.E_STATE
    assert lib.fds_run(ctx0, 12, 5, 0) == nat.E_STATE
    assert lib.fds_set_tile(ctx0, 0) == nat.E_STATE
    assert lib.fds_batch_run(handle, 12, 5, 0) == nat.E_STATE
    assert lib.fds_batch_take(handle, None) == 0          # outputs discarded
    assert lib.fds_get_state(ctx0, 0, nat.dptr(buf), buf.size) == 0
    ms = ct.c_double()
    assert lib.fds_batch_last_run_ms(handle, ct.byref(ms)) == 0 and ms.value > 0
    batch.close()
    for f in fields:
        f._runner().close()


def test_prepare_run_changes_nothing():
    """fds_prepare_run (tables, output rows, chunk graphs; no steps) leaves the field as it was:
    the run that follows gives the same bits as one without it."""
    scene = _scene('a2d_main')
    res = []
    for prep in (True, False):
        fld = _build(scene)
        fld.simulate(3)
        if prep:
            fld._runner().prepare(fld, 137)
            assert fld.step == 3
        fld.simulate(137)
        res.append(scenes.collect(fld, scene))
        fld._runner().close()
    for k in res[0]:
        np.testing.assert_array_equal(res[0][k], res[1][k], err_msg=k)
