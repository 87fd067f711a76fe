"""Failure detection (host-forced conditions; every kernel still exits normally).

* ``check_finite``: a NaN or infinity in the state -- written into ``values`` before the run,
  or grown by a time step above the stability limit -- makes ``simulate`` raise
  ``FloatingPointError`` naming the component and the step by which it was found, after the
  run's steps are counted and their outputs recorded; the tile, streaming, resident 1D and
  batch paths.  Without it, the reference's behaviour: NaN propagates silently.
* ``sim`` is ``simulate``.
"""
import re

import numpy as np
import pytest

import pyfds_amd
import scenes
from pyfds_amd.batch import BatchRunner


def _a2d(scene='a2d_main', **attrs):
    sc = scenes.SCENES[scene]()
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, sc)
    for k, v in attrs.items():
        setattr(fld, k, v)
    return fld


def test_sim_is_simulate(monkeypatch):
    calls = []
    monkeypatch.setattr(pyfds_amd.Acoustic2D, 'simulate', lambda self, n=None: calls.append(n))
    f = _a2d()
    f.sim(7)
    f.sim()
    assert calls == [7, None]


@pytest.mark.gpu
def test_sim_runs_like_simulate():
    a, b = _a2d(), _a2d()
    a.simulate(37)
    b.sim(37)
    np.testing.assert_array_equal(a.pressure.values, b.pressure.values)
    assert a.step == b.step == 37


def _found(exc):
    m = re.search(r'component (\d+), found after step (\d+)', str(exc.value))
    assert m, str(exc.value)
    return int(m.group(1)), int(m.group(2))


@pytest.mark.gpu
@pytest.mark.parametrize('engine', ['stream', 'tile'])
def test_nan_in_values_is_reported_per_chunk(engine):
    f = _a2d(check_finite=True, graph_chunk=10, **({'tile': 1} if engine == 'tile' else {}))
    p = np.zeros(f.num_points)
    p[f.num_points // 3] = np.nan
    f.pressure.values = p
    with pytest.raises(FloatingPointError) as exc:
        f.simulate(30)                      # one synchronous run of graph chunks
    comp, step = _found(exc)
    runner = f._runner()
    g = runner.lib.fds_steps_per_launch(runner.ctx)
    assert comp == 0 and step == 10 - 10 % g     # the first chunk's scan
    assert f.step == 30                     # the run completed and is counted
    assert len(f.pressure.outputs[0].signals[0]) == 30
    assert np.isnan(f.pressure.values).any()


@pytest.mark.gpu
def test_without_check_nan_propagates_silently():
    f = _a2d(graph_chunk=10)
    p = np.zeros(f.num_points)
    p[5] = np.nan
    f.pressure.values = p
    f.simulate(20)
    assert np.isnan(f.pressure.values).any() and f.step == 20


@pytest.mark.gpu
def test_unstable_time_step_blows_up_and_is_reported():
    # dt 40x the CFL limit: the field overflows to infinity / NaN within the run
    sc = scenes.SCENES['a2d_main']()
    sc['kwargs'] = dict(sc['kwargs'], t_delta=sc['kwargs']['t_delta'] * 40, t_samples=400)
    f = scenes.build_field(pyfds_amd, pyfds_amd.regions, sc)
    f.check_finite = True
    f.graph_chunk = 25
    with pytest.raises(FloatingPointError) as exc:
        f.simulate(300)
    comp, step = _found(exc)
    runner = f._runner()
    g = runner.lib.fds_steps_per_launch(runner.ctx)
    assert 0 < step <= 300 and step % (25 - 25 % g) == 0     # at a graph chunk's end


@pytest.mark.gpu
def test_pipelined_run_stops_after_the_segment_that_found_it():
    sc = scenes.SCENES['a2d_main']()
    sc['kwargs'] = dict(sc['kwargs'], t_delta=sc['kwargs']['t_delta'] * 40, t_samples=2000)
    f = scenes.build_field(pyfds_amd, pyfds_amd.regions, sc)
    f.check_finite = True
    f.graph_chunk = 10
    with pytest.raises(FloatingPointError):
        f.simulate(2000)
    assert 0 < f.step < 2000                 # segments after the hit were not enqueued
    assert len(f.pressure.outputs[0].signals[0]) == f.step


@pytest.mark.gpu
def test_line_engine_reports_nan():
    sc = scenes.SCENES['a1d_rigid']()
    f = scenes.build_field(pyfds_amd, pyfds_amd.regions, sc)
    f.check_finite = True
    v = np.zeros(f.num_points)
    v[10] = np.inf
    f.velocity.values = v
    with pytest.raises(FloatingPointError) as exc:
        f.simulate(120)
    comp, step = _found(exc)
    assert step <= 120 and f.step == 120


@pytest.mark.gpu
def test_batch_member_with_nan_is_named():
    a, b = _a2d(), _a2d('a2d_wrap')
    b.check_finite = True
    p = np.zeros(b.num_points)
    p[3] = np.nan
    b.pressure.values = p
    batch = BatchRunner([a, b], chunk=5)
    with pytest.raises(FloatingPointError, match='batch member 1'):
        batch.simulate(40)
    # a pipelined batch run stops enqueueing after the segment that found it
    assert a.step == b.step and 0 < a.step < 40
    assert len(a.pressure.outputs[0].signals[0]) == a.step
