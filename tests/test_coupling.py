"""Coupled fields and acoustic flow (pyfds/coupling.py, coupled_fields.py, acoustic_flow.py).

CPU: the reference's own coupling tests (test/test_coupling.py) restated on pyfds_amd, plus the
flow profile rules.  GPU: coupled runs and the flow model against fixtures written by running
the reference (tests/golden/coupled_*.npz), bit for bit in FP64.
"""
import os
import warnings

import numpy as np
import pytest

import pyfds_amd as fds
import coupled_scenes as cs

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')
# scenes whose device loop evaluates exp: the device's exp is correctly rounded, numpy's (which
# wrote the fixture) is within an ulp or two of it, so those scenes hold to 1e-12, not bitwise
EXP_SCENES = ('mc1d',)


def _thermal1d(**kw):
    return fds.Thermal1D(t_delta=1, t_samples=1, x_delta=1, x_samples=3,
                         material=fds.ThermalMaterial(1, 1, 1), **kw)


# ------------------------------------------------------------------ CPU
def test_synchronized_fields_exposes_components():
    acs = fds.Acoustic1D(t_delta=1, t_samples=1, x_delta=1, x_samples=3,
                         material=fds.AcousticMaterial(400, 1))
    ths = _thermal1d()
    with pytest.raises(RuntimeError):
        fds.SynchronizedFields([acs, acs], [])
    cpl = fds.SynchronizedFields([acs, ths], [])
    for comp in (acs.velocity, acs.pressure, ths.temperature, ths.heat_flux):
        assert comp in vars(cpl).values()
    assert len(cpl.material_regions) == 2
    cpl.assemble_matrices()
    assert acs.matrices_assembled and ths.matrices_assembled and cpl.matrices_assembled
    cpl.step = 4
    assert acs.step == 4 and ths.step == 4 and cpl.num_points == 3


def test_boundary_coupling_add_assign_accumulate():
    comp1 = fds.fields.FieldComponent(num_points=12)
    comp1.values = 2 * np.ones(12)
    comp2 = fds.fields.FieldComponent(num_points=12)
    comp2.values = np.ones(12)
    fds.BoundaryCoupling(comp1, comp2, lambda x: x ** 2).apply(0)
    assert np.allclose(comp2.values, 5)
    # assign every 3rd step what accumulated since: steps 1, 2, 3 -> 3 * 2 at step 3, then 2
    accu = fds.BoundaryCoupling(comp1, comp2, lambda x: x, False, True, 3)
    for step in (1, 2, 3, 4):
        accu.apply(step)
    assert np.allclose(comp2.values, 6)
    assert np.allclose(accu.accumulated_transfer, 2)


def test_material_coupling_reassembles_with_stepping_and_threshold():
    base = np.array([[-1, 1, 0], [0, -1, 1], [0, 0, -1]])
    ths1 = _thermal1d()
    ths1.assemble_matrices()
    assert np.allclose(ths1.a_t_q.toarray(), base)
    comp1 = fds.fields.FieldComponent(num_points=3)
    comp1.values = np.ones(3)
    c = fds.MaterialCoupling(comp1, ths1, 'density', lambda x: 1 + x ** 2, stepping=2)
    assert np.allclose(ths1.material_vector('density'), 2)
    assert np.allclose(ths1.material_vector('thermal_conductivity_x'), 1)
    c.apply(0)
    assert np.allclose(ths1.a_t_q.toarray(), base / 2)
    comp1.values = 2 * np.ones(3)
    c.apply(1)                       # not a coupling step
    assert np.allclose(ths1.a_t_q.toarray(), base / 2)
    c.apply(2)
    assert np.allclose(ths1.a_t_q.toarray(), base / 5)

    comp1.values = np.zeros(3)
    ths2 = _thermal1d()
    t = fds.MaterialCoupling(comp1, ths2, 'density', lambda x: 1 + x ** 2,
                             rel_change_threshold=0.4)
    t.apply(0)
    assert np.allclose(ths2.a_t_q.toarray(), base)
    comp1.values = np.ones(3)
    t.apply(1)                       # change 0.5 > 0.4
    assert np.allclose(ths2.a_t_q.toarray(), base / 2)
    comp1.values = 1.5 * np.ones(3)
    t.apply(2)                       # change (3.25 - 2) / 3.25 = 0.38 < 0.4
    assert np.allclose(ths2.a_t_q.toarray(), base / 2)


def test_exponential_and_power_law_factors():
    src = fds.fields.FieldComponent(num_points=4)
    src.values = np.array([0.0, 1.0, -2.0, 3.0])
    e = fds.MaterialCouplingExponential(src, _thermal1d(), 'density', a=0.25, b=0.5)
    assert np.allclose(e.transfer_function(src.values), 0.25 + 0.75 * np.exp(0.5 * src.values))
    assert e.transfer_function(np.zeros(1))[0] == 1.0          # p for q = 0
    p = fds.MaterialCouplingPowerLaw(src, _thermal1d(), 'density', power=3, factor=0.1)
    assert np.allclose(p.transfer_function(src.values), 1 + 0.1 * src.values ** 3)


def test_flow_profile_rules():
    kw = dict(x_samples=8, x_delta=1e-3, y_samples=4, y_delta=1e-3, t_samples=4,
              t_delta=1e-6, material=fds.AcousticMaterial(1500, 1000))
    with pytest.raises(ValueError):
        fds.AcousticFlow2D([1.0, 2.0], **kw)
    f = fds.AcousticFlow2D(100.0, **kw)
    assert np.array_equal(f.flow, np.full(4, 100.0))
    assert np.array_equal(f.flow_t_deltas, np.full(4, 10))       # 1e-3 / 100 / 1e-6
    with warnings.catch_warnings(record=True) as w:
        warnings.simplefilter('always')
        fds.AcousticFlow2D([100.0, 200.0, 1000.0, 2000.0], **kw)  # deltas 10, 5, 1, 0
    assert any('Flow velocity may be to high' in str(x.message) for x in w)


def test_apply_flow_shifts_rows_on_the_host():
    nx, ny = 6, 3
    f = fds.AcousticFlow2D([250.0, 500.0, 1000.0], x_samples=nx, x_delta=1e-3, y_samples=ny,
                           y_delta=1e-3, t_samples=4, t_delta=1e-6,
                           material=fds.AcousticMaterial(1500, 1000))
    assert list(f.flow_t_deltas) == [4, 2, 1]
    vals = np.arange(nx * ny, dtype=float) + 1
    for c in (f.pressure, f.velocity_x, f.velocity_y):
        c.values = vals.copy()
    f.step = 2                       # rows with period 2 and 1 shift, row 0 (period 4) not
    f.apply_flow()
    want = vals.reshape(ny, nx).copy()
    for r in (1, 2):
        want[r, 1:] = vals.reshape(ny, nx)[r, :-1]
        want[r, 0] = 0
    for c in (f.pressure, f.velocity_x, f.velocity_y):
        assert np.array_equal(c.values, want.ravel())


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize('name', sorted(cs.SCENES))
def test_coupled_runs_match_reference_golden(name):
    gold = np.load(os.path.join(GOLDEN, 'coupled_' + name + '.npz'))
    fld, plan, comps, params = cs.SCENES[name](fds)
    assert str(gold['fingerprint']) == cs.fingerprint(params), 'stale fixture'
    print('scene', name, flush=True)
    res = cs.collect(cs.run(fld, plan), comps)
    keys = [k for k in gold.files if k != 'fingerprint']
    assert sorted(keys) == sorted(res)
    for k in keys:
        if name in EXP_SCENES:
            np.testing.assert_allclose(res[k], gold[k], rtol=0, atol=1e-12 * np.abs(
                gold[k]).max(), err_msg='%s %s' % (name, k))
        else:
            np.testing.assert_array_equal(res[k], gold[k], err_msg='%s %s' % (name, k))


@pytest.mark.gpu
def test_material_coupling_host_protocol_matches_reference_golden_bitwise(monkeypatch):
    """The same scene through the reference's step-by-step host protocol (numpy's exp on the
    host, as the fixture was written) is bit for bit the fixture."""
    monkeypatch.setattr(fds.coupling, '_device_plan', lambda sync: None)
    gold = np.load(os.path.join(GOLDEN, 'coupled_mc1d.npz'))
    fld, plan, comps, params = cs.SCENES['mc1d'](fds)
    res = cs.collect(cs.run(fld, plan), comps)
    for k in res:
        np.testing.assert_array_equal(res[k], gold[k], err_msg=k)


@pytest.mark.gpu
@pytest.mark.parametrize('chunk', [0, 1, 7])
def test_flow_shift_on_device_for_any_graph_chunk(chunk):
    """The device shift runs inside replayed graph chunks and after plain launches alike."""
    gold = np.load(os.path.join(GOLDEN, 'coupled_flow2d.npz'))
    fld, plan, comps, params = cs.SCENES['flow2d'](fds)
    fld.graph_chunk = chunk
    res = cs.collect(cs.run(fld, plan), comps)
    for k in res:
        np.testing.assert_array_equal(res[k], gold[k], err_msg='%d %s' % (chunk, k))


@pytest.mark.gpu
def test_flow_sim_step_equals_simulate():
    gold = np.load(os.path.join(GOLDEN, 'coupled_flow2d.npz'))
    fld, plan, comps, params = cs.SCENES['flow2d'](fds)
    fld.assemble_matrices()
    for _ in range(sum(n for _, n in plan)):
        fld.sim_step()
        fld.step += 1
    res = cs.collect(fld, comps)
    for k in res:
        np.testing.assert_array_equal(res[k], gold[k], err_msg=k)


# ------------------------------------------------------------------ device-coupled loop
def _host_protocol(fld):
    """Force the reference's step-by-step host protocol on this object."""
    fld._device_coupling = lambda: None
    return fld


def _ta1d_big(ns, n=2000, steps=2300, stepping=4):
    dx = 1e-4
    dt = 0.5 * dx / 1500.0
    fld = ns.ThermoAcoustic1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                              thermal_material=ns.ThermalMaterial(4000.0, 1000.0, 0.6),
                              acoustic_material=ns.AcousticMaterial(1500.0, 1000.0,
                                                                    absorption_coef=3e-3),
                              stepping=stepping)
    i = np.arange(n)
    fld.pressure.values = 1e5 * np.exp(-0.5 * ((i - n / 3) / 15.0) ** 2)
    t = np.arange(steps) * dt
    fld.pressure.add_boundary(ns.regions.PointRegion([n // 2], None, 'src'),
                              value=1e4 * np.sin(2 * np.pi * 2e6 * t), additive=True)
    fld.velocity.add_boundary(ns.regions.PointRegion([0], None, 'wall'), value=0, additive=False)
    fld.temperature.add_output(ns.regions.Region([n // 3, n // 2], 'T'))
    fld.velocity.add_output(ns.regions.Region([n // 2 + 1], 'v'))
    return fld


@pytest.mark.gpu
@pytest.mark.parametrize('name', ['ta1d_s1', 'ta1d_s3'])
def test_thermoacoustic_runs_device_coupled(name):
    fld, plan, comps, params = cs.SCENES[name](fds)
    assert fld._device_coupling() is not None
    fld.simulate(5)
    assert fds._native.engine_of(fld.fields[0]._runner().ctx) == 'line'


@pytest.mark.gpu
@pytest.mark.parametrize('n', [400, 2000])
@pytest.mark.parametrize('stepping', [1, 4])
def test_device_coupled_equals_host_protocol_bitwise(stepping, n):
    """400 cells (one workgroup) and 2000 cells (segments), 2300 steps in runs that cross the
    2048-step launch limit, with a sim_step and an accumulation left open between them."""
    res = []
    for host in (False, True):
        fld = _ta1d_big(fds, n=n, stepping=stepping)
        if host:
            _host_protocol(fld)
        print('stepping', stepping, 'host' if host else 'device', flush=True)
        fld.simulate(2050)
        fld.sim_step()
        fld.step += 1
        fld.simulate(2300 - 2051 - 5)
        fld.simulate(5)
        r = cs.collect(fld, ['pressure', 'velocity', 'temperature', 'heat_flux'])
        r['acc'] = np.broadcast_to(np.asarray(fld.interactions[0].accumulated_transfer,
                                              dtype=float), (n,)).copy()
        res.append(r)
    for k in res[1]:
        np.testing.assert_array_equal(res[0][k], res[1][k], err_msg=k)


@pytest.mark.gpu
def test_device_coupled_signal_end_raises_like_host():
    out = []
    for host in (False, True):
        fld = _ta1d_big(fds, n=300, steps=100, stepping=3)
        if host:
            _host_protocol(fld)
        with pytest.raises(IndexError):
            fld.simulate(130)        # the source signal has 100 samples
        out.append((fld.step, np.array(fld.pressure.values), np.array(fld.temperature.values),
                    np.array(fld.temperature.outputs[0].signals)))
    assert out[0][0] == out[1][0] == 100
    for a, b in zip(out[0][1:], out[1][1:]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize('form', ['fp32', 'fp32_kahan'])
def test_device_coupled_fp32_forms(form):
    gold = np.load(os.path.join(GOLDEN, 'coupled_ta1d_s3.npz'))
    fld, plan, comps, params = cs.SCENES['ta1d_s3'](fds)
    for f in fld.fields:
        f.form = form
    res = cs.collect(cs.run(fld, plan), comps)
    for k in res:
        if k == 'step':
            continue
        ref = gold[k]
        err = np.linalg.norm(res[k] - ref) / max(np.linalg.norm(ref), 1e-300)
        print('  %-24s rel L2 %.3e' % (k, err))
        assert err <= 1e-4, (form, k, err)


# ------------------------------------------------------------------ device material coupling
def _mc_scene(ns, n=300, steps=400, threshold=1e-3, power=2.0, boundary=True):
    """Acoustic1D + Thermal1D: power-law sound velocity (every 4th step), power-law density
    with a relative-change threshold, and a linear boundary transfer from the acoustic
    pressure into the temperature, accumulated over 3 steps.  No exp: device and host protocol
    evaluate the same operations."""
    dx = 1e-4
    dt = 0.4 * dx / 1600.0
    ac = ns.Acoustic1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                       material=ns.AcousticMaterial(1500.0, 1000.0, absorption_coef=1e-3))
    th = ns.Thermal1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                      material=ns.ThermalMaterial(4000.0, 1000.0, 0.6))
    i = np.arange(n)
    ac.pressure.values = np.exp(-0.5 * ((i - n / 3) / 4.0) ** 2)
    th.temperature.values = 20.0 + 50.0 * np.exp(-0.5 * ((i - n / 2) / 20.0) ** 2)
    t = np.arange(steps) * dt
    ac.pressure.add_boundary(ns.regions.PointRegion([n // 4], None, 'src'),
                             value=0.1 * np.sin(2 * np.pi * 1e6 * t), additive=True)
    # the sound velocity grows by at most 30 % (the step stays stable for any power)
    inter = [ns.MaterialCouplingPowerLaw(th.temperature, ac, 'sound_velocity', power=power,
                                         factor=0.3 / 70.0 ** power, stepping=4),
             ns.MaterialCouplingPowerLaw(th.temperature, th, 'density', power=2, factor=1e-4,
                                         rel_change_threshold=threshold)]
    if boundary:
        inter.append(ns.BoundaryCoupling(ac.pressure, th.temperature,
                                         ns.LinearTransfer(1e-3), additive=True,
                                         accumulate=True, stepping=3))
    ac.velocity.add_output(ns.regions.Region([n // 2, n // 2 + 7], 'v'))
    th.temperature.add_output(ns.regions.Region([n // 2], 'T'))
    return ns.SynchronizedFields([ac, th], inter)


def test_device_plan_covers_the_shipped_transfer_functions():
    sync = _mc_scene(fds)
    plan = fds.coupling._device_plan(sync)
    assert [d['kind'] for d in plan] == [2, 2, 3]
    assert [d['target'] for d in plan] == [0, 1, 1]
    # a user transfer function keeps the host protocol
    sync = _mc_scene(fds, boundary=False)
    sync.interactions.append(fds.BoundaryCoupling(sync.fields[0].pressure,
                                                  sync.fields[1].temperature,
                                                  lambda v: 2 * v))
    assert fds.coupling._device_plan(sync) is None
    # the Kahan form keeps the host protocol
    sync = _mc_scene(fds)
    for f in sync.fields:
        f.form = 'fp32_kahan'
    assert fds.coupling._device_plan(sync) is None
    # two material couplings of one target keep the host protocol
    sync = _mc_scene(fds, boundary=False)
    sync.interactions.append(fds.MaterialCouplingExponential(
        sync.fields[1].temperature, sync.fields[0], 'density', a=0.5, b=0.1))
    assert fds.coupling._device_plan(sync) is None


def test_linear_transfer_is_a_scalar_multiple():
    v = np.array([1.0, -2.0, 3.5])
    np.testing.assert_array_equal(fds.LinearTransfer(0.25)(v), 0.25 * v)


def _run_mc(monkeypatch, host, plan, caplog, **kw):
    import logging
    with monkeypatch.context() as m:
        if host:
            m.setattr(fds.coupling, '_device_plan', lambda sync: None)
        sync = _mc_scene(fds, **kw)
        caplog.clear()
        with caplog.at_level(logging.INFO, logger='pyfds'):
            for n in plan:
                sync.simulate(n)
        logs = [r.getMessage() for r in caplog.records
                if r.getMessage().startswith(('Relative', 'Matrices reassembled'))]
    ac, th = sync.fields
    res = {'step': np.array(sync.step)}
    for name, c in [('p', ac.pressure), ('v', ac.velocity), ('T', th.temperature),
                    ('q', th.heat_flux)]:
        res[name] = np.array(c.values)
        for k, o in enumerate(c.outputs):
            res['%s_out%d' % (name, k)] = np.array(o.signals)
    for j, it in enumerate(sync.interactions):
        st = getattr(it, 'last_used_factors', None)
        if st is None:
            st = it.accumulated_transfer
        res['state%d' % j] = np.broadcast_to(np.asarray(st, float), (ac.num_points,)).copy()
    # the re-assembled coefficients, as the host's matrices see them
    res['a_p_v'] = ac.a_p_v.toarray()
    res['a_T_q'] = np.asarray(th._mats['g'], float)
    return res, logs


@pytest.mark.gpu
@pytest.mark.parametrize('plan', [(400,), (101, 57, 3, 239)])
def test_device_material_coupling_equals_host_protocol_bitwise(monkeypatch, caplog, plan):
    """The device loop against the reference's host protocol, over runs that split the
    coupling periods (101, 57, 3 steps), with the reassembly log lines of the threshold."""
    dev, dev_log = _run_mc(monkeypatch, False, plan, caplog)
    host, host_log = _run_mc(monkeypatch, True, plan, caplog)
    print('reassemblies logged: device %d host %d' % (len(dev_log), len(host_log)))
    assert sorted(dev) == sorted(host)
    for k in host:
        assert np.all(np.isfinite(host[k])), k
        np.testing.assert_array_equal(dev[k], host[k], err_msg=k)
    assert dev_log == host_log and len(host_log) > 0


@pytest.mark.gpu
@pytest.mark.parametrize('power', [0.5, -1.0, 1.0, 3.0])
def test_device_power_law_paths(monkeypatch, caplog, power):
    """numpy's ``values ** power`` takes fast paths for 0.5, -1, 1 (and 2, above); other
    exponents run the correctly rounded power on the device and numpy's on the host (within
    an ulp), so those hold to 1e-12."""
    dev, _ = _run_mc(monkeypatch, False, (150,), caplog, power=power, threshold=None)
    host, _ = _run_mc(monkeypatch, True, (150,), caplog, power=power, threshold=None)
    for k in host:
        assert np.all(np.isfinite(host[k])), k
        if power == 3.0:
            np.testing.assert_allclose(dev[k], host[k], rtol=0,
                                       atol=1e-12 * max(np.abs(host[k]).max(), 1e-300),
                                       err_msg=k)
        else:
            np.testing.assert_array_equal(dev[k], host[k], err_msg=k)
