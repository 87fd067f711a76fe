"""Coupled-field and flow scenes, built through the public API of ``pyfds`` or ``pyfds_amd``.

* ``ta1d_s1`` / ``ta1d_s3``: ThermoAcoustic1D (acoustic losses heat the thermal field),
  coupling every step and every third step with accumulation (``pyfds/coupled_fields.py``).
* ``mc1d``: SynchronizedFields of Acoustic1D + Thermal1D with a temperature-dependent sound
  velocity (MaterialCouplingExponential every 5 steps, re-assembling the acoustic field) and a
  power-law density coupling with a relative-change threshold (``pyfds/coupling.py``).
* ``flow2d``: AcousticFlow2D with a per-row flow profile (``pyfds/acoustic_flow.py``).

Each builder returns ``(simulated object, run plan, components to collect, params)``;
``params`` feeds the fixture fingerprint.
"""
import hashlib
import json

import numpy as np


def _pulse(n, centre, width, amp):
    i = np.arange(n)
    return amp * np.exp(-0.5 * ((i - centre) / width) ** 2)


def ta1d(ns, stepping):
    n, dx, steps = 160, 1e-4, 260
    dt = 0.5 * dx / 1500.0
    params = dict(kind='ta1d', n=n, dx=dx, dt=dt, steps=steps, stepping=stepping)
    fld = ns.ThermoAcoustic1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                              thermal_material=ns.ThermalMaterial(4000.0, 1000.0, 0.6),
                              acoustic_material=ns.AcousticMaterial(1500.0, 1000.0,
                                                                    absorption_coef=2e-3),
                              stepping=stepping)
    fld.pressure.values = _pulse(n, 80, 5.0, 1e5)
    t = np.arange(steps) * dt
    fld.pressure.add_boundary(ns.regions.PointRegion([20], None, 'src'),
                              value=1e4 * np.sin(2 * np.pi * 2e6 * t), additive=True)
    fld.temperature.add_output(ns.regions.Region([30, 80, 120], 'T'))
    fld.pressure.add_output(ns.regions.Region([60], 'p'))
    comps = ['pressure', 'velocity', 'temperature', 'heat_flux']
    return fld, [('simulate', 150), ('simulate', 101)], comps, params


def mc1d(ns):
    n, dx, steps = 120, 1e-4, 220
    dt = 0.4 * dx / 1500.0
    params = dict(kind='mc1d', n=n, dx=dx, dt=dt, steps=steps)
    ac = ns.Acoustic1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                       material=ns.AcousticMaterial(1500.0, 1000.0))
    th = ns.Thermal1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                      material=ns.ThermalMaterial(4000.0, 1000.0, 0.6))
    ac.pressure.values = _pulse(n, 40, 4.0, 1.0)
    th.temperature.values = _pulse(n, 70, 12.0, 50.0)
    speed = ns.MaterialCouplingExponential(th.temperature, ac, 'sound_velocity', a=0.8, b=-0.01,
                                           stepping=5)
    dens = ns.MaterialCouplingPowerLaw(th.temperature, th, 'density', power=2, factor=1e-4,
                                       rel_change_threshold=1e-3)
    ac.velocity.add_output(ns.regions.Region([50, 90], 'v'))
    fld = ns.SynchronizedFields([ac, th], [speed, dens])
    comps = ['pressure', 'velocity', 'temperature', 'heat_flux']
    return fld, [('simulate', 60), ('simulate', 41)], comps, params


def flow2d(ns):
    nx, ny, dx, steps = 64, 40, 1e-4, 100
    dt = 0.5 * dx / 1500.0
    flow = np.linspace(300.0, 1200.0, ny)
    params = dict(kind='flow2d', nx=nx, ny=ny, dx=dx, dt=dt, steps=steps,
                  flow=[float(v) for v in flow])
    fld = ns.AcousticFlow2D(flow, x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                            t_samples=steps, t_delta=dt,
                            material=ns.AcousticMaterial(1500.0, 1000.0))
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
    fld.pressure.values = np.exp(-0.5 * (((ix - 20) ** 2 + (iy - 20) ** 2) / 16.0)).ravel()
    fld.pressure.add_output(ns.regions.Region([5 + nx * 10, 40 + nx * 30], 'p'))
    comps = ['pressure', 'velocity_x', 'velocity_y']
    return fld, [('simulate', 45), ('simulate', 44)], comps, params


SCENES = {
    'ta1d_s1': lambda ns: ta1d(ns, 1),
    'ta1d_s3': lambda ns: ta1d(ns, 3),
    'mc1d': mc1d,
    'flow2d': flow2d,
}


def run(fld, plan):
    for action, arg in plan:
        assert action == 'simulate'
        fld.simulate(arg)
    return fld


def collect(fld, comps):
    res = {'step': np.array(fld.step)}
    for c in comps:
        comp = getattr(fld, c)
        res[c] = np.array(comp.values, dtype=float)
        for k, out in enumerate(comp.outputs):
            res['%s_out%d' % (c, k)] = np.array(out.signals, dtype=float)
    return res


def fingerprint(params):
    return hashlib.sha256(json.dumps(params, sort_keys=True).encode()).hexdigest()
