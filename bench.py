"""Benchmark: Acoustic2D 2048x2048, 2000 steps, on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--forms fp64,fp32,fp32_kahan]
                    [--cpu-steps S] [--no-cpu]

Workload (one step = one full time step of the field): water background, one seeded steel
inclusion (c = 5900 m/s, rho = 7850 kg/m^3) of 128-512 cells per side, a 5-cycle 1 MHz sine
burst as an additive pressure point source at the centre, dt at CFL 0.5 on the steel sound
velocity, 64 recorded pressure points; everything is set up through the pyfds-compatible API
of pyfds_amd and timed through ``Field.simulate`` (device run + output download).

Prints one JSON line.  ``value`` is the FP64 drop-in form (bit-identical to the reference's
float64 arithmetic); the FP32 and FP32-Kahan forms are reported under ``forms`` with their
relative L2 error against the FP64 result.  ``roofline`` uses the algorithmic bytes per cell
update (FP64: p, vx, vy read + write and the two coefficients K, B = 8 words = 64 B) and the
kernel time from HIP events recorded on the library's stream around the timed launches.
``cpu_baseline`` times the numpy oracle (the plain method, single core) on a bounded
number of steps of the same workload.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

NX = NY = 2048
STEPS = 2000
HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = {'fp64': 64, 'fp32': 32, 'fp32_kahan': 56}
METRIC = 'Gcell-updates/s & % HBM roofline; FP32/batched/learned-BC vs plain FP64'


def workload(nx=NX, ny=NY, steps=STEPS, seed=1):
    """The Acoustic2D 2048^2 configuration; the source signal covers ``steps`` steps."""
    dx = 1e-4
    dt = 0.5 * dx / 5900.0
    t = np.arange(max(steps, STEPS)) * dt
    rng = np.random.default_rng(seed)
    w, h = (int(v) for v in rng.integers(128, 513, size=2))
    x0, y0 = int(rng.integers(0, nx - w)), int(rng.integers(0, ny - h))
    outs = [int(v) for v in rng.choice(nx * ny, 64, replace=False)]
    return dict(nx=nx, ny=ny, dx=dx, dt=dt, incl=(x0, y0, w, h), outs=outs,
                source=nx // 2 + nx * (ny // 2),
                signal=np.sin(2 * np.pi * 1e6 * t) * (t < 5e-6))


def build_field(ns, wl, form='fp64', t_samples=None):
    fld = ns.Acoustic2D(x_samples=wl['nx'], x_delta=wl['dx'], y_samples=wl['ny'],
                        y_delta=wl['dx'], t_samples=t_samples or STEPS, t_delta=wl['dt'],
                        material=ns.AcousticMaterial(1500.0, 1000.0))
    fld.form = form
    x0, y0, w, h = wl['incl']
    xs = np.arange(x0, x0 + w)
    ys = np.arange(y0, y0 + h)
    idx = (xs[:, None] + ys[None, :] * wl['nx']).ravel()
    fld.add_material_region(ns.regions.RectRegion(idx, wl['incl'], 'steel'),
                            ns.AcousticMaterial(5900.0, 7850.0))
    fld.pressure.add_boundary(ns.regions.PointRegion([wl['source']], None, 'source'),
                              value=wl['signal'], additive=True)
    fld.pressure.add_output(ns.regions.Region(wl['outs'], 'probes'))
    return fld


def time_form(ns, wl, form, steps, warmup):
    import torch
    fld = build_field(ns, wl, form, t_samples=steps + warmup)
    fld.assemble_matrices()
    if warmup:
        fld.simulate(warmup)
    runner = fld._runner()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fld.simulate(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes as ct
    ms = ct.c_double()
    runner.lib.fds_last_run_ms(runner.ctx, ct.byref(ms))
    cells = wl['nx'] * wl['ny']
    kernel_s = ms.value / 1e3 / steps       # per step, last run segment = whole timed run
    achieved = BYTES_PER_CELL[form] * cells / kernel_s / 1e9
    p = np.array(fld.pressure.values)
    return dict(gcells=cells * steps / wall / 1e9, wall_s=wall, ms_per_step=wall / steps * 1e3,
                kernel_us=kernel_s * 1e6, achieved_gbs=achieved, p=p)


def cpu_baseline(wl, steps):
    """The numpy oracle (plain FP64 method) on the same workload, single core."""
    import oracle
    n = wl['nx'] * wl['ny']
    x0, y0, w, h = wl['incl']
    xs = np.arange(x0, x0 + w)
    ys = np.arange(y0, y0 + h)
    incl = (xs[:, None] + ys[None, :] * wl['nx']).ravel()
    mats = [(np.arange(n), dict(sound_velocity=1500.0, density=1000.0, absorption_coef=0)),
            (incl, dict(sound_velocity=5900.0, density=7850.0, absorption_coef=0))]
    fld = oracle.PlainAcoustic2D(wl['nx'], wl['dx'], steps, wl['dt'], mats, ny=wl['ny'],
                                 dy=wl['dx'])
    fld.boundaries['pressure'] = [dict(indices=[wl['source']], value=wl['signal'],
                                       additive=True)]
    fld.outputs['pressure'] = [dict(indices=wl['outs'], signals=[])]
    fld.assemble()
    t0 = time.perf_counter()
    fld.simulate(steps)
    wall = time.perf_counter() - t0
    return n * steps / wall / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=STEPS)
    ap.add_argument('--warmup', type=int, default=20)
    ap.add_argument('--forms', default='fp64,fp32,fp32_kahan')
    ap.add_argument('--cpu-steps', type=int, default=20)
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py measures one GPU only: this build has no multi-GPU path '
              '(run with --gpus 1)', file=sys.stderr)
        sys.exit(2)

    import pyfds_amd as ns
    wl = workload(steps=args.steps + args.warmup)
    forms = [f for f in args.forms.split(',') if f]
    res = {f: time_form(ns, wl, f, args.steps, args.warmup) for f in forms}
    main_form = 'fp64' if 'fp64' in res else forms[0]
    ref = res[main_form]
    out_forms = {}
    for f, r in res.items():
        entry = dict(value=round(r['gcells'], 3), ms_per_step=round(r['ms_per_step'], 4),
                     kernel_us=round(r['kernel_us'], 2),
                     roofline_frac=round(r['achieved_gbs'] / HBM_PEAK_GBS, 4),
                     bytes_per_cell=BYTES_PER_CELL[f])
        if f != main_form and main_form == 'fp64':
            den = np.linalg.norm(ref['p'])
            entry['rel_l2_vs_fp64'] = float(np.linalg.norm(r['p'] - ref['p']) / den)
        out_forms[f] = entry
    cpu = None
    if not args.no_cpu:
        cpu_val = cpu_baseline(wl, args.cpu_steps)
        cpu = dict(value=round(cpu_val, 5), unit='Gcell-updates/s', cores=1, kind='port',
                   sample='numpy oracle (plain FP64 method), Acoustic2D 2048x2048, same scene, '
                          '%d steps (per-step cost is constant)' % args.cpu_steps)
    line = {
        'metric': METRIC,
        'value': round(ref['gcells'], 3),
        'unit': 'Gcell-updates/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': round(ref['ms_per_step'], 4),
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f64' if main_form == 'fp64' else 'f32',
        'data': 'synthetic',
        'config': {'workload': 'Acoustic2D 2048x2048 %s drop-in, steel inclusion, 1 MHz '
                               'burst source, 64 recorded points' % main_form.upper(),
                   'nx': NX, 'ny': NY, 'form': main_form, 'recorded_points': 64},
        'roofline': {'bound': 'hbm', 'achieved': round(ref['achieved_gbs'], 1),
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': round(ref['achieved_gbs'] / HBM_PEAK_GBS, 4),
                     'traffic': None},
        'cpu_baseline': cpu,
        'forms': out_forms,
    }
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
