"""Benchmark on one MI355X: Acoustic2D 2048x2048 (headline), batched sweep, Thermal2D 4096^2.

    python bench.py [--gpus 1] [--steps K] [--warmup W (600)] [--forms fp64,fp32,fp32_kahan]
                    [--sections acoustic2d,batched,thermal] [--cpu-steps S] [--no-cpu]
                    [--batch-steps 1000] [--thermal-steps 2000]

Workload (one step = one full time step of the field): water background, one seeded steel
inclusion (c = 5900 m/s, rho = 7850 kg/m^3) of 128-512 cells per side, a 5-cycle 1 MHz sine
burst as an additive pressure point source at the centre, dt at CFL 0.5 on the steel sound
velocity, 64 recorded pressure points; everything is set up through the pyfds-compatible API
of pyfds_amd and timed through ``Field.simulate`` (device run + output download).

Prints one JSON line.  ``value`` is the FP64 drop-in form (bit-identical to the reference's
float64 arithmetic); the FP32 and FP32-Kahan forms are reported under ``forms`` with their
relative L2 error against the FP64 result.  ``roofline`` uses the algorithmic bytes per cell
update (FP64: p, vx, vy read + write and the two coefficients K, B = 8 words = 64 B) times the
cell updates of one launch (a launch of the streaming engine advances 1-3 fused time steps), over
the launch time from HIP events recorded on the library's stream around the timed launches.
With fused steps the physical HBM traffic per launch (``traffic``, measured with rocprofv3) is
below the algorithmic bytes, so ``frac`` can pass what one pass per step could reach.
``cpu_baseline`` times the numpy oracle (the plain method, single core) on a bounded
number of steps of the same workload.

Further sections (reported under ``lossy``, ``batched`` and ``thermal2d``):

* lossy: the same Acoustic2D 2048^2 scene with absorbing materials (shear viscosity 1e-3 Pa s,
  bulk viscosity 2.4e-3 Pa s everywhere), FP64 and FP32, autotuned (lossy acoustics stream one
  step per launch); algorithmic bytes 72 / 36 B per cell update (the viscous factor M adds one
  word).

* batched: 256 Acoustic2D / Acoustic3DAxi instances, sides uniform in 128..1024 (seed 2),
  1-4 seeded inclusions each (c 1000-6000 m/s, rho 800-8000 kg/m^3), dt at CFL 0.5 on each
  instance's own c_max, a 1 MHz burst source at each centre and 16 recorded points, FP32 and
  FP32-Kahan, one launch per step over all instances' tiles in HIP-graph chunks of 50; each
  form's pressure after the run against the same sweep in FP64 (the oracle's bits):
  relative L2 per member (largest, median) and over all cells (``rel_l2_vs_fp64``).
* learned_bc: Acoustic2D 1024x1024 FP32 interior + 16-cell learned layer (BF16-MFMA conv net,
  61 440 parameters) trained on plain 3072x3072 FP32 runs with random Ricker sources and
  inclusions (seeds 100..), judged on a held-out scene against the plain 3072^2 run:
  reflected energy, the same small grid without the net (ablation), wall-clock of both methods.
* acoustic1d: Acoustic1D, 2000 cells, water, dx 1e-4, dt at CFL 0.5, Gaussian pulse, rigid
  ends (velocity Dirichlet 0 at both ends), 3 recorded pressure points, 5000 steps; FP64 on
  the resident 1D engine (its default layout of segments, and the other layouts) and on the
  LDS-tile engine (one launch per step in HIP-graph chunks of 50).  Latency-bound: reported
  as time per step.  Also ThermoAcoustic1D on the same line (acoustic losses heating a
  Thermal1D field every 4th step, accumulated in between): the device-coupled loop (one
  workgroup steps both fields and the transfer) against the reference's host protocol over the
  same device fields.
* thermal2d: Thermal2D 4096x4096, k in {0.6, 50} W/mK in seeded 256-cell blocks (seed 3),
  (16x16), rho c_p identical, Fourier number 0.2 on k_max, Gaussian hot spot (sigma 64 cells); FP64
  and FP32-Kahan over a bounded number of the 100 000 steps (per-step cost is constant), with
  the Kahan result's relative L2 error against FP64.
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
LOSSY_BYTES_PER_CELL = {'fp64': 72, 'fp32': 36, 'fp32_kahan': 60}
THERMAL_BYTES_PER_CELL = {'fp64': 32, 'fp32': 16, 'fp32_kahan': 24}
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


def make_lossy(ns, fld):
    for mr in fld.material_regions:
        for m in mr.materials:
            m.shear_viscosity, m.bulk_viscosity = 1e-3, 2.4e-3
    return fld


def time_lossy(ns, wl, form, steps, warmup, fixed=None):
    fld = make_lossy(ns, build_field(ns, wl, form, t_samples=steps + warmup + 50))
    fld.assemble_matrices()
    tune = tuned(fld, fixed)
    runner = fld._runner()
    runner.prepare(fld, steps)      # one-time set-up (chunk graphs) outside the timed steps
    fld.simulate(warmup)            # then the warm-up, right before the timed steps
    t0 = time.perf_counter()
    fld.simulate(steps)
    wall = time.perf_counter() - t0
    import ctypes as ct
    ms = ct.c_double()
    runner.lib.fds_last_run_ms(runner.ctx, ct.byref(ms))
    cells = wl['nx'] * wl['ny']
    kernel_s = ms.value / 1e3 / steps
    achieved = LOSSY_BYTES_PER_CELL[form] * cells / kernel_s / 1e9
    return dict(value=round(cells * steps / wall / 1e9, 3), ms_per_step=round(wall / steps * 1e3, 4),
                kernel_us=round(kernel_s * 1e6, 2), achieved_gbs=round(achieved, 1),
                roofline_frac=round(achieved / HBM_PEAK_GBS, 4),
                bytes_per_cell=LOSSY_BYTES_PER_CELL[form], autotune=tune)


def tuned(fld, fixed=None):
    """Pick the shape / strip height / graph chunk by timing (state restored afterwards), or
    use ``fixed`` = (tile, strip)."""
    from pyfds_amd import autotune, _native as nat
    if fixed is not None:
        fld.tile, fld.strip_height, fld.graph_chunk = fixed[0], fixed[1], 50
        tile, chunk, table = fixed[0], 50, None
    else:
        tile, chunk, table = autotune.tune_field(fld, steps=40, use_cache=False)
        table = {k: round(v, 2) for k, v in table.items()}
    info = nat.tile_table()[tile]
    # steps per launch as the context runs them (lossy fields: one, whatever the shape)
    runner = fld._runner()
    runner.sync(fld)
    spl = nat.check(runner.lib.fds_steps_per_launch(runner.ctx))
    return dict(tile=tile, strip=getattr(fld, 'strip_height', 0) or 0, chunk=chunk,
                shape='%s %dx%d, %d step(s) per launch%s' % (
                    info['engine'], info['tx'], getattr(fld, 'strip_height', 0) or info['ty'],
                    spl, ', non-temporal' if info['nontemporal'] else '')
                + (', prefetch %d rows' % info['prefetch'] if info['prefetch'] > 1 else '')
                + (' via LDS rings' if info['lds_ring'] else ''),
                steps_per_launch=spl, table=table)


COPY_CAL_BYTES = 1 << 30      # plain copy used to scale the PMC byte counters
COPY_CAL_VARIANT = 0          # fds_stream_copy_bench variant (grid-stride copy)


def _pmc_rows(outdir):
    """(kernel name, counter value) of every dispatch in a rocprofv3 counter CSV."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(outdir, '**', '*counter_collection.csv'), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append((r.get('Kernel_Name', ''), float(r.get('Counter_Value', 'nan'))))
    return rows


def measure_traffic(kind, form, tile, strip, launches=20):
    """HBM bytes per launch of the step kernel of ``kind`` ('acoustic2d' or 'thermal') at the
    shape ``tile`` / ``strip``: two rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE) of this
    script's --traffic-probe mode, which runs ``launches`` full launches after a 1 GiB plain
    copy; each counter is scaled by what it reads for that copy (known bytes).  Returns a dict
    (bytes per launch, the raw KiB and the scale factors) or None when rocprofv3 is missing or
    a run fails."""
    import shutil
    import subprocess
    import tempfile
    prof = shutil.which('rocprofv3') or '/opt/rocm/bin/rocprofv3'
    if not os.path.exists(prof):
        return None
    family = {'acoustic2d': 'k_stream_acoustic<', 'thermal': 'k_stream_thermal<'}[kind]
    out = {}
    for counter in ('FETCH_SIZE', 'WRITE_SIZE'):
        d = tempfile.mkdtemp(prefix='pyfds_pmc_', dir='/tmp')
        cmd = ['timeout', '-s', 'KILL', '150', prof, '--pmc', counter, '-d', d, '-o', 'pmc',
               '--output-format', 'csv', '--', sys.executable, os.path.abspath(__file__),
               '--traffic-probe', '%s:%s:%d:%d:%d' % (kind, form, tile, strip, launches)]
        env = dict(os.environ, TMPDIR='/tmp')
        res = subprocess.run(cmd, cwd='/tmp', env=env, capture_output=True, text=True)
        rows = _pmc_rows(d)
        shutil.rmtree(d, ignore_errors=True)
        # rocprofv3 may end with a signal after the program finished: use complete results
        if 'traffic-probe done' not in res.stdout or not rows:
            print('traffic probe (%s) failed: rc %d %s' % (counter, res.returncode,
                                                          res.stderr[-400:]), file=sys.stderr)
            return None
        copy = [v for k, v in rows if 'k_stream_copy' in k]
        kern = {}
        for k, v in rows:
            if family in k:
                kern.setdefault(k, []).append(v)
        if not copy or not kern:
            return None
        name, vals = max(kern.items(), key=lambda kv: len(kv[1]))
        scale = COPY_CAL_BYTES / (float(np.median(copy)) * 1024.0)
        out[counter] = dict(kib=float(np.median(vals)), scale=round(scale, 4), kernel=name,
                            dispatches=len(vals))
        pat = [v for k, v in rows if 'k_stream_pattern<' in k]
        if pat:
            out[counter]['pattern_kib'] = float(np.median(pat))
    fetch, write = out['FETCH_SIZE'], out['WRITE_SIZE']
    read_b = fetch['kib'] * 1024 * fetch['scale']
    write_b = write['kib'] * 1024 * write['scale']
    res = dict(bytes=int(round(read_b + write_b)), read_bytes=int(round(read_b)),
               write_bytes=int(round(write_b)), fetch_kib=round(fetch['kib'], 1),
               write_kib=round(write['kib'], 1), fetch_scale=fetch['scale'],
               write_scale=write['scale'], kernel=fetch['kernel'],
               dispatches=fetch['dispatches'])
    if 'pattern_kib' in fetch and 'pattern_kib' in write:
        # the same counters for the memory-only kernel of the same access pattern
        pr = fetch['pattern_kib'] * 1024 * fetch['scale']
        pw = write['pattern_kib'] * 1024 * write['scale']
        res['pattern'] = dict(bytes=int(round(pr + pw)), read_bytes=int(round(pr)),
                              write_bytes=int(round(pw)))
    return res


def traffic_probe(spec):
    """Child of measure_traffic (runs under rocprofv3): a 1 GiB calibration copy, then
    ``launches`` full launches of the step kernel at the given shape and, for an LDS-ring
    shape, as many launches of the memory-only kernel of its access pattern."""
    import ctypes as ct
    import pyfds_amd as ns
    from pyfds_amd import _native as nat
    kind, form, tile, strip, launches = spec.split(':')
    tile, strip, launches = int(tile), int(strip), int(launches)
    lib = nat.load()
    gbs, ms = ct.c_double(), ct.c_double()
    nat.check(lib.fds_stream_copy_bench(COPY_CAL_BYTES, 2, COPY_CAL_VARIANT, ct.byref(gbs),
                                        ct.byref(ms)))
    if kind == 'acoustic2d':
        fld = build_field(ns, workload(steps=400), form, t_samples=400)
    else:
        fld = thermal_field(ns, form, 400)
    fld.tile, fld.strip_height, fld.graph_chunk = tile, strip, 50
    fld.assemble_matrices()
    runner = fld._runner()
    runner.sync(fld)
    spl = nat.check(lib.fds_steps_per_launch(runner.ctx))
    fld.simulate(spl * launches)
    info = nat.tile_table()[tile]
    if info['lds_ring'] and info['tx'] == 64 and spl in (2, 3) and form in ('fp64', 'fp32'):
        # the memory-only kernel of the same access pattern, for its own byte counters
        ms, b = ct.c_double(), ct.c_double()
        nat.check(lib.fds_stream_pattern_bench(
            fld.x.samples, fld.y.samples, 8 if form == 'fp64' else 4, spl, spl - 1,
            strip or info['ty'], 3 if kind == 'acoustic2d' else 1, launches, ct.byref(ms),
            ct.byref(b)))
    print('traffic-probe done', flush=True)


def launch_timing(fld, launches=30):
    """Average duration (us) of one full launch of the field's step kernel: HIP events on the
    library's stream around ``launches`` launches (graph-replayed, run set up beforehand)."""
    import ctypes as ct
    runner = fld._runner()
    runner.sync(fld)
    spl = runner.lib.fds_steps_per_launch(runner.ctx)
    n = spl * launches
    runner.prepare(fld, n)
    fld.simulate(n)
    ms = ct.c_double()
    runner.lib.fds_last_run_ms(runner.ctx, ct.byref(ms))
    return ms.value * 1e3 / launches, spl


def pattern_floor(tune, nx, ny, elem, n_state, launch_us):
    """For an LDS-ring shape: the time per launch of the memory-only kernel with the step
    kernel's access pattern (``fds_stream_pattern_bench``: its tiles, tile order, ring loads,
    LDS footprint and stores, no arithmetic) at the tuned strip height, and the step kernel's
    launch time over it (1 = the kernel runs at the floor its memory traffic sets)."""
    import ctypes as ct
    from pyfds_amd import _native as nat
    info = nat.tile_table()[tune['tile']]
    spl = tune['steps_per_launch']
    if not info['lds_ring'] or info['tx'] != 64 or spl not in (2, 3):
        return None
    strip = tune['strip'] or info['ty']
    lib = nat.load()
    ms, b = ct.c_double(), ct.c_double()
    nat.check(lib.fds_stream_pattern_bench(nx, ny, elem, spl, spl - 1, strip, n_state, 20,
                                           ct.byref(ms), ct.byref(b)))
    us = ms.value * 1e3
    return dict(us_per_launch=round(us, 2), requested_gbs=round(b.value / us / 1e3, 1),
                kernel_over_floor=round(launch_us / us, 3),
                geometry='64-column tiles, strips of %d, %d ghost rows/columns per side, '
                         '%d-stage rings' % (strip, spl - 1, spl))


def roofline_entry(bytes_per_update, cells, launch_us, spl, traffic):
    """The roofline object: physical HBM bytes per launch (measured) over the launch time
    against 8 TB/s; without a measurement the one-pass compulsory bytes (a lower bound of the
    physical traffic).  Detail: the compulsory and the per-step algorithmic rates."""
    compulsory = bytes_per_update * cells            # state r+w + coefficients, one pass
    algorithmic = bytes_per_update * cells * spl     # the per-step figure x cell updates
    detail = dict(launch_us=round(launch_us, 3), steps_per_launch=spl,
                  compulsory_bytes_per_launch=compulsory,
                  compulsory_gbs=round(compulsory / launch_us / 1e3, 1),
                  compulsory_frac=round(compulsory / launch_us / 1e3 / HBM_PEAK_GBS, 4),
                  algorithmic_bytes_per_launch=algorithmic,
                  algorithmic_gbs=round(algorithmic / launch_us / 1e3, 1),
                  algorithmic_frac=round(algorithmic / launch_us / 1e3 / HBM_PEAK_GBS, 4))
    if traffic is not None:
        phys = traffic['bytes'] / launch_us / 1e3
        detail.update(traffic_over_compulsory=round(traffic['bytes'] / compulsory, 3),
                      pmc=traffic)
        roof = dict(bound='hbm', achieved=round(phys, 1), peak=HBM_PEAK_GBS, unit='GB/s',
                    frac=round(phys / HBM_PEAK_GBS, 4), traffic=traffic['bytes'],
                    basis='measured HBM bytes per launch (rocprofv3 FETCH_SIZE/WRITE_SIZE, '
                          'copy-calibrated) / launch time')
    else:
        roof = dict(bound='hbm', achieved=detail['compulsory_gbs'], peak=HBM_PEAK_GBS,
                    unit='GB/s', frac=detail['compulsory_frac'], traffic=None,
                    basis='one-pass compulsory bytes / launch time (traffic not measured)')
    return roof, detail


def time_form(ns, wl, form, steps, warmup, fixed=None):
    import torch
    fld = build_field(ns, wl, form, t_samples=steps + warmup + 200)
    fld.assemble_matrices()
    tune = tuned(fld, fixed)
    runner = fld._runner()
    runner.prepare(fld, steps)      # one-time set-up (chunk graphs) outside the timed steps
    if warmup:                      # then the warm-up, right before the timed steps
        fld.simulate(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fld.simulate(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes as ct
    ms = ct.c_double()
    runner.lib.fds_last_run_ms(runner.ctx, ct.byref(ms))
    cells = wl['nx'] * wl['ny']
    kernel_s = ms.value / 1e3 / steps       # per step, summed over the run's segments
    achieved = BYTES_PER_CELL[form] * cells / kernel_s / 1e9
    p = np.array(fld.pressure.values)
    launch_us, _ = launch_timing(fld)
    return dict(gcells=cells * steps / wall / 1e9, wall_s=wall, ms_per_step=wall / steps * 1e3,
                kernel_us=kernel_s * 1e6, achieved_gbs=achieved, p=p, tune=tune,
                launch_us=launch_us)


def batched_params(n_inst=256, seed=2):
    """Per-instance parameters of the sweep: (nx, ny, axi, inclusions, probe seed)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_inst):
        nx, ny = (int(v) for v in rng.integers(128, 1025, size=2))
        incl = []
        for _ in range(int(rng.integers(1, 5))):
            w = int(rng.integers(max(2, nx // 10), max(3, nx // 3)))
            h = int(rng.integers(max(2, ny // 10), max(3, ny // 3)))
            x0, y0 = int(rng.integers(0, nx - w)), int(rng.integers(0, ny - h))
            incl.append((x0, y0, w, h, float(rng.uniform(1000, 6000)),
                         float(rng.uniform(800, 8000))))
        probes = [int(v) for v in rng.choice(nx * ny, 16, replace=False)]
        out.append((nx, ny, i % 2 == 1, incl, probes))
    return out


def batched_field(ns, form, params, steps):
    nx, ny, axi, incl, probes = params
    dx = 1e-4
    cmax = max([1500.0] + [c for *_, c, _ in incl])
    dt = 0.5 * dx / cmax
    cls = ns.Acoustic3DAxi if axi else ns.Acoustic2D
    fld = cls(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx, t_samples=steps,
              t_delta=dt, material=ns.AcousticMaterial(1500.0, 1000.0))
    fld.form = form
    for x0, y0, w, h, c, rho in incl:
        xs, ys = np.arange(x0, x0 + w), np.arange(y0, y0 + h)
        fld.add_material_region(ns.regions.RectRegion(
            (xs[:, None] + ys[None, :] * nx).ravel(), (x0, y0, w, h)),
            ns.AcousticMaterial(c, rho))
    t = np.arange(steps + 64) * dt
    fld.pressure.add_boundary(ns.regions.PointRegion([nx // 2 + nx * (ny // 2)], None),
                              value=np.sin(2 * np.pi * 1e6 * t) * (t < 5e-6), additive=True)
    fld.pressure.add_output(ns.regions.Region(probes))
    return fld


def batched_fields(ns, form, n_inst=256, seed=2, steps=1000):
    """The 256-instance parameter sweep (see module docstring)."""
    return [batched_field(ns, form, p, steps) for p in batched_params(n_inst, seed)]


def time_batched(ns, form, steps, warmup, chunk=50, fixed=None):
    import torch
    from pyfds_amd.batch import BatchRunner
    from pyfds_amd import autotune
    # signals long enough for the timed run and the schedule A/B after it
    fields = batched_fields(ns, form, steps=steps + warmup + 64 + 210)
    table = None
    if fixed is None:
        tile, chunk, table = autotune.tune_batch(fields, steps=20, chunks=(chunk,),
                                                 use_cache=False)
    else:
        tile = fixed[0]
        for f in fields:
            f.strip_height = fixed[1]
    batch = BatchRunner(fields, tile=tile, chunk=chunk)
    if warmup:
        batch.simulate(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch.simulate(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    cells = batch.cells
    kernel_s = batch.last_run_ms() / 1e3 / steps
    achieved = BYTES_PER_CELL[form] * cells / kernel_s / 1e9
    batch.close()
    # every member's pressure after the warm-up and the timed steps, before the A/B below
    # advances the same fields further
    values = [np.array(f.pressure.values) for f in fields]
    # A/B of the launch schedules on the same shape (the fields run on; timings only):
    # one workgroup per tile ('dispatch') against the persistent work queue ('queue')
    sched = {'dispatch': round(kernel_s * 1e6, 2)}
    ab_steps = max(3, min(steps, 200))
    try:
        qb = BatchRunner(fields, tile=tile, chunk=chunk, schedule='queue')
        qb.simulate(5)
        qb.simulate(ab_steps)
        sched['queue'] = round(qb.last_run_ms() * 1e3 / ab_steps, 2)
        qb.close()
    except ValueError as e:          # a shape without work-queue kernels
        sched['queue'] = 'n/a: %s' % e
    res = dict(gcells=cells * steps / wall / 1e9, ms_per_step=wall / steps * 1e3,
               kernel_us=kernel_s * 1e6, achieved_gbs=achieved, cells=cells, tile=tile,
               tune_us_per_step=table, schedule_us_per_step=sched)
    return res, fields, values


def batched_reference(ns, steps):
    """Pressure of every member of the sweep after ``steps`` steps of the FP64 batch (bit for
    bit the oracle's), against which the FP32 forms' error is measured."""
    from pyfds_amd.batch import BatchRunner
    fields = batched_fields(ns, 'fp64', steps=steps + 64)
    batch = BatchRunner(fields, chunk=50)
    batch.simulate(steps)
    out = [np.array(f.pressure.values) for f in fields]
    batch.close()
    for f in fields:
        f._runner().close()
    return out


def batched_error(values, ref):
    """Relative L2 error of each member's pressure (``values``, one array per member) against
    the FP64 reference: the largest, the median and the one over all members' cells together."""
    num = np.array([np.linalg.norm(v - r) for v, r in zip(values, ref)])
    den = np.array([np.linalg.norm(r) for r in ref])
    per = num / np.where(den > 0, den, 1.0)
    return dict(max=float(per.max()), median=float(np.median(per)),
                all_cells=float(np.sqrt((num ** 2).sum() / (den ** 2).sum())))


def thermal_field(ns, form, steps, nx=4096, seed=3):
    dx = 1e-3
    rho, cp, kmax = 1000.0, 4000.0, 50.0
    dt = 0.2 * dx ** 2 * rho * cp / kmax
    fld = ns.Thermal2D(x_samples=nx, x_delta=dx, y_samples=nx, y_delta=dx, t_samples=steps,
                       t_delta=dt, material=ns.ThermalMaterial(cp, rho, 0.6))
    fld.form = form
    rng = np.random.default_rng(seed)
    nb = nx // 16                      # 16x16 = 256-cell blocks
    hot = rng.random((nb, nb)) < 0.5
    blk = np.kron(hot, np.ones((16, 16), bool)).ravel()
    fld.add_material_region(ns.regions.Region(np.flatnonzero(blk)),
                            ns.ThermalMaterial(cp, rho, kmax))
    iy, ix = np.divmod(np.arange(nx * nx), nx)
    fld.temperature.values = np.exp(-0.5 * ((ix - nx / 2) ** 2 + (iy - nx / 2) ** 2) / 64.0 ** 2)
    fld.temperature.add_output(ns.regions.Region(
        [int(v) for v in rng.choice(nx * nx, 4, replace=False)]))
    return fld


def time_thermal(ns, form, steps, warmup, fixed=None):
    import torch
    fld = thermal_field(ns, form, steps + warmup + 200)
    fld.assemble_matrices()
    tune = tuned(fld, fixed)
    runner = fld._runner()
    runner.prepare(fld, steps)      # one-time set-up (chunk graphs) outside the timed steps
    if warmup:                      # then the warm-up, right before the timed steps
        fld.simulate(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    fld.simulate(steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    import ctypes as ct
    ms = ct.c_double()
    runner.lib.fds_last_run_ms(runner.ctx, ct.byref(ms))
    cells = fld.num_points
    kernel_s = ms.value / 1e3 / steps
    achieved = THERMAL_BYTES_PER_CELL[form] * cells / kernel_s / 1e9
    t = np.array(fld.temperature.values)
    launch_us, _ = launch_timing(fld)
    return dict(gcells=cells * steps / wall / 1e9, ms_per_step=wall / steps * 1e3,
                kernel_us=kernel_s * 1e6, achieved_gbs=achieved, t=t, tune=tune,
                launch_us=launch_us, cells=cells)


A1D_N, A1D_STEPS = 2000, 5000


def acoustic1d_field(ns, form='fp64', n=A1D_N, steps=A1D_STEPS):
    dx = 1e-4
    fld = ns.Acoustic1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=0.5 * dx / 1500.0,
                        material=ns.AcousticMaterial(1500.0, 1000.0))
    fld.form = form
    i = np.arange(n)
    fld.pressure.values = np.exp(-0.5 * ((i - n // 2) / 20.0) ** 2)
    fld.velocity.add_boundary(ns.regions.PointRegion([0], None, 'left'), value=0,
                              additive=False)
    fld.velocity.add_boundary(ns.regions.PointRegion([n - 1], None, 'right'), value=0,
                              additive=False)
    fld.pressure.add_output(ns.regions.Region([n // 4, n // 2, 3 * n // 4], 'probes'))
    return fld


def time_acoustic1d(ns, form, steps, tile=None, shape=None):
    import ctypes as ct
    from pyfds_amd import _native as nat
    fld = acoustic1d_field(ns, form)
    if tile is not None:
        fld.tile = tile
    if shape is not None:
        fld.line_shape = shape
    fld.simulate(10)
    runner = fld._runner()
    t0 = time.perf_counter()
    fld.simulate(steps)
    wall = time.perf_counter() - t0
    ms = ct.c_double()
    nat.check(runner.lib.fds_last_run_ms(runner.ctx, ct.byref(ms)))
    return dict(us_per_step=round(wall / steps * 1e6, 3),
                gpu_us_per_step=round(ms.value * 1e3 / steps, 3),
                gcells=round(fld.num_points * steps / wall / 1e9, 4),
                engine=nat.engine_of(runner.ctx), p=np.array(fld.pressure.values))


# resident-engine layouts timed: (cells per thread, threads, ghost); threads 0 = one workgroup
A1D_SHAPES = {'wave_segments_2x64_g32': (2, 64, 32), 'wave_segments_4x64_g96': (4, 64, 96),
              'segments_2x128_g64': (2, 128, 64), 'segments_2x256_g128': (2, 256, 128),
              'one_wg_cpt2': (2, 0, 0), 'one_wg_cpt4': (4, 0, 0)}


def acoustic1d_section(ns, steps, no_cpu):
    out = {}
    dflt = time_acoustic1d(ns, 'fp64', steps)
    out['default'] = {k: v for k, v in dflt.items() if k != 'p'}
    best = ('default', dflt)
    for name, shape in A1D_SHAPES.items():
        r = time_acoustic1d(ns, 'fp64', steps, shape=shape)
        out[name] = {k: v for k, v in r.items() if k != 'p'}
        out[name]['equal_to_default'] = bool(np.array_equal(r['p'], dflt['p']))
        if r['us_per_step'] < best[1]['us_per_step']:
            best = (name, r)
    tile = time_acoustic1d(ns, 'fp64', steps, tile=5)
    out['tile_engine'] = {k: v for k, v in tile.items() if k != 'p'}
    out['line_equals_tile_bitwise'] = bool(np.array_equal(dflt['p'], tile['p']))
    for form in ('fp32', 'fp32_kahan'):
        r = time_acoustic1d(ns, form, steps)
        out['line_' + form] = {k: v for k, v in r.items() if k != 'p'}
        out['line_' + form]['rel_l2_vs_fp64'] = float(
            np.linalg.norm(r['p'] - dflt['p']) / np.linalg.norm(dflt['p']))
    out['fp64'] = dict(value=dflt['gcells'], us_per_step=dflt['us_per_step'],
                       gpu_us_per_step=dflt['gpu_us_per_step'], layout='default',
                       fastest_layout=best[0],
                       speedup_vs_tile_engine=round(tile['us_per_step']
                                                    / dflt['us_per_step'], 2))
    out['config'] = ('Acoustic1D %d cells, rigid ends, 3 recorded points, %d steps, FP64 '
                     '(wall time of simulate(), output download included)' % (A1D_N, steps))
    out['thermoacoustic1d'] = thermoacoustic_section(ns, steps, 200)
    if not no_cpu:
        out['cpu_baseline'] = dict(value=round(cpu_baseline_acoustic1d(steps), 5),
                                   unit='Gcell-updates/s', cores=1, kind='port',
                                   sample='oracle (scipy dia_matrix products), the same '
                                          'Acoustic1D run, all %d steps' % steps)
    return out


def thermoacoustic_field(ns, steps, n=A1D_N):
    dx = 1e-4
    dt = 0.5 * dx / 1500.0
    fld = ns.ThermoAcoustic1D(x_samples=n, x_delta=dx, t_samples=steps, t_delta=dt,
                              thermal_material=ns.ThermalMaterial(4000.0, 1000.0, 0.6),
                              acoustic_material=ns.AcousticMaterial(1500.0, 1000.0,
                                                                    absorption_coef=3e-3),
                              stepping=4)
    i = np.arange(n)
    fld.pressure.values = 1e5 * np.exp(-0.5 * ((i - n // 2) / 20.0) ** 2)
    fld.velocity.add_boundary(ns.regions.PointRegion([0], None, 'left'), value=0,
                              additive=False)
    fld.velocity.add_boundary(ns.regions.PointRegion([n - 1], None, 'right'), value=0,
                              additive=False)
    fld.temperature.add_output(ns.regions.Region([n // 4, n // 2, 3 * n // 4], 'probes'))
    return fld


def thermoacoustic_section(ns, steps, host_steps):
    out = {}
    for mode, k in (('device_coupled', steps), ('host_protocol', host_steps)):
        fld = thermoacoustic_field(ns, steps + 10)
        if mode == 'host_protocol':
            fld._device_coupling = lambda: None
        fld.simulate(5)
        t0 = time.perf_counter()
        fld.simulate(k)
        wall = time.perf_counter() - t0
        out[mode] = dict(us_per_step=round(wall / k * 1e6, 3), steps=k)
    out['speedup'] = round(out['host_protocol']['us_per_step']
                           / out['device_coupled']['us_per_step'], 1)
    out['config'] = ('ThermoAcoustic1D %d cells, loss coupling every 4th step (accumulated), '
                     'FP64, wall time per coupled step' % A1D_N)
    return out


def cpu_baseline_acoustic1d(steps):
    import oracle
    oracle.Dia.use_scipy = True
    n = A1D_N
    o = oracle.PlainAcoustic1D(n, 1e-4, steps, 0.5 * 1e-4 / 1500.0,
                               [(np.arange(n), dict(sound_velocity=1500.0, density=1000.0,
                                                    absorption_coef=0))])
    i = np.arange(n)
    o.components['pressure'] = np.exp(-0.5 * ((i - n // 2) / 20.0) ** 2)
    o.boundaries['velocity'] = [dict(indices=[0], value=0, additive=False),
                                dict(indices=[n - 1], value=0, additive=False)]
    o.outputs['pressure'] = [dict(indices=[n // 4, n // 2, 3 * n // 4], signals=[])]
    o.assemble()
    t0 = time.perf_counter()
    o.simulate(steps)
    return n * steps / (time.perf_counter() - t0) / 1e9


def _oracle_of(f, steps):
    """The oracle of one sweep member (same materials, source and probes)."""
    import oracle
    mats = [(np.asarray(m.region.indices), dict(sound_velocity=m.materials[0].sound_velocity,
                                                 density=m.materials[0].density,
                                                 absorption_coef=0))
            for m in f.material_regions]
    cls = oracle.PlainAcoustic2D if type(f).__name__ == 'Acoustic2D' \
        else oracle.PlainAcoustic3DAxi
    o = cls(f.x.samples, f.x.increment, steps, f.t.increment, mats, ny=f.y.samples,
            dy=f.y.increment)
    b = f.pressure.boundaries[0]
    o.boundaries['pressure'] = [dict(indices=b.region.indices, value=b.value, additive=True)]
    o.outputs['pressure'] = [dict(indices=f.pressure.outputs[0].region.indices, signals=[])]
    o.assemble()
    return o


def _batched_worker(job):
    """Pool worker: step the sweep members ``ids`` on the oracle; (cell updates, seconds)."""
    ids, steps = job
    import oracle
    import pyfds_amd as ns
    oracle.Dia.use_scipy = True
    params = batched_params()
    cells, busy = 0, 0.0
    for i in ids:
        o = _oracle_of(batched_field(ns, 'fp64', params[i], steps), steps)
        t0 = time.perf_counter()
        o.simulate(steps)
        busy += time.perf_counter() - t0
        cells += o.n * steps
    return cells, busy


def cpu_baseline_batched(steps, n_inst=32):
    """Oracle over the first ``n_inst`` instances of the sweep, ``steps`` steps each, one core."""
    cells, busy = _batched_worker((list(range(n_inst)), steps))
    return cells / busy / 1e9


def cpu_baseline_batched_pool(steps, workers):
    """All instances of the sweep, ``steps`` steps each, on a pool of ``workers`` processes
    (one per core; instances are independent): cell updates over the busiest worker's time."""
    import multiprocessing as mp
    jobs = [(list(range(w, 256, workers)), steps) for w in range(workers)]
    with mp.get_context('spawn').Pool(workers) as pool:
        res = pool.map(_batched_worker, jobs)
    return sum(c for c, _ in res) / max(b for _, b in res) / 1e9


def cpu_baseline_thermal(steps):
    import oracle
    oracle.Dia.use_scipy = True
    import pyfds_amd as ns
    f = thermal_field(ns, 'fp64', steps)
    mats = [(np.asarray(m.region.indices), dict(heat_capacity=m.materials[0].heat_capacity,
                                                 density=m.materials[0].density,
                                                 thermal_conductivity_x=m.materials[0]
                                                 .thermal_conductivity_x,
                                                 thermal_conductivity_y=m.materials[0]
                                                 .thermal_conductivity_y))
            for m in f.material_regions]
    o = oracle.PlainThermal2D(f.x.samples, f.x.increment, steps, f.t.increment, mats,
                              ny=f.y.samples, dy=f.y.increment)
    o.components['temperature'] = np.array(f.temperature.values)
    o.assemble()
    t0 = time.perf_counter()
    o.simulate(steps)
    return f.num_points * steps / (time.perf_counter() - t0) / 1e9


def learned_section(no_cpu, runs=3, iters=1500):
    import numpy as np
    from pyfds_amd.learned import LearnedBoundary, Scene
    t0 = time.perf_counter()
    lb = LearnedBoundary(n_int=1024, big=3072, seed=0)
    shape = lb.generate(n_runs=runs, seed=100, after=1200)
    t_gen = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = lb.train(iters=iters, batch=32, lr=2e-3, log_every=max(1, iters // 5))
    t_train = time.perf_counter() - t0
    scene = Scene.random(np.random.default_rng(7), 1024)
    res = lb.validate(scene, tolerance=0.01, after=1200)
    steps = res['learned_steps']
    macs = 4 * lb.G * (48 * 32 * 5 + 32 * 48 * 5)
    out = dict(reflected_energy=round(res['learned'], 6),
               reflected_energy_without_net=round(res['plain_small'], 6),
               window_steps=list(res['window']), finite=res['finite'],
               learned_s=round(res['learned_s'], 4), enlarged_s=round(res['enlarged_s'], 4),
               speedup_vs_enlarged=round(res['speedup'], 3), timed_steps=steps,
               learned_gcells=round(lb.G ** 2 * steps / res['learned_s'] / 1e9, 3),
               enlarged_gcells=round(lb.big ** 2 * steps / res['enlarged_s'] / 1e9, 3),
               params=int(lb.n_params), net_gflop_per_step=round(2 * macs / 1e9, 4),
               training=dict(runs=runs, frames=list(shape), iters=iters, batch=32,
                             loss=[(int(i), round(v, 5)) for i, v in hist],
                             data_s=round(t_gen, 2), train_s=round(t_train, 2)),
               uses=('learned' if lb.validated <= 0.01 else 'enlarged (fallback)'))
    if not no_cpu:
        out['cpu_baseline'] = dict(value=round(cpu_baseline_enlarged(lb, scene, 3), 5),
                                   unit='Gcell-updates/s', cores=1, kind='port',
                                   sample='oracle (scipy dia_matrix products), plain 3072x3072 enlarged '
                                          'run, 3 steps')
    return out


def cpu_baseline_enlarged(lb, scene, steps):
    import oracle
    oracle.Dia.use_scipy = True
    f = lb.field(scene, lb.big, lb.off, steps)
    mats = [(np.asarray(m.region.indices), dict(sound_velocity=m.materials[0].sound_velocity,
                                                 density=m.materials[0].density,
                                                 absorption_coef=0))
            for m in f.material_regions]
    o = oracle.PlainAcoustic2D(lb.big, lb.dx, steps, lb.dt, mats, ny=lb.big, dy=lb.dx)
    b = f.pressure.boundaries[0]
    o.boundaries['pressure'] = [dict(indices=b.region.indices, value=b.value, additive=True)]
    o.assemble()
    t0 = time.perf_counter()
    o.simulate(steps)
    return lb.big ** 2 * steps / (time.perf_counter() - t0) / 1e9


def cpu_host():
    """The host the CPU baseline ran on: model, logical CPUs, the affinity mask's size and the
    threads the baseline used (numpy's element-wise operations and scipy's dia_matvec are
    single-threaded; no BLAS call is made)."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    aff = sorted(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else None
    return dict(model=model, cpu_count=os.cpu_count(),
                affinity=len(aff) if aff is not None else None,
                affinity_first=aff[:4] if aff else None, threads=1,
                omp_num_threads=os.environ.get('OMP_NUM_THREADS'))


def _pinned_worker(job):
    """Pool worker (spawned with OMP_NUM_THREADS=1): pin this process to one CPU, run the
    headline workload on the oracle; (cell updates, seconds)."""
    cpu, wl, steps = job
    if cpu is not None and hasattr(os, 'sched_setaffinity'):
        os.sched_setaffinity(0, {cpu})
    return cpu_baseline(wl, steps, rate=False)


def cpu_baseline_pinned(wl, steps, cpus):
    """The oracle on the headline workload in len(cpus) processes at once, each pinned to one
    CPU of ``cpus`` with OMP_NUM_THREADS (and the BLAS thread counts) at 1, each on its own
    copy of the workload (the reference has no intra-run parallelism; independent runs are how
    it uses more cores): total cell updates / the slowest process's time, Gcell-updates/s."""
    import multiprocessing as mp
    keys = ('OMP_NUM_THREADS', 'OPENBLAS_NUM_THREADS', 'MKL_NUM_THREADS')
    old = {k: os.environ.get(k) for k in keys}
    os.environ.update({k: '1' for k in keys})
    try:
        with mp.get_context('spawn').Pool(len(cpus)) as pool:
            res = pool.map(_pinned_worker, [(c, wl, steps) for c in cpus])
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
    return sum(c for c, _ in res) / max(t for _, t in res) / 1e9


def cpu_baseline(wl, steps, rate=True):
    """The oracle (plain FP64 method, scipy dia_matrix products as the reference uses) on the
    same workload, in this process (one thread: numpy's element-wise operations and scipy's
    dia_matvec do not spawn threads).  Gcell-updates/s, or (cell updates, seconds)."""
    import oracle
    oracle.Dia.use_scipy = True
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
    return n * steps / wall / 1e9 if rate else (n * steps, wall)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=STEPS)
    # untimed steps of the same run first: the first ~20 ms of sustained load on a fresh
    # process run up to 8 % slower (measured on 2048^2 FP64), 600 steps cover it
    ap.add_argument('--warmup', type=int, default=600)
    ap.add_argument('--forms', default='fp64,fp32,fp32_kahan')
    ap.add_argument('--cpu-steps', type=int, default=20)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--sections', default='acoustic2d,lossy,acoustic1d,batched,thermal,learned')
    ap.add_argument('--lossy-steps', type=int, default=500)
    ap.add_argument('--lossy-forms', default='fp64,fp32')
    ap.add_argument('--lossy-tile', type=int, default=None)
    ap.add_argument('--lossy-strip', type=int, default=0)
    ap.add_argument('--a1d-steps', type=int, default=A1D_STEPS)
    ap.add_argument('--learned-iters', type=int, default=1500)
    ap.add_argument('--batch-steps', type=int, default=1000)
    ap.add_argument('--batch-forms', default='fp32,fp32_kahan')
    ap.add_argument('--thermal-steps', type=int, default=2000)
    ap.add_argument('--thermal-forms', default='fp64,fp32,fp32_kahan')
    ap.add_argument('--tile', type=int, default=None,
                    help='Acoustic2D section: use this shape index instead of autotuning')
    ap.add_argument('--strip', type=int, default=0,
                    help='with --tile: strip height of a streaming shape (0: its own)')
    ap.add_argument('--thermal-tile', type=int, default=None)
    ap.add_argument('--thermal-strip', type=int, default=0)
    ap.add_argument('--batch-tile', type=int, default=None)
    ap.add_argument('--batch-strip', type=int, default=0)
    ap.add_argument('--no-traffic', action='store_true',
                    help='skip the rocprofv3 HBM traffic measurement of the chosen shapes')
    ap.add_argument('--traffic-probe', default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.traffic_probe:
        traffic_probe(args.traffic_probe)
        return
    if args.gpus != 1:
        print('bench.py measures one GPU only: pyfds_amd runs each field on one GPU and has '
              'no multi-GPU path (run with --gpus 1)', file=sys.stderr)
        sys.exit(2)

    import pyfds_amd as ns
    wl = workload(steps=args.steps + args.warmup + 200)
    forms = [f for f in args.forms.split(',') if f]
    fixed = None if args.tile is None else (args.tile, args.strip)
    res = {f: time_form(ns, wl, f, args.steps, args.warmup, fixed) for f in forms}
    main_form = 'fp64' if 'fp64' in res else forms[0]
    ref = res[main_form]
    out_forms = {}
    for f, r in res.items():
        spl = r['tune']['steps_per_launch']
        cells = NX * NY
        entry = dict(value=round(r['gcells'], 3), ms_per_step=round(r['ms_per_step'], 4),
                     kernel_us=round(r['kernel_us'], 2), launch_us=round(r['launch_us'], 2),
                     compulsory_frac=round(BYTES_PER_CELL[f] * cells / r['launch_us'] / 1e3
                                           / HBM_PEAK_GBS, 4),
                     algorithmic_frac=round(r['achieved_gbs'] / HBM_PEAK_GBS, 4),
                     bytes_per_cell=BYTES_PER_CELL[f], autotune=r['tune'])
        if f != main_form and main_form == 'fp64':
            den = np.linalg.norm(ref['p'])
            entry['rel_l2_vs_fp64'] = float(np.linalg.norm(r['p'] - ref['p']) / den)
        out_forms[f] = entry
    traffic = None if args.no_traffic else measure_traffic(
        'acoustic2d', main_form, ref['tune']['tile'], ref['tune']['strip'])
    roof, detail = roofline_entry(BYTES_PER_CELL[main_form], NX * NY, ref['launch_us'],
                                  ref['tune']['steps_per_launch'], traffic)
    detail['kernel'] = ref['tune']['shape']
    if main_form in ('fp64', 'fp32'):
        floor = pattern_floor(ref['tune'], NX, NY, 8 if main_form == 'fp64' else 4, 3,
                              ref['launch_us'])
        if floor is not None:
            if traffic is not None and 'pattern' in traffic:
                floor['hbm_bytes'] = traffic['pattern']['bytes']
                floor['physical_gbs'] = round(traffic['pattern']['bytes'] /
                                              floor['us_per_launch'] / 1e3, 1)
            detail['pattern_floor'] = floor
    detail['note'] = ('frac: measured HBM bytes per launch / launch time / 8 TB/s; a launch '
                      'advances steps_per_launch fused steps in one pass, so the per-step '
                      'algorithmic rate (algorithmic_frac, 64 B per cell update) may pass 1 '
                      'while the physical one cannot')
    cpu = cpu_all = None
    if not args.no_cpu:
        host = cpu_host()
        cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, 'sched_getaffinity') else [None]
        cpu_val = cpu_baseline_pinned(wl, args.cpu_steps, cpus[:1])
        host.update(threads=1, omp_num_threads='1 (set for the baseline processes)',
                    pinned_to=cpus[:1])
        cpu = dict(value=round(cpu_val, 5), unit='Gcell-updates/s', cores=1, kind='port',
                   sample='oracle (plain FP64 method with scipy dia_matrix products, as the '
                          'reference), Acoustic2D 2048x2048, same scene, %d steps (per-step '
                          'cost is constant), one process pinned to one CPU, OMP_NUM_THREADS=1'
                          % args.cpu_steps, host=host)
        workers = cpus[:16]
        cpu_all = dict(value=round(cpu_baseline_pinned(wl, args.cpu_steps, workers), 5),
                       unit='Gcell-updates/s', cores=len(workers), kind='port',
                       sample='the same oracle run in %d processes at once, one pinned to each '
                              'CPU, each on its own copy of the workload (%d steps); total cell '
                              'updates / slowest process' % (len(workers), args.cpu_steps))
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
        'roofline': roof,
        'roofline_detail': detail,
        'cpu_baseline': cpu,
        'cpu_baseline_all_cores': cpu_all,
        'forms': out_forms,
    }
    sections = set(args.sections.split(','))
    if 'lossy' in sections:
        lfix = None if args.lossy_tile is None else (args.lossy_tile, args.lossy_strip)
        line['lossy'] = {f: time_lossy(ns, wl, f, args.lossy_steps, 10, lfix)
                         for f in args.lossy_forms.split(',') if f}
        line['lossy']['config'] = ('Acoustic2D 2048x2048 with absorbing materials, %d steps'
                                   % args.lossy_steps)
    if 'acoustic1d' in sections:
        line['acoustic1d'] = acoustic1d_section(ns, args.a1d_steps, args.no_cpu)
    if 'batched' in sections:
        bat = {}
        # every member after the warm-up and the timed steps, FP64 (the oracle's bits)
        ref = batched_reference(ns, args.batch_steps + 10)
        for f in [x for x in args.batch_forms.split(',') if x]:
            r, flds, vals = time_batched(ns, f, args.batch_steps, 10,
                                   fixed=None if args.batch_tile is None
                                   else (args.batch_tile, args.batch_strip))
            bat[f] = dict(value=round(r['gcells'], 3), ms_per_step=round(r['ms_per_step'], 4),
                          kernel_us=round(r['kernel_us'], 2),
                          achieved_gbs=round(r['achieved_gbs'], 1),
                          roofline_frac=round(r['achieved_gbs'] / HBM_PEAK_GBS, 4),
                          bytes_per_cell=BYTES_PER_CELL[f], cells=r['cells'],
                          tile=r['tile'], autotune_us_per_step=r['tune_us_per_step'],
                          schedule_kernel_us_per_step=r['schedule_us_per_step'],
                          rel_l2_vs_fp64={k: float('%.3g' % v)
                                          for k, v in batched_error(vals, ref).items()})
            for fl in flds:
                fl._runner().close()
        del ref
        bat['config'] = ('256 Acoustic2D/Acoustic3DAxi instances, sides 128-1024, %d steps, '
                         'HIP-graph chunks of 50' % args.batch_steps)
        if not args.no_cpu:
            host = cpu_host()
            workers = max(1, min(16, host['affinity'] or 1))
            bat['cpu_baseline'] = dict(value=round(cpu_baseline_batched(2), 5),
                                       unit='Gcell-updates/s', cores=1, kind='port',
                                       sample='oracle (plain FP64 method, scipy dia_matrix '
                                              'products), first 32 of the 256 instances, '
                                              '2 steps each', host=host)
            bat['cpu_baseline_all_cores'] = dict(
                value=round(cpu_baseline_batched_pool(2, workers), 5), unit='Gcell-updates/s',
                cores=workers, kind='port',
                sample='the same oracle on all 256 instances, 2 steps each, one process per '
                       'core (instances are independent)')
        line['batched'] = bat
    if 'thermal' in sections:
        th = {}
        tvals = {}
        for f in [x for x in args.thermal_forms.split(',') if x]:
            r = time_thermal(ns, f, args.thermal_steps, 10,
                             None if args.thermal_tile is None
                             else (args.thermal_tile, args.thermal_strip))
            tvals[f] = r['t']
            traffic = None
            if f == 'fp64' and not args.no_traffic:
                traffic = measure_traffic('thermal', f, r['tune']['tile'], r['tune']['strip'])
            roof, detail = roofline_entry(THERMAL_BYTES_PER_CELL[f], r['cells'], r['launch_us'],
                                          r['tune']['steps_per_launch'], traffic)
            if f in ('fp64', 'fp32'):
                n = int(round(r['cells'] ** 0.5))
                floor = pattern_floor(r['tune'], n, n, 8 if f == 'fp64' else 4, 1,
                                      r['launch_us'])
                if floor is not None:
                    if traffic is not None and 'pattern' in traffic:
                        floor['hbm_bytes'] = traffic['pattern']['bytes']
                        floor['physical_gbs'] = round(traffic['pattern']['bytes'] /
                                                      floor['us_per_launch'] / 1e3, 1)
                    detail['pattern_floor'] = floor
            th[f] = dict(value=round(r['gcells'], 3), ms_per_step=round(r['ms_per_step'], 4),
                         kernel_us=round(r['kernel_us'], 2),
                         algorithmic_frac=round(r['achieved_gbs'] / HBM_PEAK_GBS, 4),
                         roofline=roof, roofline_detail=detail,
                         bytes_per_cell=THERMAL_BYTES_PER_CELL[f], autotune=r['tune'])
        if 'fp64' in tvals:
            for f, t in tvals.items():
                if f != 'fp64':
                    th[f]['rel_l2_vs_fp64'] = float(np.linalg.norm(t - tvals['fp64'])
                                                    / np.linalg.norm(tvals['fp64']))
        th['config'] = ('Thermal2D 4096x4096, k 0.6/50 in 256-cell blocks, Fo 0.2, %d of the '
                        '100000 steps' % args.thermal_steps)
        if not args.no_cpu:
            th['cpu_baseline'] = dict(value=round(cpu_baseline_thermal(3), 5),
                                      unit='Gcell-updates/s', cores=1, kind='port',
                                      sample='oracle (scipy dia_matrix products), Thermal2D '
                                             '4096x4096, 3 steps')
        line['thermal2d'] = th
    if 'learned' in sections:
        line['learned_bc'] = learned_section(args.no_cpu, iters=args.learned_iters)
    print(json.dumps(line), flush=True)


if __name__ == '__main__':
    main()
