"""Autotuner: picks the step kernels' shape, strip height and HIP-graph chunk by timing.

For a field (or a batch of fields) it times a few steps of every admissible shape of the
compiled shape table (``fds_tile_count``/``fds_tile_shape``/``fds_tile_engine``; streaming
shapes also at a few strip heights, ``fds_set_strip_height``) and graph-chunk length with the
GPU's own clock (HIP events around the launches, ``fds_last_run_ms`` /
``fds_batch_last_run_ms``), keeps the fastest and caches the choice in a JSON file keyed by
device, model, form, grid and lossy/Axi variant.  Timing runs advance a scratch copy of the
state: the caller's field is restored to its values before tuning.
"""
import ctypes as ct
import json
import os

import numpy as np

from . import _native as nat

CACHE = os.environ.get('PYFDS_AMD_TUNE_CACHE',
                       os.path.join(os.path.expanduser('~'), '.cache', 'pyfds_amd_tune.json'))


def _device_name():
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:
        return 'hip-device'


def _load_cache():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save_cache(cache):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, 'w') as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass


def _key(fields):
    f = fields[0]
    lossy = any(getattr(x, '_coefs', {}).get(4) is not None for x in fields)
    cells = sum(x.num_points for x in fields)
    return '|'.join(str(v) for v in (_device_name(), type(f).__name__, f.form, len(fields),
                                     cells, f._device_grid(), lossy, 'v4'))


def _snapshot(fields):
    out = []
    for f in fields:
        comps = [getattr(f, n) for n in f._component_names if n]
        out.append(([np.array(c.values, copy=True) for c in comps],
                    [[[list(s) for s in o.signals] for o in c.outputs] for c in comps],
                    f.step))
    return out


def _restore(fields, snap):
    for f, (vals, sigs, step) in zip(fields, snap):
        comps = [getattr(f, n) for n in f._component_names if n]
        for c, v, s in zip(comps, vals, sigs):
            c.values = v
            for o, so in zip(c.outputs, s):
                o.signals = [list(x) for x in so]
        f.step = step


# strip heights also tried for the 16-row and the 32-row streaming shapes (the others: their own
# height); the time per step depends on the height irregularly (on 2048^2 acoustics the
# register-path shapes did worst at powers of two, the LDS-ring shapes best at 8 or 12 rows;
# on 4096^2 thermal the ring shapes at 16), so the tuner times each
STRIP_HEIGHTS = {16: (6, 12, 20, 24), 32: (8, 12, 16, 20, 24)}


def tune_field(fld, steps=40, chunks=(50,), use_cache=True):
    """Choose ``fld.tile`` / ``fld.strip_height`` / ``fld.graph_chunk`` for one field; returns
    (tile, chunk, table) where table maps 'tile,strip,chunk' -> microseconds per step (strip 0:
    the shape's own height)."""
    return _tune([fld], steps, chunks, use_cache, batch=False)


def tune_batch(fields, steps=20, chunks=(50,), use_cache=True):
    """Choose one tile shape and chunk for a ``BatchRunner`` over ``fields``."""
    return _tune(fields, steps, chunks, use_cache, batch=True)


def _tune(fields, steps, chunks, use_cache, batch):
    lib = nat.load()
    key = _key(fields) + ('|batch' if batch else '')
    cache = _load_cache() if use_cache else {}
    if key in cache:
        tile, chunk = cache[key]['tile'], cache[key]['chunk']
        _apply(fields, tile, chunk, cache[key].get('strip', 0))
        return tile, chunk, cache[key]['table']
    for f in fields:
        if not f.matrices_assembled:
            f.assemble_matrices()
    snap = _snapshot(fields)
    shapes, engines = nat.tile_shapes(), nat.tile_engines()
    single_only = [d['single_only'] for d in nat.tile_table()]
    one_row = fields[0]._device_grid()[1] == 1
    lossy = any(getattr(x, '_coefs', {}).get(4) is not None for x in fields)
    cand = []
    for i, (tx, ty) in enumerate(shapes):
        if (ty == 1) != one_row:
            continue
        if engines[i] == 'stream':
            # 1D models run on the resident engine; batches stream lossless members only
            # (a lossy field runs any streaming shape one step per launch); the shapes with
            # deeper prefetch or the LDS ring serve single lossless fields
            if one_row or (single_only[i] and (lossy or batch)) or (lossy and batch):
                continue
            cand.append((i, 0))
            cand += [(i, h) for h in STRIP_HEIGHTS.get(ty, ())]
        else:
            cand.append((i, 0))
    table = {}
    for tile, strip in cand:
        for chunk in chunks:
            try:
                _set_strip(fields, strip)
                if batch:
                    from .batch import BatchRunner
                    br = BatchRunner(fields, tile=tile, chunk=chunk)
                    br.simulate(min(steps, 5))
                    br.simulate(steps)
                    ms = br.last_run_ms()
                    br.close()
                else:
                    f = fields[0]
                    r = f._runner()
                    r.sync(f)
                    nat.check(lib.fds_set_tile(r.ctx, tile))
                    r.chunk = chunk
                    r.run_steps(f, f.step, min(steps, 5))
                    r.run_steps(f, f.step, steps)
                    out = ct.c_double()
                    nat.check(lib.fds_last_run_ms(r.ctx, ct.byref(out)))
                    ms = out.value
            except ValueError:        # tile does not fit this field's LDS budget
                continue
            finally:
                _restore(fields, snap)
            table['%d,%d,%d' % (tile, strip, chunk)] = ms * 1e3 / steps
    if not table:
        raise RuntimeError('no tile shape could be timed')
    best = min(table, key=table.get)
    tile, strip, chunk = (int(v) for v in best.split(','))
    _apply(fields, tile, chunk, strip)
    cache[key] = dict(tile=tile, strip=strip, chunk=chunk, table=table)
    if use_cache:
        _save_cache(cache)
    return tile, chunk, table


def _set_strip(fields, strip):
    lib = nat.load()
    for f in fields:
        f.strip_height = strip
        r = f.__dict__.get('_device_runner')
        if r is not None:
            r.strip_height = strip
            if r.ctx is not None:
                nat.check(lib.fds_set_strip_height(r.ctx, strip))


def _apply(fields, tile, chunk, strip=0):
    lib = nat.load()
    _set_strip(fields, strip)
    for f in fields:
        f.tile = tile
        f.graph_chunk = chunk
        r = f.__dict__.get('_device_runner')
        if r is not None:
            r.tile, r.chunk = tile, chunk
            if r.ctx is not None:
                nat.check(lib.fds_set_tile(r.ctx, tile))
