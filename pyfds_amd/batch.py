"""Batched runs: many independent fields (a parameter sweep) advanced together.

pyfds steps one field per Python loop.  ``BatchRunner`` keeps that per-field API -- every
member is an ordinary ``pyfds_amd`` field with its own materials, boundaries and outputs -- but
advances all members with one kernel launch per time step over the tiles of every member
(``fds_batch_run``), replayed in HIP-graph chunks.  Members may differ in model (Acoustic2D,
Acoustic3DAxi, Thermal2D, ...) and grid size; they must share the numeric form and device,
and they step in lock step from a common step counter.
"""
import collections
import ctypes as ct

import numpy as np

from . import _native as nat
from .device import pipeline_sizes


SCHEDULES = {'dispatch': 0, 'queue': 1}


class BatchRunner:
    """Advance ``fields`` together; ``tile`` picks one shape of the tile table for all.

    ``schedule`` (streaming shapes): ``'dispatch'`` launches one workgroup per tile and lets
    the hardware dispatcher hand them out; ``'queue'`` launches a persistent grid (as many
    workgroups as the device holds at once) that takes tiles from a device work queue,
    largest first.  Both give the same bits."""

    def __init__(self, fields, tile=None, chunk=50, schedule='dispatch'):
        if schedule not in SCHEDULES:
            raise ValueError('schedule is one of {}'.format(sorted(SCHEDULES)))
        self.schedule = schedule
        self.fields = list(fields)
        if not self.fields:
            raise ValueError('empty batch')
        forms = {getattr(f, 'form', 'fp64') for f in self.fields}
        if len(forms) != 1:
            raise ValueError('batch members must share one form, got {}'.format(sorted(forms)))
        self.tile = -1 if tile is None else int(tile)
        self.chunk = int(chunk)
        self.lib = nat.load()
        self.handle = None
        self._members = None

    def close(self):
        if self.handle is not None:
            self.lib.fds_batch_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _prepare(self):
        runners = []
        for f in self.fields:
            if not f.matrices_assembled:
                f.assemble_matrices()
            r = f._runner()
            r.sync(f)
            runners.append(r)
        members = tuple(r.ctx.value for r in runners)
        if self.handle is None or members != self._members:
            self.close()
            arr = (ct.c_void_p * len(members))(*members)
            handle = ct.c_void_p()
            nat.check(self.lib.fds_batch_create(ct.byref(handle), arr, len(members),
                                                self.tile))
            self.handle, self._members = handle, members
            if self.schedule != 'dispatch':
                nat.check(self.lib.fds_batch_set_schedule(handle, SCHEDULES[self.schedule]))
        return runners

    def simulate(self, num_steps):
        """Advance every member ``num_steps`` steps (all members must be at the same step)."""
        steps = {f.step for f in self.fields}
        if len(steps) != 1:
            raise ValueError('batch members are at different steps: {}'.format(sorted(steps)))
        start = steps.pop()
        runners = self._prepare()
        valid = [v for v in (r._valid_steps() for r in runners) if v is not None]
        if valid and start + num_steps > min(valid):
            raise IndexError('a boundary signal ends at step {}'.format(min(valid)))
        sizes = pipeline_sizes(num_steps, num_steps, self.chunk,
                               sum(r._output_count() for r in runners),
                               nat.check(self.lib.fds_batch_steps_per_launch(self.handle)))
        if sizes:
            self._run_pipelined(runners, start, sizes)
            return
        rc = self.lib.fds_batch_run(self.handle, start, num_steps, self.chunk)
        if rc != nat.E_NONFINITE:
            nat.check(rc)
        msg = nat.error_message() if rc == nat.E_NONFINITE else None
        for f, r in zip(self.fields, runners):
            r.collect(f, num_steps)
            f.step += num_steps
        if msg is not None:        # the run completed; a member met a NaN or an infinity
            raise FloatingPointError(msg)

    def _run_pipelined(self, runners, start, sizes):
        """Asynchronous batch runs of ``sizes`` steps, the next one enqueued before the outputs
        of the current one are appended (``fds_batch_run_async`` / ``fds_batch_take``)."""
        lib, handle = self.lib, self.handle
        counts = [[c.size if c is not None else 0 for c in r.out_cells] for r in runners]
        queue = collections.deque()
        nxt = [start]

        def enqueue(k):
            nat.check(lib.fds_batch_run_async(handle, nxt[0], k, self.chunk))
            queue.append(k)
            nxt[0] += k

        nonfinite = []

        def take():
            k = queue[0]
            data = [[np.empty((k, m)) if m else None for m in cs] for cs in counts]
            ptrs = (ct.c_void_p * (3 * len(data)))(
                *[None if d is None else d.ctypes.data for ds in data for d in ds])
            rc = lib.fds_batch_take(handle, ptrs)
            if rc != nat.E_NONFINITE:
                nat.check(rc)
            queue.popleft()
            for f in self.fields:
                f.step += k
            for f, r, ds in zip(self.fields, runners, data):
                r._append_outputs(f, ds)
            if rc == nat.E_NONFINITE:
                nonfinite.append(nat.error_message())

        try:
            enqueue(sizes[0])
            for k in sizes[1:]:
                if nonfinite:
                    break
                enqueue(k)
                take()
            if queue:
                take()
        finally:
            while queue:
                take()
        if nonfinite:
            raise FloatingPointError(nonfinite[0])

    def last_run_ms(self):
        ms = ct.c_double()
        nat.check(self.lib.fds_batch_last_run_ms(self.handle, ct.byref(ms)))
        return ms.value

    @property
    def cells(self):
        return int(np.sum([f.num_points for f in self.fields]))
