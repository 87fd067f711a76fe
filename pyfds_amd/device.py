"""Device side of a field: one HIP context per field, kept in step with the host objects.

``DeviceRunner`` replaces the reference's per-step work of ``Field.simulate`` +
``sim_step`` (``pyfds/fields.py:67-95``, ``pyfds/acoustics.py:111-128`` ...): before a run it
uploads whatever the host changed (coefficients after ``assemble_matrices``, component
values the user read or assigned, the boundary and output lists), then hands the whole run
to ``fds_run`` and appends the recorded samples to the ``Output`` objects.

Boundary lists are compiled into per-cell programs with the reference's semantics
(``pyfds/fields.py:591-600``, ``pyfds/regions.py:125-145``): boundaries apply in list order;
within one boundary all old values are read before any write and, for a repeated index, the
last occurrence wins; the value is a scalar, ``signal[step]`` or one signal per point.
"""
import collections
import ctypes as ct

import numpy as np

from . import _native as nat

INT64_MAX = np.iinfo(np.int64).max
OUTPUT_BYTES_PER_SEGMENT = 256 << 20
# A run with recorded outputs is cut into about PIPELINE_SEGMENTS asynchronous runs (each
# staging at most PIPELINE_STAGE_BYTES of outputs) so that the host builds the per-sample
# lists of one while the device computes the next.
PIPELINE_SEGMENTS = 8
PIPELINE_STAGE_BYTES = 16 << 20


def pipeline_sizes(n, seg, chunk, m, group=1):
    """Step counts of the asynchronous runs of an ``n``-step run with ``m`` recorded values per
    step (``seg``: the longest run the caller allows), or None for one synchronous run.  Runs
    are whole graph chunks of ``chunk`` steps rounded to ``group`` (steps per launch), so that
    only the last run can end in shorter launches."""
    group = max(1, group)
    unit = max(group, chunk - chunk % group) if chunk > 0 else group
    if m == 0 or n < 4 * unit:
        return None
    size = -(-n // PIPELINE_SEGMENTS)
    size = -(-size // unit) * unit                # whole graph chunks
    size = min(size, seg, max(1, PIPELINE_STAGE_BYTES // (8 * m)))
    if size >= n:
        return None
    sizes = [size] * (n // size)
    if n % size:
        sizes.append(n % size)
    return sizes


class BoundaryProgram:
    """A component's boundary list compiled for the device (see fds_set_boundaries)."""

    __slots__ = ('cells', 'ptr', 'a', 'off', 'stride', 'len', 'table')

    def __init__(self, cells, ptr, a, off, stride, length, table):
        self.cells, self.ptr, self.a, self.off = cells, ptr, a, off
        self.stride, self.len, self.table = stride, length, table

    def same_as(self, other):
        return other is not None and all(
            np.array_equal(getattr(self, f), getattr(other, f)) for f in self.__slots__)

    @property
    def valid_steps(self):
        """Steps with a value for every entry (None when all values are scalars)."""
        finite = self.len[self.len < INT64_MAX]
        return int(finite.min()) if finite.size else None


def normalise_indices(indices, n, what):
    """Region indices as int64 with numpy's negative-index rule; IndexError if out of range."""
    idx = np.asarray(indices, dtype=np.int64).ravel()
    if idx.size and (idx.min() < -n or idx.max() >= n):
        bad = idx[(idx < -n) | (idx >= n)][0]
        raise IndexError('index {} is out of bounds for axis 0 with size {} ({})'
                         .format(bad, n, what))
    return np.where(idx < 0, idx + n, idx)


def compile_boundaries(boundaries, n):
    """Compile a boundary list into a :class:`BoundaryProgram`."""
    cells, a_all, off_all, stride_all, len_all = [], [], [], [], []
    table, toff = [], 0
    for b in boundaries:
        idx = normalise_indices(b.region.indices, n, 'boundary')
        k = idx.size
        if k == 0:
            continue
        a = float(b.additive)
        v = b.value
        if np.ndim(v) == 0:
            seg = np.array([float(v)])
            off = np.full(k, toff, np.int64)
            stride = np.zeros(k, np.int64)
            length = np.full(k, INT64_MAX, np.int64)
        elif isinstance(v, np.ndarray):
            arr = np.asarray(v, dtype=np.float64)
            if arr.ndim == 1:
                seg = arr
                off = np.full(k, toff, np.int64)
                stride = np.ones(k, np.int64)
                length = np.full(k, arr.shape[0], np.int64)
            elif arr.ndim == 2 and arr.shape[1] in (1, k):
                # value[step] is a row: one column per point (or one shared column)
                seg = np.ascontiguousarray(arr).ravel()
                col = np.arange(k) if arr.shape[1] == k else np.zeros(k, np.int64)
                off = toff + col.astype(np.int64)
                stride = np.full(k, arr.shape[1], np.int64)
                length = np.full(k, arr.shape[0], np.int64)
            else:
                raise ValueError('boundary signal of shape {} does not match {} points'
                                 .format(arr.shape, k))
        else:
            sigs = [np.asarray(s, dtype=np.float64).ravel() for s in v]
            if len(sigs) != k:
                raise ValueError('shape mismatch: {} per-point signals for {} boundary points'
                                 .format(len(sigs), k))
            lens = np.array([s.shape[0] for s in sigs], np.int64)
            seg = np.concatenate(sigs) if sigs else np.zeros(0)
            off = toff + np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
            stride = np.ones(k, np.int64)
            length = lens
        # repeated index inside one boundary: the last occurrence wins
        rev_u, rev_first = np.unique(idx[::-1], return_index=True)
        last = k - 1 - rev_first
        cells.append(rev_u)
        a_all.append(np.full(rev_u.size, a))
        off_all.append(off[last])
        stride_all.append(stride[last])
        len_all.append(length[last])
        table.append(seg)
        toff += seg.size
    if not cells:
        z = np.zeros(0, np.int64)
        return BoundaryProgram(z, np.zeros(0, np.int32), np.zeros(0), z, z, z, np.zeros(0))
    cell = np.concatenate(cells)
    order = np.argsort(cell, kind='stable')      # keeps list order per cell
    cell = cell[order]
    ucell, counts = np.unique(cell, return_counts=True)
    ptr = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    return BoundaryProgram(ucell.astype(np.int64), ptr, np.concatenate(a_all)[order],
                           np.concatenate(off_all)[order], np.concatenate(stride_all)[order],
                           np.concatenate(len_all)[order],
                           np.ascontiguousarray(np.concatenate(table), dtype=np.float64))


def _copy_value(v):
    """A boundary value as the program compiler reads it, copied (scalar, signal array or
    per-point signals), to detect in-place changes between runs."""
    if np.ndim(v) == 0:
        return ('s', float(v))
    if isinstance(v, np.ndarray):
        return ('a', v.copy())
    return ('l', [np.array(x, dtype=np.float64, copy=True) for x in v])


def _same_value(v, snap):
    kind, data = snap
    if kind == 's':
        return np.ndim(v) == 0 and float(v) == data
    if kind == 'a':
        # by bits: equal bits are equal values; a difference only in bits (-0.0, NaN payloads)
        # costs a recompile, which the program comparison then finds unchanged
        return isinstance(v, np.ndarray) and v.shape == data.shape and v.dtype == data.dtype \
            and v.tobytes() == data.tobytes()
    if np.ndim(v) == 0 or isinstance(v, np.ndarray) or len(v) != len(data):
        return False
    return all(np.array_equal(np.asarray(x, dtype=np.float64), d) for x, d in zip(v, data))


def _int_list_copy(idx):
    """A copy of ``idx`` when it is a flat Python list of ints (regions hold their indices so),
    else None.  Comparing such a list with its copy by list equality gives the answer that
    comparing the two as arrays gives, without rebuilding an array from the list."""
    if type(idx) is not list or not all(type(x) is int for x in idx):
        return None
    return list(idx)


def _snapshot(items, value_of=None):
    """What a boundary or output list holds, by identity and by content: the objects, their
    regions, copies of the index arrays (and of flat int lists as lists) and (boundaries) of
    the values."""
    snap = []
    for it in items:
        idx = it.region.indices
        entry = (it, it.region, _int_list_copy(idx), np.array(idx, copy=True))
        if value_of is not None:
            entry += (bool(it.additive), it.value, _copy_value(it.value))
        snap.append(entry)
    return snap


def _unchanged(items, snap, boundaries):
    """True when ``items`` still hold exactly what ``snap`` recorded (same objects, regions,
    index values and, for boundaries, additive flags and values)."""
    if snap is None or len(items) != len(snap):
        return False
    for it, entry in zip(items, snap):
        if it is not entry[0] or it.region is not entry[1]:
            return False
        idx = it.region.indices
        if entry[2] is not None and type(idx) is list:
            if idx != entry[2]:
                return False
        else:
            arr = np.asarray(idx)
            if arr.shape != entry[3].shape or not np.array_equal(arr, entry[3]):
                return False
        if boundaries:
            if bool(it.additive) != entry[4] or not _same_value(it.value, entry[6]):
                return False
    return True


class DeviceRunner:
    """Owns the ``fds_ctx`` of one field."""

    def __init__(self, field):
        self.lib = nat.load()
        self.ctx = None
        self.coef_version = None
        self.programs = [None, None, None]
        self.bsnap = [None, None, None]     # boundary lists the programs were compiled from
        self.osnap = [None, None, None]     # output lists the recorded cells came from
        self.out_cells = [None, None, None]
        self.out_layout = [[], [], []]
        self.owners = [None, None, None]
        self.flow = None
        self.nf_check = False
        self.chunk = getattr(field, 'graph_chunk', 50)
        self.tile = getattr(field, 'tile', None)
        self.strip_height = getattr(field, 'strip_height', None)
        self.line_shape = getattr(field, 'line_shape', None)

    # -- lifetime -------------------------------------------------------------------------
    def _create(self, field):
        form = nat.FORMS.get(getattr(field, 'form', 'fp64'))
        if form is None:
            raise ValueError('unknown form {!r}; use one of {}'.format(field.form,
                                                                        sorted(nat.FORMS)))
        nx, ny = field._device_grid()
        ctx = ct.c_void_p()
        nat.check(self.lib.fds_create(ct.byref(ctx), field._model_id, form, nx, ny,
                                      int(getattr(field, 'device', 0))))
        self.ctx = ctx
        self.n = nx * ny
        if self.tile is not None:
            nat.check(self.lib.fds_set_tile(self.ctx, int(self.tile)))
        if self.strip_height is not None:
            nat.check(self.lib.fds_set_strip_height(self.ctx, int(self.strip_height)))
        if self.line_shape is not None:
            cpt, threads, ghost = (int(v) for v in self.line_shape)
            nat.check(self.lib.fds_set_line_shape(self.ctx, cpt, threads, ghost))

    def close(self):
        if self.ctx is not None and self.lib is not None:
            self.lib.fds_destroy(self.ctx)
        self.ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- host -> device ----------------------------------------------------------------
    def _components(self, field):
        return [getattr(field, name) if name else None for name in field._component_names]

    def sync(self, field):
        if self.ctx is None:
            self._create(field)
        lib, ctx = self.lib, self.ctx
        if self.coef_version != field._coef_version:
            coefs, scalars = field._device_coefficients()
            for slot in range(nat.NUM_COEFF):
                arr = coefs.get(slot)
                if arr is None:
                    nat.check(lib.fds_set_coeff(ctx, slot, None, 0))
                else:
                    arr = nat.f64(np.broadcast_to(arr, (self.n,)))
                    nat.check(lib.fds_set_coeff(ctx, slot, nat.dptr(arr), arr.size))
            s = nat.f64(np.zeros(nat.NUM_SCALARS))
            for i, v in scalars.items():
                s[i] = v
            nat.check(lib.fds_set_scalars(ctx, nat.dptr(s), s.size))
            nat.check(lib.fds_commit(ctx))
            self.coef_version = field._coef_version
        want = bool(getattr(field, 'check_finite', False))
        if want != self.nf_check:
            nat.check(lib.fds_set_nonfinite_check(ctx, int(want)))
            self.nf_check = want
        periods = getattr(field, '_device_flow_periods', None)
        if periods is not None:
            p = nat.i64(periods())
            if self.flow is None or not np.array_equal(p, self.flow):
                nat.check(lib.fds_set_flow(ctx, nat.lptr(p), p.size))
                self.flow = p.copy()
        comps = self._components(field)
        for k, comp in enumerate(comps):
            if comp is None:
                continue
            if comp._host_dirty or self.owners[k] is not comp or comp._device is not None \
                    and comp._device[0] is not self:
                vals = nat.f64(comp._values)
                if vals.shape != (self.n,):
                    raise ValueError('component values must hold {} points, got shape {}'
                                     .format(self.n, np.shape(comp._values)))
                nat.check(lib.fds_set_state(ctx, k, nat.dptr(vals), vals.size))
                comp._host_dirty = False
                comp._device = None
                self.owners[k] = comp
            # boundary program: recompiled only when the list, a region, its indices, a flag
            # or a value (signals are compared by content: users may edit them in place)
            # changed since the last upload
            if not _unchanged(comp.boundaries, self.bsnap[k], True):
                prog = compile_boundaries(comp.boundaries, self.n)
                if not prog.same_as(self.programs[k]):
                    nat.check(lib.fds_set_boundaries(
                        ctx, k, prog.cells.size, nat.lptr(prog.cells), nat.iptr(prog.ptr),
                        prog.a.size, nat.dptr(prog.a), nat.lptr(prog.off),
                        nat.lptr(prog.stride), nat.lptr(prog.len), prog.table.size,
                        nat.dptr(prog.table)))
                    self.programs[k] = prog
                self.bsnap[k] = _snapshot(comp.boundaries, True)
            if not _unchanged(comp.outputs, self.osnap[k], False) or self.out_cells[k] is None:
                layout, cells, start = [], [], 0
                for out in comp.outputs:
                    idx = normalise_indices(out.region.indices, self.n, 'output')
                    layout.append((out, start, idx.size))
                    cells.append(idx)
                    start += idx.size
                cells = np.concatenate(cells) if cells else np.zeros(0, np.int64)
                if self.out_cells[k] is None or not np.array_equal(cells, self.out_cells[k]):
                    cells = nat.i64(cells)
                    nat.check(lib.fds_set_outputs(ctx, k, cells.size, nat.lptr(cells)))
                    self.out_cells[k] = cells
                self.out_layout[k] = layout
                self.osnap[k] = _snapshot(comp.outputs)
        return comps

    # -- running --------------------------------------------------------------------------
    def _valid_steps(self):
        v = [p.valid_steps for p in self.programs if p is not None]
        v = [x for x in v if x is not None]
        return min(v) if v else None

    def _segment(self, num_steps, progress):
        m = sum(c.size for c in self.out_cells if c is not None)
        seg = num_steps if m == 0 else max(1, OUTPUT_BYTES_PER_SEGMENT // (8 * m))
        if progress is not None:
            seg = min(seg, max(self.chunk, -(-num_steps // 20)))
        return max(1, min(seg, num_steps))

    def _output_count(self):
        return sum(c.size for c in self.out_cells if c is not None)

    def _run_pipelined(self, field, sizes, progress):
        """Run ``sizes`` steps after steps as asynchronous runs, the next one enqueued before
        the outputs of the current one are appended (``fds_run_async`` / ``fds_run_take``)."""
        lib, ctx, chunk = self.lib, self.ctx, int(self.chunk)
        counts = [c.size if c is not None else 0 for c in self.out_cells]
        queue = collections.deque()
        start = [field.step]

        def enqueue(k):
            nat.check(lib.fds_run_async(ctx, start[0], k, chunk))
            queue.append(k)
            start[0] += k

        nonfinite = []

        def take():
            k = queue[0]
            data = [np.empty((k, m)) if m else None for m in counts]
            rc = lib.fds_run_take(ctx, *[None if d is None else nat.dptr(d) for d in data])
            if rc != nat.E_NONFINITE:
                nat.check(rc)
            # the device has run these steps: count them before anything else can fail
            queue.popleft()
            first = field.step
            field.step += k
            self._append_outputs(field, data)
            if progress is not None:
                for s in range(first, first + k):
                    progress.log(s)
            if rc == nat.E_NONFINITE:
                nonfinite.append(nat.error_message())

        try:
            enqueue(sizes[0])
            for k in sizes[1:]:
                if nonfinite:          # a non-finite value was found: enqueue no more steps
                    break
                enqueue(k)
                take()
            if queue:
                take()
        finally:
            # after an error, keep the host in step with what the device ran
            while queue:
                take()
        if nonfinite:
            raise FloatingPointError(nonfinite[0])

    def run_steps(self, field, start, nsteps):
        """Run ``nsteps`` steps from pyfds step ``start``; append the recorded outputs."""
        self.sync(field)
        self._run(field, start, nsteps)

    def _run(self, field, start, nsteps):
        """``run_steps`` after ``sync``."""
        rc = self.lib.fds_run(self.ctx, start, nsteps, int(self.chunk))
        if rc == nat.E_NONFINITE:       # the run completed: keep its outputs, then raise
            msg = nat.error_message()
            self.collect(field, nsteps)
            raise FloatingPointError(msg)
        nat.check(rc)
        self.collect(field, nsteps)

    def collect(self, field, nsteps):
        """After a device run of ``nsteps`` steps: append the recorded samples to the
        ``Output`` objects and mark the components as held by the device."""
        data = []
        for k, comp in enumerate(self._components(field)):
            m = self.out_cells[k].size if comp is not None else 0
            d = None
            if m:
                d = np.empty((nsteps, m))
                nat.check(self.lib.fds_read_outputs(self.ctx, k, 0, nsteps, nat.dptr(d)))
            data.append(d)
        self._append_outputs(field, data)

    def _append_outputs(self, field, data):
        """Append one run's recorded samples (``data[k]``: steps x outputs of component k, or
        None) to the ``Output`` objects as ``write_outputs`` would have, step by step
        (``pyfds/fields.py:602-611``), and mark the components as held by the device."""
        for k, comp in enumerate(self._components(field)):
            if comp is None:
                continue
            if data[k] is not None:
                for out, s0, cnt in self.out_layout[k]:
                    out._record(data[k][:, s0:s0 + cnt])
            comp._device = (self, k)
            comp._host_dirty = False

    def run_recording(self, field, start, nsteps, cells):
        """Run ``nsteps`` steps recording ``cells[k]`` (flat indices) of component k every
        step; returns one float32 array [nsteps][len(cells[k])] per component.  The field's own
        ``Output`` objects are not touched (they are re-synchronised on the next run)."""
        self.sync(field)
        for k, idx in enumerate(cells):
            idx = nat.i64(idx)
            nat.check(self.lib.fds_set_outputs(self.ctx, k, idx.size, nat.lptr(idx)))
            self.out_cells[k] = None
        nat.check(self.lib.fds_run(self.ctx, start, nsteps, int(self.chunk)))
        out = []
        for k, idx in enumerate(cells):
            data = np.empty((nsteps, len(idx)))
            nat.check(self.lib.fds_read_outputs(self.ctx, k, 0, nsteps, nat.dptr(data)))
            out.append(data.astype(np.float32))
        for k, comp in enumerate(self._components(field)):
            if comp is not None:
                comp._device = (self, k)
                comp._host_dirty = False
        return out

    def simulate(self, field, start, num_steps, progress):
        """``Field.simulate``'s loop; runs until the shortest boundary signal ends and then
        raises IndexError like the reference's ``signal[step]`` would."""
        self.sync(field)
        valid = self._valid_steps()
        n_ok = num_steps if valid is None else max(0, min(num_steps, valid - start))
        seg = self._segment(num_steps, progress)
        m = self._output_count()
        sizes = pipeline_sizes(n_ok, seg, int(self.chunk), m,
                               nat.check(self.lib.fds_steps_per_launch(self.ctx)))
        if sizes is None and m and n_ok:
            # runs with outputs take the asynchronous path even unpipelined: the outputs come
            # back through pinned staging behind the last step instead of a copy afterwards
            sizes = [seg] * (n_ok // seg) + ([n_ok % seg] if n_ok % seg else [])
        if sizes:
            self._run_pipelined(field, sizes, progress)
        done = n_ok if sizes else 0
        while done < n_ok:
            k = min(seg, n_ok - done)
            try:
                self._run(field, field.step, k)
            except FloatingPointError:
                field.step += k          # the device ran these steps
                raise
            if progress is not None:
                for s in range(field.step, field.step + k):
                    progress.log(s)
            field.step += k
            done += k
        if n_ok < num_steps:
            raise IndexError('index {} is out of bounds for axis 0 with size {}'
                             .format(start + n_ok, valid))

    def prepare(self, field, num_steps):
        """Do the one-time set-up of ``simulate(num_steps)`` from the field's current step
        (tables, output rows, the graphs its runs replay) without stepping."""
        self.sync(field)
        seg = self._segment(num_steps, None)
        sizes = pipeline_sizes(num_steps, seg, int(self.chunk), self._output_count(),
                               nat.check(self.lib.fds_steps_per_launch(self.ctx)))
        longest = max(sizes) if sizes else min(seg, num_steps)
        nat.check(self.lib.fds_prepare_run(self.ctx, field.step, longest, int(self.chunk)))

    def single_step(self, field):
        """One ``sim_step`` at ``field.step`` (the caller advances the step counter)."""
        self.sync(field)
        valid = self._valid_steps()
        if valid is not None and field.step >= valid:
            raise IndexError('index {} is out of bounds for axis 0 with size {}'
                             .format(field.step, valid))
        self.run_steps(field, field.step, 1)

    # -- device -> host -------------------------------------------------------------------
    def download(self, k, out=None):
        """Component k's device state into ``out`` (a new array if None)."""
        if out is None:
            out = np.empty(self.n)
        nat.check(self.lib.fds_get_state(self.ctx, k, nat.dptr(out), self.n))
        return out

    def download_all(self, field):
        for comp in self._components(field):
            if comp is not None and comp._device is not None:
                comp.values  # noqa: B018 (property access downloads)
