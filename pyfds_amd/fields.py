"""Fields, dimensions and field components (drop-in for ``pyfds.fields``).

The classes keep the reference's constructors, attributes and methods
(``pyfds/fields.py:9-668``).  What changes is where the time stepping runs:

* ``Field.simulate`` / ``sim_step`` of the models hand the whole step loop to the HIP
  library through ``pyfds_amd.device.DeviceRunner`` (matrix-free stencil kernels, boundaries
  and outputs applied on the device); there is no CPU time-stepping path.
* ``FieldComponent.values`` is a property backed by the device copy: it downloads lazily
  after a device run and marks the host copy for upload when it is read or assigned, so
  user code that reads or modifies ``values`` between runs keeps the reference semantics.
* ``d_x/d_y/d_x2/d_y2`` still return scipy DIA matrices (public API, used for the ``a_*``
  attributes), built exactly like the reference's.
* region builders are vectorised with numpy and return int64 arrays of indices in the
  reference's order.
* two optional keyword-only constructor arguments select the numeric form
  (``form='fp64' | 'fp32' | 'fp32_kahan'``) and the HIP device.
* ``check_finite = True`` on a field makes its device runs scan the state for NaN and
  infinity (the first component after every graph chunk of steps, every component after the
  run); ``simulate`` then raises ``FloatingPointError`` naming the component and the step by
  which it was found, after the steps run so far are counted and their outputs recorded (a
  pipelined run stops enqueueing further steps).  Off by default: the reference propagates
  non-finite values silently.
* ``sim`` is a short name of ``simulate``.
"""
import logging as lo

import numpy as np
import scipy.sparse as sp

from . import regions as reg

logger = lo.getLogger('pyfds')


class Field:
    """Base class for all fields (``pyfds/fields.py:9-127``)."""

    check_finite = False

    def __init__(self):
        self.material_regions = []
        self.step = 0
        self.matrices_assembled = False
        self.t = None

    @property
    def num_points(self):
        raise NotImplementedError

    def get_index(self, position):
        raise NotImplementedError

    def material_vector(self, mat_parameter):
        """Vector of ``mat_parameter`` for every point: zeros, then each material region in
        list order (later regions overwrite); ``KeyError`` if no material defines it
        (``pyfds/fields.py:34-57``)."""
        found = False
        vec = np.zeros(self.num_points)
        for mat_reg in self.material_regions:
            for mat in mat_reg.materials:
                if hasattr(mat, mat_parameter):
                    self._fill(vec, mat_reg.region.indices, getattr(mat, mat_parameter))
                    found = True
        if not found:
            raise KeyError('Material parameter {} not found in set materials.'
                           .format(mat_parameter))
        return vec

    def _fill(self, vec, indices, value):
        """``vec[indices] = value`` (the order of ``indices`` does not matter here)."""
        vec[indices] = value

    def assemble_matrices(self):
        raise NotImplementedError

    def sim_step(self):
        raise NotImplementedError

    # -- device step loop ------------------------------------------------------------------
    def _device_loop_ok(self):
        """True when ``sim_step`` is the model's own device step (not overridden by a
        subclass), so ``simulate`` may hand whole chunks of steps to the device."""
        own = getattr(type(self), '_device_sim_step', None)
        return own is not None and type(self).sim_step is own

    def simulate(self, num_steps=None):
        """Run ``num_steps`` steps (``t.samples`` if falsy) from ``self.step``
        (``pyfds/fields.py:67-95``).  Assembles once; logs progress every 5 % for
        un-segmented runs."""
        if not num_steps:
            num_steps = self.t.samples
            progress_logger = ProgressLogger(num_steps)
        else:
            progress_logger = None

        if not self.matrices_assembled:
            self.assemble_matrices()
            logger.info('Matrices created.')

        logger.info('Starting simulation of {} steps.'.format(num_steps))

        start_step = self.step
        if self._device_loop_ok():
            self._runner().simulate(self, start_step, num_steps, progress_logger)
        else:
            while self.step < start_step + num_steps:
                self.sim_step()
                if progress_logger:
                    progress_logger.log(self.step)
                self.step += 1

        logger.info('Simulation of {} steps completed.'.format(num_steps))

    def sim(self, num_steps=None):
        """Short name of :meth:`simulate` (same arguments and behaviour)."""
        return self.simulate(num_steps)

    def _runner(self):
        runner = self.__dict__.get('_device_runner')
        if runner is None:
            from .device import DeviceRunner
            runner = DeviceRunner(self)
            self.__dict__['_device_runner'] = runner
        return runner

    def __getstate__(self):
        """Picklable (the reference's Animator pickles whole fields): device state is
        downloaded into the components and the device context is dropped."""
        state = dict(self.__dict__)
        runner = state.pop('_device_runner', None)
        if runner is not None:
            runner.download_all(self)
        return state

    def get_point_region(self, position, name=''):
        """Point region at ``position`` (``pyfds/fields.py:97-108``)."""
        return reg.PointRegion([self.get_index(position)], position, name=name)

    def add_material_region(self, *args, **kwargs):
        """Append a ``MaterialRegion`` (``pyfds/fields.py:110-119``)."""
        new_material_region = reg.MaterialRegion(*args, **kwargs)
        self.material_regions.append(new_material_region)
        logger.info('Material region {} added.'.format(new_material_region.region.name))

    def reset(self):
        """Zero every component, keep boundaries and outputs, step 0
        (``pyfds/fields.py:121-127``)."""
        for name in dir(self):
            comp = getattr(self, name, None)
            if isinstance(comp, FieldComponent):
                comp.values = np.zeros(comp.num_points)
        self.step = 0


def _dia(diags, offsets, n):
    return sp.dia_matrix((np.array(diags), offsets), shape=(n, n))


def _first_difference(factors, n, stride, variant):
    """Column-indexed first difference: offsets [0, s] forward, [-s, s] central, [-s, 0]
    backward (``pyfds/fields.py:158-184, 273-328``)."""
    if variant == 'forward':
        return _dia([-factors, factors], [0, stride], n)
    if variant == 'central':
        return _dia([-factors / 2, factors / 2], [-stride, stride], n)
    if variant == 'backward':
        return _dia([-factors, factors], [-stride, 0], n)
    raise ValueError('Unknown difference quotient variant {}.'.format(variant))


def _second_difference(factors, n, stride):
    """Three-point second difference (``pyfds/fields.py:186-202, 330-365``)."""
    return _dia([factors, -2 * factors, factors], [-stride, 0, stride], n)


class Field1D(Field):
    """One-dimensional field (``pyfds/fields.py:130-241``)."""

    def __init__(self, x_samples, x_delta, t_samples, t_delta, material, *, form='fp64',
                 device=0):
        super().__init__()
        self.x = Dimension(x_samples, x_delta)
        self.t = Dimension(t_samples, t_delta)
        self.form = form
        self.device = device
        self.add_material_region(self.get_line_region((0, max(self.x.vector)), name='main'),
                                 material)

    @property
    def num_points(self):
        return self.x.samples

    def d_x(self, factors=None, variant='forward'):
        if factors is None:
            factors = np.array(1).repeat(self.num_points)
        return _first_difference(factors, self.num_points, 1, variant)

    def d_x2(self, factors=None):
        if factors is None:
            factors = np.array(1).repeat(self.num_points)
        return _second_difference(factors, self.num_points, 1)

    def get_index(self, position):
        return self.x.get_index(position)

    def get_position(self, index):
        return self.x.vector[index]

    def get_line_region(self, position, name=''):
        """Inclusive index range between two x coordinates (``pyfds/fields.py:228-241``)."""
        return reg.LineRegion(np.arange(self.get_index(position[0]),
                                        self.get_index(position[1]) + 1, dtype=np.int64),
                              position, name=name)


class Field2D(Field):
    """Two-dimensional field, flat index ``ix + nx*iy`` (``pyfds/fields.py:244-533``)."""

    def __init__(self, x_samples, x_delta, y_samples, y_delta, t_samples, t_delta, material,
                 *, form='fp64', device=0):
        super().__init__()
        self.x = Dimension(x_samples, x_delta)
        self.y = Dimension(y_samples, y_delta)
        self.t = Dimension(t_samples, t_delta)
        self.form = form
        self.device = device
        self.add_material_region(self.get_rect_region(
            (0, 0, max(self.x.vector), max(self.y.vector)), name='main'), material)

    @property
    def num_points(self):
        return self.x.samples * self.y.samples

    def _fill(self, vec, indices, value):
        """``vec[indices] = value``.  Rectangles from ``get_rect_region`` (the main region among
        them) list their cells column by column, so a plain scatter strides through memory by a
        whole row per element (0.47 s for 4096^2 cells); when the indices are exactly such a
        rectangle, the same cells are written as a 2D slice instead."""
        idx = indices
        nx, ny = self.x.samples, self.y.samples
        if isinstance(idx, np.ndarray) and idx.ndim == 1 and idx.size >= 1 << 16 and \
                idx.dtype.kind in 'iu' and nx > 1:
            first = int(idx[0])
            # rows of the rectangle: the first column steps by nx until it ends
            col = idx[:min(idx.size, ny) + 1] - first
            steps = np.nonzero(col != np.arange(col.size) * nx)[0]
            h = int(steps[0]) if steps.size else col.size
            w = idx.size // h if h else 0
            x0, y0 = first % nx, first // nx
            if 0 < h and w * h == idx.size and x0 + w <= nx and y0 + h <= ny:
                expect = (first + np.arange(w, dtype=idx.dtype)[:, None]
                          + np.arange(h, dtype=idx.dtype)[None, :] * nx).ravel()
                if np.array_equal(idx, expect):
                    vec.reshape(ny, nx)[y0:y0 + h, x0:x0 + w] = value
                    return
        vec[idx] = value

    def d_x(self, factors=None, variant='forward'):
        if factors is None:
            factors = np.array(1).repeat(self.num_points)
        return _first_difference(factors, self.num_points, 1, variant)

    def d_y(self, factors=None, variant='forward'):
        if factors is None:
            factors = np.array(1).repeat(self.num_points)
        return _first_difference(factors, self.num_points, self.x.samples, variant)

    def d_x2(self, factors=None):
        if factors is None:
            factors = np.array(1).repeat(self.num_points)
        return _second_difference(factors, self.num_points, 1)

    def d_y2(self, factors=None):
        if factors is None:
            factors = np.array(1).repeat(self.num_points)
        return _second_difference(factors, self.num_points, self.x.samples)

    def get_index(self, position):
        return self.x.get_index(position[0]) + self.y.get_index(position[1]) * self.x.samples

    def get_position(self, index):
        return self.x.vector[index % self.x.samples], self.y.vector[int(index / self.x.samples)]

    def _positions(self, flat):
        """Coordinates of many flat indices (vectorised ``get_position``)."""
        nx = self.x.samples
        return self.x.vector[flat % nx], self.y.vector[flat // nx]

    def get_line_region(self, position, name=''):
        """Rasterised line between two points, inclusive (``pyfds/fields.py:391-418``)."""
        nx = self.x.samples
        start_idx = self.get_index(position[:2])
        end_idx = self.get_index(position[2:])
        x_diff = start_idx % nx - end_idx % nx
        y_diff = int(start_idx / nx) - int(end_idx / nx)
        num_points = max(np.abs([x_diff, y_diff]))
        if num_points == 0:
            # the reference evaluates 0/0 -> nan and fails converting it to an index
            raise ValueError('cannot convert float NaN to integer')
        ii = np.arange(num_points + 1)
        frac = ii / num_points
        x_pos = start_idx % nx - np.round(frac * x_diff)
        y_pos = int(start_idx / nx) - np.round(frac * y_diff)
        return reg.LineRegion((x_pos + nx * y_pos).astype(np.int64), position, name=name)

    def get_rect_region(self, position, name=''):
        """Inclusive rectangle (origin, size); indices ordered x-major like the reference's
        nested comprehension (``pyfds/fields.py:420-441``)."""
        x_start = self.x.get_index(position[0])
        y_start = self.y.get_index(position[1])
        x_end = self.x.get_index(position[0] + position[2])
        y_end = self.y.get_index(position[1] + position[3])
        x_start, x_end = min(x_start, x_end), max(x_start, x_end)
        y_start, y_end = min(y_start, y_end), max(y_start, y_end)
        xs = np.arange(x_start, x_end + 1, dtype=np.int64)
        ys = np.arange(y_start, y_end + 1, dtype=np.int64)
        return reg.RectRegion((xs[:, None] + ys[None, :] * self.x.samples).ravel(),
                              position, name)

    def get_tri_region(self, position, name=''):
        """Points inside a triangle, inclusive (``pyfds/fields.py:443-502``)."""
        if (position[2] - position[0]) * (position[5] - position[1]) - \
                (position[3] - position[1]) * (position[4] - position[0]) > 0:
            position = (position[0], position[1], position[4], position[5],
                        position[2], position[3])
        p = position
        corners = (self.get_index(p[:2]), self.get_index(p[2:4]), self.get_index(p[4:]))
        flat = np.arange(min(corners), max(corners) + 1, dtype=np.int64)
        px, py = self._positions(flat)
        e01 = (px - p[0]) * (p[3] - p[1]) - (py - p[1]) * (p[2] - p[0])
        e12 = (px - p[2]) * (p[5] - p[3]) - (py - p[3]) * (p[4] - p[2])
        e20 = (px - p[4]) * (p[1] - p[5]) - (py - p[5]) * (p[0] - p[4])
        return reg.TriRegion(flat[(e01 >= 0) & (e12 >= 0) & (e20 >= 0)], position, name)

    def get_ellipse_region(self, centre, radii, name=''):
        """Points inside an ellipse (``pyfds/fields.py:504-533``); like the reference the
        scanned flat range stops before the top point's index."""
        if np.isscalar(radii):
            radii = (radii, radii)
        lo_idx = self.get_index((centre[0], centre[1] - radii[1]))
        hi_idx = self.get_index((centre[0], centre[1] + radii[1]))
        flat = np.arange(lo_idx, hi_idx, dtype=np.int64)
        px, py = self._positions(flat)
        inside = (centre[0] - px) ** 2 / radii[0] ** 2 + (centre[1] - py) ** 2 / radii[1] ** 2 <= 1
        return reg.EllipseRegion(flat[inside], centre, radii, name)


class Dimension:
    """A space or time axis (``pyfds/fields.py:536-571``)."""

    def __init__(self, samples, increment):
        self.samples = int(samples)
        self.increment = increment
        self.snap_radius = np.finfo(float).eps * 10

    @property
    def vector(self):
        return np.arange(start=0, stop=self.samples) * self.increment

    def get_index(self, value):
        """Index of the sample within ``snap_radius`` of ``value`` (asserts exactly one)."""
        index, = np.where(np.abs(self.vector - value) <= self.snap_radius)
        assert len(index) < 2, "Multiple points found within snap radius of given value."
        assert len(index) > 0, "No point found within snap radius of given value."
        return int(index[0])


class FieldComponent:
    """One component of a field (``pyfds/fields.py:574-633``).

    ``values`` is backed by the device after a run: reading it downloads (once) and marks the
    host copy as the one to upload before the next run, so in-place edits are honoured.

    ``in_place``: the reference updates this component in place (``values -= ...``), so an
    array obtained before a run holds the new state afterwards; the download then goes into
    that same array (when it is a writeable float64 array of the right size).  Components the
    reference assigns anew each step (thermal fluxes, the nonlinear models' pressure) get a
    new array, as there.
    """

    def __init__(self, num_points, in_place=True):
        self._values = np.zeros(num_points)
        self._device = None          # (runner, component index) holding newer values
        self._host_dirty = True      # host copy must be uploaded before the next run
        self._in_place = in_place
        self.boundaries = []
        self.outputs = []

    @property
    def num_points(self):
        return len(self._values)

    @property
    def values(self):
        if self._device is not None:
            runner, k = self._device
            self._device = None
            v = self._values
            if self._in_place and isinstance(v, np.ndarray) and v.dtype == np.float64 \
                    and v.shape == (runner.n,) and v.flags.c_contiguous and v.flags.writeable:
                runner.download(k, out=v)
            else:
                self._values = runner.download(k)
        self._host_dirty = True
        return self._values

    @values.setter
    def values(self, new_values):
        self._device = None
        self._host_dirty = True
        self._values = new_values

    def apply_bounds(self, step):
        """Apply the boundaries to ``values`` on the host (standalone-component API of
        ``pyfds/fields.py:591-600``; field time stepping applies them on the device)."""
        vals = self.values
        for bound in self.boundaries:
            vals[bound.region.indices] = bound.apply(vals[bound.region.indices], step=step)

    def write_outputs(self):
        """Append the current values to the outputs (standalone-component API of
        ``pyfds/fields.py:602-611``)."""
        vals = self.values
        for output in self.outputs:
            if not output.signals:
                output.signals = [[vals[index]] for index in output.region.indices]
            else:
                for index, signal in zip(output.region.indices, output.signals):
                    signal.append(vals[index])

    def add_boundary(self, *args, **kwargs):
        new_bound = reg.Boundary(*args, **kwargs)
        self.boundaries.append(new_bound)
        logger.info('Boundary {} added.'.format(new_bound.region.name))

    def add_output(self, *args, **kwargs):
        new_output = reg.Output(*args, **kwargs)
        self.outputs.append(new_output)
        logger.info('Output region {} added.'.format(new_output.region.name))


class ProgressLogger:
    """Logs progress every ``log_increment`` percent (``pyfds/fields.py:636-668``)."""

    def __init__(self, num_steps, log_increment=5, logger_instance=None):
        self.num_steps = num_steps
        self.log_increment = log_increment
        self.logger = logger_instance if logger_instance else logger
        self._last_message_at = None

    def log(self, current_step):
        percent = int(current_step / self.num_steps * 100)
        if int(current_step / self.num_steps * 100 % self.log_increment) == 0 and \
                self._last_message_at != percent:
            self.logger.info('Simulating. {} % completed.'.format(percent))
            self._last_message_at = percent
