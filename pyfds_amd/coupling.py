"""Coupled fields (drop-in for ``pyfds.coupling``, ``pyfds/coupling.py:14-300``).

``SynchronizedFields`` advances several fields of equal spatial and temporal resolution one
step at a time and applies its interactions after every step.  The interactions act on
``FieldComponent.values`` (``BoundaryCoupling``) or re-scale a material parameter and
re-assemble the target field (``MaterialCoupling`` and its exponential / power-law forms).

Two ways to run it:

* **on the device** (``simulate``, when every interaction is a shipped one the device knows:
  ``MaterialCouplingExponential``, ``MaterialCouplingPowerLaw``, or a ``BoundaryCoupling``
  whose transfer is a :class:`LinearTransfer`; fields of one form, FP64 or FP32, on one
  device): each step runs every field's step kernel and then the interactions as element-wise
  kernels (``fds_coupling_run``), with no host round trip until the run ends.  Afterwards the
  interactions' state (``last_used_factors``, ``accumulated_transfer``), the re-assembled
  targets' coefficients and the re-assembly log lines are brought back to the host;
* **the host protocol** otherwise (any other transfer function, ``sim_step`` calls): each
  field steps on the device; ``values`` reads download the device state and writes are
  uploaded before the field's next step, so the host-side interactions see and change exactly
  what the reference's do.
"""
import ctypes as ct
import logging as lo

import numpy as np

from . import _native as nat
from . import fields as fld

__all__ = [
    'SynchronizedFields', 'BoundaryCoupling', 'MaterialCoupling', 'MaterialCouplingExponential',
    'MaterialCouplingPowerLaw', 'LinearTransfer',
]

logger = lo.getLogger('pyfds')


class SynchronizedFields(fld.Field):
    """Fields stepped together, interactions applied after each step
    (``pyfds/coupling.py:14-90``)."""

    def __init__(self, fields, interactions):
        self.fields = fields
        self.interactions = interactions
        # expose every component of every field under its own name (the gfx animator looks
        # components up on the simulated object); a name used twice is not supported
        for field in self.fields:
            for name, value in vars(field).items():
                if not isinstance(value, fld.FieldComponent):
                    continue
                if hasattr(self, name):
                    raise RuntimeError('Coupling of fields with identically named components '
                                       'is currently not possible')
                setattr(self, name, value)
        first = self.fields[0]
        self.t = first.t
        self.x = first.x
        if hasattr(first, 'y'):
            self.y = first.y

    @property
    def step(self):
        return self.fields[0].step

    @step.setter
    def step(self, value):
        for field in self.fields:
            field.step = value

    @property
    def num_points(self):
        return self.fields[0].num_points

    @property
    def material_regions(self):
        """All fields' material regions, concatenated (read only: add regions to the
        individual fields)."""
        out = []
        for field in self.fields:
            out += field.material_regions
        return out

    def assemble_matrices(self):
        for field in self.fields:
            field.assemble_matrices()

    @property
    def matrices_assembled(self):
        return all(field.matrices_assembled for field in self.fields)

    def sim_step(self):
        """Step every field, then apply every interaction at the current step."""
        for field in self.fields:
            field.sim_step()
        for interaction in self.interactions:
            interaction.apply(self.step)

    # -- device loop ----------------------------------------------------------------------
    def simulate(self, num_steps=None):
        """``Field.simulate`` (``pyfds/fields.py:67-95``): on the device when the fields and
        interactions allow it (see the module docstring), else one host-protocol step at a
        time."""
        plan = _device_plan(self)
        if plan is None:
            return super().simulate(num_steps)
        if not num_steps:
            num_steps = self.t.samples
            progress_logger = fld.ProgressLogger(num_steps)
        else:
            progress_logger = None
        if not self.matrices_assembled:
            self.assemble_matrices()
            logger.info('Matrices created.')
        logger.info('Starting simulation of {} steps.'.format(num_steps))
        _run_on_device(self, plan, num_steps, progress_logger)
        logger.info('Simulation of {} steps completed.'.format(num_steps))


class BoundaryCoupling:
    """Feed ``transfer_function(source.values)`` into ``target.values``, added or assigned,
    every ``stepping``-th step, optionally accumulating the transfer in between
    (``pyfds/coupling.py:93-146``)."""

    def __init__(self, source_component, target_component, transfer_function,
                 additive=True, accumulate=False, stepping=1):
        self.source_component = source_component
        self.target_component = target_component
        self.transfer_function = transfer_function
        self.additive = additive
        self.accumulate = accumulate
        self.stepping = stepping
        self.accumulated_transfer = 0

    def _deliver(self, transfer):
        if self.additive is True:
            self.target_component.values += transfer
        else:
            self.target_component.values = transfer

    def apply(self, step):
        if self.accumulate is True:
            self.accumulated_transfer += self.transfer_function(self.source_component.values)
        if step % self.stepping != 0:
            return
        if self.accumulate is False:
            self._deliver(self.transfer_function(self.source_component.values))
        else:
            self._deliver(self.accumulated_transfer)
            self.accumulated_transfer = 0


class MaterialCoupling:
    """Multiply one material parameter of ``target_field`` by
    ``transfer_function(source.values)`` and re-assemble the target every ``stepping``-th step
    when the factors moved by more than ``rel_change_threshold`` (``pyfds/coupling.py:149-226``).

    The target's ``material_vector`` is wrapped (the original stays reachable as
    ``static_material_vector``), so every later assembly of the target sees the coupled
    parameter; a re-assembly re-uploads the target's coefficients before its next step.
    """

    def __init__(self, source_component, target_field, target_parameter,
                 transfer_function, rel_change_threshold=None, stepping=1):
        self.source_component = source_component
        self.target_field = target_field
        self.target_parameter = target_parameter
        self.transfer_function = transfer_function
        self.rel_change_threshold = rel_change_threshold
        self.stepping = stepping
        self.last_used_factors = 0
        self.target_field.static_material_vector = self.target_field.material_vector
        self.target_field.material_vector = self._material_vector

    def _material_vector(self, mat_parameter):
        base = self.target_field.static_material_vector(mat_parameter)
        if mat_parameter != self.target_parameter:
            return base
        return base * self.transfer_function(self.source_component.values)

    def apply(self, step):
        if step % self.stepping != 0:
            return
        factors = self.transfer_function(self.source_component.values)
        rel_change = max(abs((factors - self.last_used_factors) / factors))
        if self.rel_change_threshold is not None and not rel_change > self.rel_change_threshold:
            return
        self.target_field.assemble_matrices()     # sees the wrapped material_vector
        self.last_used_factors = factors
        if self.rel_change_threshold is not None:
            logger.info(f"Relative change in parameters is {rel_change}.")
            logger.info(f"Matrices reassembled in step {step}.")


class MaterialCouplingExponential(MaterialCoupling):
    """Factor ``a + (1 - a) * exp(b * q)`` (``pyfds/coupling.py:229-270``)."""

    def __init__(self, source_component, target_field, target_parameter, a, b,
                 rel_change_threshold=None, stepping=1):
        self.a = a
        self.b = b
        super().__init__(source_component=source_component, target_field=target_field,
                         target_parameter=target_parameter,
                         transfer_function=self.transfer_function,
                         rel_change_threshold=rel_change_threshold, stepping=stepping)

    def transfer_function(self, values):
        return self.a + (1 - self.a) * np.exp(self.b * values)


class MaterialCouplingPowerLaw(MaterialCoupling):
    """Factor ``1 + factor * q ** power`` (``pyfds/coupling.py:273-300``)."""

    def __init__(self, source_component, target_field, target_parameter, power, factor,
                 rel_change_threshold=None, stepping=1):
        self.power = power
        self.factor = factor
        super().__init__(source_component=source_component, target_field=target_field,
                         target_parameter=target_parameter,
                         transfer_function=self.transfer_function,
                         rel_change_threshold=rel_change_threshold, stepping=stepping)

    def transfer_function(self, values):
        return 1 + self.factor * values ** self.power


class LinearTransfer:
    """``transfer(values) = factor * values``: a ``BoundaryCoupling`` transfer function that
    the device loop of :class:`SynchronizedFields` applies itself."""

    def __init__(self, factor):
        self.factor = float(factor)

    def __call__(self, values):
        return self.factor * values


# material vectors a re-assembly reads, per model family (FDS_P_* order)
_ACOUSTIC_PARAMS = ('sound_velocity', 'density', 'absorption_coef')
_THERMAL_PARAMS = {'density': 1, 'heat_capacity': 3, 'thermal_conductivity_x': 4,
                   'thermal_conductivity_y': 5}
_KINDS = {'exp': 1, 'power': 2, 'linear': 3}


def _family(field):
    mid = getattr(field, '_model_id', None)
    if mid in (nat.ACOUSTIC1D, nat.ACOUSTIC2D, nat.ACOUSTIC_AXI):
        return 'acoustic'
    if mid in (nat.THERMAL1D, nat.THERMAL2D, nat.THERMAL_AXI):
        return 'thermal'
    return None


def _component_of(fields, comp):
    for i, f in enumerate(fields):
        for k, name in enumerate(f._component_names):
            if name and getattr(f, name) is comp:
                return i, k
    return None


def _y_slots_consistent(field, family):
    """Re-assembly keeps a y coefficient array aliased to its x array only when both have the
    same factor: the field must store the y arrays whenever dx != dy."""
    if not hasattr(field, 'y'):
        return True
    dt, dx, dy = field.t.increment, field.x.increment, field.y.increment
    pairs = ([(0, 1, dt / dx, dt / dy), (2, 3, dt / dx, dt / dy)] if family == 'acoustic'
             else [(0, 1, dt / dx, dt / dy), (2, 3, 1 / dx, 1 / dy)])
    for xs, ys, sx, sy in pairs:
        if field._coefs.get(ys) is None and sx != sy:
            return False
    return True


def _device_plan(sync):
    """The interaction descriptors of the device loop, or None when the host protocol must
    run (an unknown transfer function or field, a mixed form, the Kahan form)."""
    fields = list(sync.fields)
    if not fields:
        return None
    forms = {getattr(f, 'form', 'fp64') for f in fields}
    devices = {int(getattr(f, 'device', 0)) for f in fields}
    if len(forms) != 1 or forms & {'fp32_kahan'} or len(devices) != 1:
        return None
    for f in fields:
        if _family(f) is None or not f._device_loop_ok() \
                or getattr(f, '_device_flow_periods', None) is not None \
                or getattr(f, 'learned_boundary', None) is not None \
                or f.num_points != fields[0].num_points:
            return None
    descs, targets = [], set()
    for it in sync.interactions:
        src = _component_of(fields, getattr(it, 'source_component', None))
        if src is None:
            return None
        d = dict(source=src, stepping=int(it.stepping), threshold=float('nan'),
                 additive=0, accumulate=0, target_comp=-1, param=-1, p0=0.0, p1=0.0)
        if type(it) in (MaterialCouplingExponential, MaterialCouplingPowerLaw):
            if it.transfer_function.__func__ is not type(it).transfer_function:
                return None
            if it.target_field not in fields or id(it.target_field) in targets:
                return None
            targets.add(id(it.target_field))
            fam = _family(it.target_field)
            if fam == 'acoustic':
                if it.target_parameter not in _ACOUSTIC_PARAMS:
                    return None
                param = _ACOUSTIC_PARAMS.index(it.target_parameter)
                if it.target_field.x.samples == 1 and hasattr(it.target_field, 'y'):
                    return None
            else:
                if it.target_parameter not in _THERMAL_PARAMS:
                    return None
                param = _THERMAL_PARAMS[it.target_parameter]
            if it.rel_change_threshold is not None:
                d['threshold'] = float(it.rel_change_threshold)
            exp = type(it) is MaterialCouplingExponential
            d.update(kind=_KINDS['exp' if exp else 'power'],
                     target=fields.index(it.target_field), param=param,
                     p0=float(it.a if exp else it.power), p1=float(it.b if exp else it.factor))
        elif type(it) is BoundaryCoupling and isinstance(it.transfer_function, LinearTransfer):
            tgt = _component_of(fields, it.target_component)
            if tgt is None or it.additive not in (True, False) \
                    or it.accumulate not in (True, False):
                return None
            d.update(kind=_KINDS['linear'], target=tgt[0], target_comp=tgt[1],
                     p0=it.transfer_function.factor, additive=int(it.additive is True),
                     accumulate=int(it.accumulate is True))
        else:
            return None
        if d['stepping'] < 1:
            return None
        d['interaction'] = it
        descs.append(d)
    return descs


class _Interaction(ct.Structure):
    _fields_ = [('kind', ct.c_int), ('source_field', ct.c_int), ('source_comp', ct.c_int),
                ('target_field', ct.c_int), ('target_comp', ct.c_int),
                ('target_param', ct.c_int), ('p0', ct.c_double), ('p1', ct.c_double),
                ('stepping', ct.c_int64), ('threshold', ct.c_double), ('additive', ct.c_int),
                ('accumulate', ct.c_int)]


def _static_vectors(field):
    """The target's material vectors without the coupling (``static_material_vector``, the
    reference's original ``material_vector``) and the expression scalars of its assembly."""
    mv = field.static_material_vector
    dt, dx = field.t.increment, field.x.increment
    dy = field.y.increment if hasattr(field, 'y') else dx
    if _family(field) == 'acoustic':
        vecs = {i: mv(name) for i, name in enumerate(_ACOUSTIC_PARAMS)}
        scalars = [dt / dx, dt / dy, dt / dx ** 2, dt / dy ** 2, dt]
    else:
        vecs = {1: mv('density'), 3: mv('heat_capacity'), 4: mv('thermal_conductivity_x')}
        if hasattr(field, 'y'):
            vecs[5] = mv('thermal_conductivity_y')
        scalars = [dt / dx, dt / dy, 1 / dx, 1 / dy]
    return vecs, scalars


# the host's assembled arrays (``_mats`` names) of each model: (name, coefficient slot, slot
# it aliases when the field does not store its own)
_MATS_OF = {
    nat.ACOUSTIC1D: [('k', 0, None), ('b', 2, None), ('m', 4, None)],
    nat.ACOUSTIC2D: [('kx', 0, None), ('ky', 1, 0), ('bx', 2, None), ('by', 3, 2),
                     ('mx', 4, None), ('my', 5, 4)],
    nat.ACOUSTIC_AXI: [('kx', 0, None), ('ky', 1, 0), ('bx', 2, None), ('by', 3, 2),
                       ('mx', 4, None), ('my', 5, 4), ('mc', 6, None)],
    nat.THERMAL1D: [('g', 0, None), ('q', 2, None)],
    nat.THERMAL2D: [('gx', 0, None), ('gy', 1, 0), ('qx', 2, None), ('qy', 3, 2)],
    nat.THERMAL_AXI: [('gx', 0, None), ('gy', 1, 0), ('qx', 2, None), ('qy', 3, 2)],
}


def _refresh_coefficients(field):
    """After device re-assemblies: the device's coefficient arrays become the host's (the
    ``a_*`` matrices are rebuilt from them on access), without a re-upload."""
    r = field._runner()
    got = {}
    for slot in sorted(field._coefs):
        if field._coefs[slot] is None:
            continue
        arr = np.empty(field.num_points)
        nat.check(r.lib.fds_get_coeff(r.ctx, slot, nat.dptr(arr), arr.size))
        field._coefs[slot] = got[slot] = arr
    for name, slot, alias in _MATS_OF[field._model_id]:
        if slot in got:
            field._mats[name] = got[slot]
        elif alias is not None and alias in got:
            field._mats[name] = got[alias]
    field.__dict__.pop('_matrix_cache', None)


def _run_on_device(sync, descs, num_steps, progress):
    fields = list(sync.fields)
    lib = nat.load()
    runners = []
    for f in fields:
        if not f.matrices_assembled:
            f.assemble_matrices()
        r = f._runner()
        r.sync(f)
        runners.append(r)
    key = tuple(r.ctx.value for r in runners) + tuple(
        (id(d['interaction']), d['kind'], d['p0'], d['p1'], d['stepping'],
         str(d['threshold']), d['additive'], d['accumulate']) for d in descs)
    handle = sync.__dict__.get('_coupling')
    if handle is None or sync.__dict__.get('_coupling_key') != key:
        if handle is not None:
            lib.fds_coupling_destroy(handle)
        arr = (_Interaction * max(1, len(descs)))()
        for j, d in enumerate(descs):
            arr[j] = _Interaction(d['kind'], d['source'][0], d['source'][1], d['target'],
                                  d['target_comp'], d['param'], d['p0'], d['p1'],
                                  d['stepping'], d['threshold'], d['additive'], d['accumulate'])
        ctxs = (ct.c_void_p * len(fields))(*[r.ctx.value for r in runners])
        handle = ct.c_void_p()
        nat.check(lib.fds_coupling_create(ct.byref(handle), ctxs, len(fields), arr, len(descs)))
        sync.__dict__['_coupling'], sync.__dict__['_coupling_key'] = handle, key
    n = fields[0].num_points
    for j, d in enumerate(descs):
        it = d['interaction']
        if d['kind'] != _KINDS['linear']:
            vecs, scalars = _static_vectors(it.target_field)
            for p, v in vecs.items():
                v = nat.f64(np.broadcast_to(v, (n,)))
                nat.check(lib.fds_coupling_set_material(handle, d['target'], p, nat.dptr(v), n))
            s = nat.f64(scalars)
            nat.check(lib.fds_coupling_set_scalars(handle, d['target'], nat.dptr(s), s.size))
            state = it.last_used_factors
        else:
            state = it.accumulated_transfer
        state = nat.f64(np.broadcast_to(np.asarray(state, dtype=np.float64), (n,)))
        nat.check(lib.fds_coupling_state(handle, j, 1, nat.dptr(state), n))
    start = fields[0].step
    cap = sum(num_steps // d['stepping'] + 1 for d in descs if d['kind'] != _KINDS['linear'])
    log = np.zeros((max(cap, 1), 2))
    nlog = ct.c_int64()
    rc = lib.fds_coupling_run(handle, start, num_steps, nat.dptr(log), log.shape[0],
                              ct.byref(nlog))
    nat.check(rc)
    for f, r in zip(fields, runners):
        r.collect(f, num_steps)
        f.step = start + num_steps
    if progress is not None:
        for s_ in range(start, start + num_steps):
            progress.log(s_)
    reassembled = set()
    for j, d in enumerate(descs):
        it = d['interaction']
        out = np.empty(n)
        nat.check(lib.fds_coupling_state(handle, j, 0, nat.dptr(out), n))
        if d['kind'] == _KINDS['linear']:
            if it.accumulate is True:
                it.accumulated_transfer = out if np.any(out != 0) else 0
        else:
            if np.any(out != 0) or not np.isscalar(it.last_used_factors):
                it.last_used_factors = out
                reassembled.add(d['target'])
    for t in reassembled:
        _refresh_coefficients(fields[t])
    for step, rel in log[:min(nlog.value, log.shape[0])]:
        logger.info(f"Relative change in parameters is {np.float64(rel)}.")
        logger.info(f"Matrices reassembled in step {int(step)}.")
