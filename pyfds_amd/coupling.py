"""Coupled fields (drop-in for ``pyfds.coupling``, ``pyfds/coupling.py:14-300``).

``SynchronizedFields`` advances several fields of equal spatial and temporal resolution one
step at a time and applies its interactions after every step.  The interactions act on
``FieldComponent.values`` (``BoundaryCoupling``) or re-scale a material parameter and
re-assemble the target field (``MaterialCoupling`` and its exponential / power-law forms).
Each field steps on the device; ``values`` reads download the device state and writes are
uploaded before the field's next step, so the host-side interactions see and change exactly
what the reference's do.
"""
import logging as lo

import numpy as np

from . import fields as fld

__all__ = [
    'SynchronizedFields', 'BoundaryCoupling', 'MaterialCoupling', 'MaterialCouplingExponential',
    'MaterialCouplingPowerLaw',
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
