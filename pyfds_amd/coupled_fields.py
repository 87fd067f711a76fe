"""Thermo-acoustic co-simulation (drop-in for ``pyfds.coupled_fields``,
``pyfds/coupled_fields.py:10-65``).

``simulate`` runs the whole coupled step loop on the device (``fds_couple_loss_run``: one
workgroup steps the acoustic field, the thermal field and the loss transfer together) when the
object is in the configuration the reference's constructor builds: its own acoustic-loss
coupling as the only interaction, both fields' own ``sim_step`` and ``material_vector``, one
form, no tile shape forced on a field.  Any other configuration (a replaced interaction list or transfer
function, a subclass overriding ``sim_step``, a field forced onto a tile shape, ...) runs the
reference's host protocol over the device-stepped fields, step by step.
"""
import ctypes as ct

import numpy as np

from . import _native as nat
from . import acoustics as ac
from . import coupling as cp
from . import fields as fld
from . import thermal as th

__all__ = [
    'ThermoAcoustic1D',
]

# longest line the device-coupled loop takes (the resident 1D engine's limit)
COUPLED_MAX_CELLS = 1 << 24


class ThermoAcoustic1D(cp.SynchronizedFields):
    """An Acoustic1D and a Thermal1D field on the same grid; the acoustic losses heat the
    thermal field every ``stepping``-th step (accumulated in between)."""

    def __init__(self, x_samples, x_delta, t_samples, t_delta,
                 thermal_material, acoustic_material, stepping=1):
        acoustic_field = ac.Acoustic1D(x_samples, x_delta, t_samples, t_delta, acoustic_material)
        thermal_field = th.Thermal1D(x_samples, x_delta, t_samples, t_delta, thermal_material)
        acoustic_loss = cp.BoundaryCoupling(
            source_component=acoustic_field.velocity,
            target_component=thermal_field.temperature,
            transfer_function=self._loss_coupling,
            additive=True,
            accumulate=stepping > 1,
            stepping=stepping)
        super().__init__([acoustic_field, thermal_field], [acoustic_loss])
        self.get_index = acoustic_field.get_index
        self.get_position = acoustic_field.get_position
        self.get_line_region = acoustic_field.get_line_region

    def _loss_coupling(self, velocity):
        """Heat added per step by viscous loss: mu / rho_t / c_p * (dv/dx)^2 * dt, with dv/dx
        from the acoustic field's velocity-to-pressure operator with its density factor
        undone (multiplying v by rho first), in the reference's evaluation order
        (``pyfds/coupled_fields.py:47-65``)."""
        acoustic, thermal = self.fields
        derivative = acoustic.a_v_p.dot(velocity * acoustic.material_vector('density'))
        loss = self._loss_factor()
        return loss * derivative ** 2 * self.t.increment

    def _loss_factor(self):
        acoustic, thermal = self.fields
        return acoustic.material_vector('absorption_coef') \
            / thermal.material_vector('density') \
            / thermal.material_vector('heat_capacity')

    # -- device-coupled step loop ----------------------------------------------------------
    def _device_coupling(self):
        """The loss coupling when the whole loop may run on the device, else None."""
        if type(self).sim_step is not cp.SynchronizedFields.sim_step:
            return None
        if type(self)._loss_coupling is not ThermoAcoustic1D._loss_coupling:
            return None
        if len(self.fields) != 2 or len(self.interactions) != 1:
            return None
        acoustic, thermal = self.fields
        coupling = self.interactions[0]
        tf = coupling.transfer_function
        if type(coupling) is not cp.BoundaryCoupling or \
                getattr(tf, '__func__', None) is not ThermoAcoustic1D._loss_coupling or \
                getattr(tf, '__self__', None) is not self:
            return None
        if coupling.source_component is not getattr(acoustic, 'velocity', None) or \
                coupling.target_component is not getattr(thermal, 'temperature', None) or \
                coupling.additive is not True or coupling.accumulate not in (True, False):
            return None
        if not isinstance(coupling.stepping, (int, np.integer)) or coupling.stepping < 1:
            return None
        if type(acoustic) is not ac.Acoustic1D or type(thermal) is not th.Thermal1D:
            return None
        for f in (acoustic, thermal):
            if not f._device_loop_ok() or 'material_vector' in vars(f):
                return None
            if getattr(f, 'tile', None) is not None or getattr(f, 'strip_height', None):
                return None
        if getattr(acoustic, 'form', 'fp64') != getattr(thermal, 'form', 'fp64'):
            return None
        if acoustic.num_points != thermal.num_points or acoustic.num_points > COUPLED_MAX_CELLS:
            return None
        return coupling

    def simulate(self, num_steps=None):
        """``Field.simulate`` (``pyfds/fields.py:67-95``); the loop runs on the device when the
        configuration allows (see the module docstring)."""
        coupling = self._device_coupling()
        if coupling is None:
            return super().simulate(num_steps)
        if not num_steps:
            num_steps = self.t.samples
            progress_logger = fld.ProgressLogger(num_steps)
        else:
            progress_logger = None
        if not self.matrices_assembled:
            self.assemble_matrices()
            fld.logger.info('Matrices created.')
        fld.logger.info('Starting simulation of {} steps.'.format(num_steps))
        start = self.step
        self._device_run(coupling, start, num_steps, progress_logger)
        # a boundary signal that ends inside the run: the host protocol reaches the same
        # step and raises the reference's IndexError there
        while self.step < start + num_steps:
            self.sim_step()
            if progress_logger:
                progress_logger.log(self.step)
            self.step += 1
        fld.logger.info('Simulation of {} steps completed.'.format(num_steps))

    def _device_run(self, coupling, start, num_steps, progress):
        acoustic, thermal = self.fields
        ra, rt = acoustic._runner(), thermal._runner()
        ra.sync(acoustic)
        rt.sync(thermal)
        valid = [v for v in (ra._valid_steps(), rt._valid_steps()) if v is not None]
        n_ok = num_steps if not valid else max(0, min(num_steps, min(valid) - start))
        if n_ok == 0:
            return 0
        n = acoustic.num_points
        rho = nat.f64(acoustic.material_vector('density'))
        loss = nat.f64(np.broadcast_to(self._loss_factor(), (n,)))
        acc_in = coupling.accumulated_transfer
        acc = nat.f64(np.broadcast_to(np.asarray(acc_in, dtype=float), (n,)).copy())
        seg = max(1, min(ra._segment(n_ok, progress), rt._segment(n_ok, progress)))
        lib = ra.lib
        done = 0
        while done < n_ok:
            k = min(seg, n_ok - done)
            step = self.step
            nat.check(lib.fds_couple_loss_run(
                ra.ctx, rt.ctx, nat.dptr(rho), nat.dptr(loss), ct.c_double(self.t.increment),
                int(coupling.stepping), 1 if coupling.accumulate else 0, nat.dptr(acc),
                step, k))
            ra.collect(acoustic, k)
            rt.collect(thermal, k)
            if progress is not None:
                for s in range(step, step + k):
                    progress.log(s)
            self.step = step + k
            done += k
        if coupling.accumulate:
            coupling.accumulated_transfer = acc
        return done
