"""Thermo-acoustic co-simulation (drop-in for ``pyfds.coupled_fields``,
``pyfds/coupled_fields.py:10-65``)."""
from . import acoustics as ac
from . import coupling as cp
from . import thermal as th

__all__ = [
    'ThermoAcoustic1D',
]


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
        loss = acoustic.material_vector('absorption_coef') \
            / thermal.material_vector('density') \
            / thermal.material_vector('heat_capacity')
        return loss * derivative ** 2 * self.t.increment
