"""Acoustics in a moving medium (drop-in for ``pyfds.acoustic_flow``,
``pyfds/acoustic_flow.py:12-57``)."""
import logging as lo
import warnings as wn

import numpy as np

from . import acoustics as acs

__all__ = [
    'AcousticFlow2D',
]

logger = lo.getLogger('pyfds')


class AcousticFlow2D(acs.Acoustic2D):
    """Acoustic2D whose rows are shifted one cell in +x every ``flow_t_deltas[row]`` steps
    (the reference's approximation of a flow along x).  ``simulate`` and ``sim_step`` shift on
    the device after every step; ``apply_flow`` is the same shift on the host values."""

    def __init__(self, flow, *args, **kwargs):
        super().__init__(*args, **kwargs)
        ny = self.y.samples
        if isinstance(flow, (list, np.ndarray)) and len(flow) == ny:
            self.flow = np.asarray(flow)
        elif isinstance(flow, (float, int)):
            self.flow = np.ones(ny) * flow
        else:
            raise ValueError('Flow must either be scalar or a vector with length of y_samples.')
        # steps between two shifts of a row
        self.flow_t_deltas = (self.x.increment / self.flow / self.t.increment).astype(int)
        if np.any(self.flow_t_deltas == 1) or np.any(self.flow_t_deltas == 0):
            wn.warn('Flow velocity may be to high. Consider reducing t_delta.', stacklevel=2)
            logger.warning('Flow velocity may be to high. Consider reducing t_delta.')

    def _device_sim_step(self):
        """One step followed by the flow shift, both on the device (the runner hands the
        shift periods to the library, which shifts after every step it runs)."""
        acs.Acoustic2D._device_sim_step(self)

    sim_step = _device_sim_step

    def _device_flow_periods(self):
        return self.flow_t_deltas

    def apply_flow(self):
        """Shift every row whose period divides the current step one cell to +x, entering a
        zero at the row start, for p, vx and vy."""
        nx = self.x.samples
        for component in (self.pressure, self.velocity_x, self.velocity_y):
            vals = component.values
            for n, f in enumerate(self.flow_t_deltas):
                if self.step % f == 0:
                    vals[n * nx + 1:(n + 1) * nx] = vals[n * nx:(n + 1) * nx - 1].copy()
                    vals[n * nx] = 0
