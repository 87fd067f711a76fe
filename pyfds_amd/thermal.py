"""Heat conduction (drop-in for ``pyfds.thermal``).

Coefficients follow ``pyfds/thermal.py:29-38, 75-90, 144-159`` term by term; the flux
components are recomputed (assigned, not accumulated) every step on the device and stored
to memory at the end of each run, which is exactly when the host can observe them.
"""
import numpy as np

from . import _native as nat
from . import fields as fld
from .acoustics import _DeviceModel, _alias

__all__ = [
    'Thermal1D', 'Thermal2D', 'Thermal3DAxi', 'ThermalMaterial'
]

_GX, _GY, _QX, _QY = range(4)


class Thermal1D(_DeviceModel, fld.Field1D):
    """One-dimensional heat conduction (``pyfds/thermal.py:12-51``)."""

    _model_id = nat.THERMAL1D
    _component_names = ('temperature', 'heat_flux', None)

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.temperature = fld.FieldComponent(self.num_points)
        self.heat_flux = fld.FieldComponent(self.num_points, in_place=False)

    def _device_grid(self):
        return self.num_points, 1

    def assemble_matrices(self):
        dt, dx = self.t.increment, self.x.increment
        g = dt / dx / self.material_vector('density') / self.material_vector('heat_capacity')
        q = 1 / dx * self.material_vector('thermal_conductivity_x')
        self._store(dict(g=g, q=q), {_GX: g, _QX: q})

    def _build_matrix(self, name):
        if name == 'a_t_q':
            return self.d_x(factors=self._mats['g'])
        return self.d_x(factors=self._mats['q'], variant='backward')

    a_t_q = property(lambda self: self._matrix('a_t_q'))
    a_q_t = property(lambda self: self._matrix('a_q_t'))


class Thermal2D(_DeviceModel, fld.Field2D):
    """Two-dimensional heat conduction (``pyfds/thermal.py:54-107``)."""

    _model_id = nat.THERMAL2D
    _component_names = ('temperature', 'heat_flux_x', 'heat_flux_y')
    _axi = False

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.temperature = fld.FieldComponent(self.num_points)
        self.heat_flux_x = fld.FieldComponent(self.num_points, in_place=False)
        self.heat_flux_y = fld.FieldComponent(self.num_points, in_place=False)

    def _device_grid(self):
        return self.x.samples, self.y.samples

    def _radii(self):
        """Distance of each heat-flux point from the axis (``pyfds/thermal.py:132-142``)."""
        return np.tile(self.x.vector, self.y.samples) + self.x.increment / 2

    def assemble_matrices(self):
        dt, dx, dy = self.t.increment, self.x.increment, self.y.increment
        rho = self.material_vector('density')
        cp = self.material_vector('heat_capacity')
        gx = dt / dx / rho / cp
        gy = dt / dy / rho / cp
        qx = 1 / dx * self.material_vector('thermal_conductivity_x')
        qy = 1 / dy * self.material_vector('thermal_conductivity_y')
        coefs = {}
        _alias(coefs, _GX, _GY, gx, gy)
        _alias(coefs, _QX, _QY, qx, qy)
        self._store(dict(gx=gx, gy=gy, qx=qx, qy=qy), coefs, {0: dx, 1: dx / 2})

    def _build_matrix(self, name):
        m = self._mats
        if name == 'a_t_qx':
            return self.d_x(factors=m['gx'] / self._radii() if self._axi else m['gx'])
        if name == 'a_t_qy':
            return self.d_y(factors=m['gy'])
        if name == 'a_qx_t':
            return self.d_x(factors=m['qx'], variant='backward')
        return self.d_y(factors=m['qy'], variant='backward')

    a_t_qx = property(lambda self: self._matrix('a_t_qx'))
    a_t_qy = property(lambda self: self._matrix('a_t_qy'))
    a_qx_t = property(lambda self: self._matrix('a_qx_t'))
    a_qy_t = property(lambda self: self._matrix('a_qy_t'))


class Thermal3DAxi(Thermal2D):
    """Axial-symmetric heat conduction, x is the radius (``pyfds/thermal.py:110-176``)."""

    _model_id = nat.THERMAL_AXI
    _axi = True


class ThermalMaterial:
    """Thermal material parameters (``pyfds/thermal.py:179-211``); the conductivity is a
    scalar or an (x, y) pair."""

    def __init__(self, heat_capacity, density, thermal_conductivity):
        self.heat_capacity = heat_capacity
        self.density = density
        self.thermal_conductivity_x = None
        self.thermal_conductivity_y = None
        self.thermal_conductivity = thermal_conductivity

    @property
    def thermal_conductivity(self):
        return self.thermal_conductivity_x, self.thermal_conductivity_y

    @thermal_conductivity.setter
    def thermal_conductivity(self, value):
        if isinstance(value, (list, tuple, np.ndarray)) and len(value) == 2:
            self.thermal_conductivity_x = value[0]
            self.thermal_conductivity_y = value[1]
        elif isinstance(value, (float, int)):
            self.thermal_conductivity_x = value
            self.thermal_conductivity_y = value
        else:
            raise ValueError('Thermal conductivity must either be scalar or a 2 element vector.')
