"""1D nonlinear acoustics (drop-in for ``pyfds.nonlinear_acoustics``).

Coefficients and buffered material parameters follow
``pyfds/nonlinear_acoustics.py:249-264, 344-358``; the step runs in the device's 1D
nonlinear kernel.
"""
import numpy as np

from . import _native as nat
from . import fields as fld
from .acoustics import AcousticMaterial, _DeviceModel

__all__ = [
    'IdealGas1D', 'Acoustic2ndOrder1D', 'AcousticMaterial2ndOrder',
]

_RHO0, _M, _EOS1, _EOS2, _CSQ = range(5)
_S_DTDX, _S_DTDX_4, _S_FLAGS = 2, 3, 4
_CONVECTIVE, _NL_STATE, _SECOND_ORDER = 1, 2, 4


class _Nonlinear1D(_DeviceModel, fld.Field1D):
    _model_id = nat.IDEALGAS1D
    _component_names = ('pressure', 'velocity', 'density')

    def _device_grid(self):
        return self.num_points, 1

    def _common(self):
        dt, dx = self.t.increment, self.x.increment
        self.stat_density = self.material_vector('density')
        m = dt / dx ** 2 * self.material_vector('absorption_coef')
        mats = dict(dv=dt / dx * np.ones(self.x.samples),
                    vp=(dt / dx) * np.ones(self.x.samples),
                    vv=m, vv2=(dt / dx / 2) * np.ones(self.x.samples))
        scalars = {_S_DTDX: dt / dx, _S_DTDX_4: (dt / dx / 2) / 2}
        return mats, {_RHO0: self.stat_density, _M: m}, scalars

    def _build_matrix(self, name):
        m = self._mats
        if name == 'a_d_v':
            return self.d_x(factors=m['dv'])
        if name == 'a_v_p':
            return self.d_x(factors=m['vp'], variant='backward')
        if name == 'a_v_v':
            return self.d_x2(factors=m['vv'])
        return self.d_x(factors=m['vv2'], variant='central')

    a_d_v = property(lambda self: self._matrix('a_d_v'))
    a_v_p = property(lambda self: self._matrix('a_v_p'))
    a_v_v = property(lambda self: self._matrix('a_v_v'))
    a_v_v2 = property(lambda self: self._matrix('a_v_v2'))


class IdealGas1D(_Nonlinear1D):
    """Nonlinear acoustics in ideal gases (``pyfds/nonlinear_acoustics.py:10-100``)."""

    def __init__(self, convective=True, nl_state=True, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.convective = convective
        self.nl_state = nl_state
        self.pressure = fld.FieldComponent(self.num_points, in_place=False)
        self.velocity = fld.FieldComponent(self.num_points)
        self.density = fld.FieldComponent(self.num_points)
        self.stat_density = None
        self.sound_velocity = None
        self.heat_cap_ratio = None

    def assemble_matrices(self):
        mats, coefs, scalars = self._common()
        self.sound_velocity = self.material_vector('sound_velocity')
        self.heat_cap_ratio = (self.material_vector('isobaric_heat_cap')
                               / self.material_vector('isochoric_heat_cap'))
        coefs[_EOS1] = self.stat_density * self.sound_velocity ** 2 / self.heat_cap_ratio
        coefs[_EOS2] = self.heat_cap_ratio
        coefs[_CSQ] = self.sound_velocity ** 2
        scalars[_S_FLAGS] = ((_CONVECTIVE if self.convective else 0)
                             | (_NL_STATE if self.nl_state else 0))
        self._store(mats, coefs, scalars)

    def is_stable(self):
        """CFL check with 1 % headroom (``pyfds/nonlinear_acoustics.py:92-100``)."""
        return np.all(self.material_vector('sound_velocity')
                      < 0.99 * self.x.increment / self.t.increment)


class Acoustic2ndOrder1D(_Nonlinear1D):
    """Second-order nonlinear acoustics (``pyfds/nonlinear_acoustics.py:103-171``)."""

    _model_id = nat.ACOUSTIC2ND1D

    def __init__(self, *args, **kwargs):
        if not (hasattr(kwargs['material'], 'd_rho_p')
                and hasattr(kwargs['material'], 'd_rho2_p')):
            raise RuntimeError('Material with properties d_rho_p and d_rho2_p is required for '
                               '2nd order nonlinear acoustic simulation.')
        super().__init__(*args, **kwargs)
        self.pressure = fld.FieldComponent(self.num_points, in_place=False)
        self.velocity = fld.FieldComponent(self.num_points)
        self.density = fld.FieldComponent(self.num_points)
        self.stat_density = None
        self.d_rho_p = None
        self.d_rho2_p = None

    def assemble_matrices(self):
        mats, coefs, scalars = self._common()
        self.d_rho_p = self.material_vector('d_rho_p')
        self.d_rho2_p = self.material_vector('d_rho2_p')
        coefs[_EOS1] = self.d_rho_p
        coefs[_EOS2] = self.d_rho2_p / 2
        scalars[_S_FLAGS] = _CONVECTIVE | _SECOND_ORDER
        self._store(mats, coefs, scalars)


class AcousticMaterial2ndOrder(AcousticMaterial):
    """Acoustic material with pressure-density derivatives
    (``pyfds/nonlinear_acoustics.py:174-190``).  ``d_rho_p`` defaults to
    ``sound_velocity**2`` (the reference's default path raises AttributeError instead)."""

    def __init__(self, d_rho_p=None, d_rho2_p=0, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.d_rho_p = d_rho_p if d_rho_p else self.sound_velocity ** 2
        self.d_rho2_p = d_rho2_p
