"""Linear acoustics (drop-in for ``pyfds.acoustics``).

``assemble_matrices`` computes the per-cell coefficient arrays with the reference's
expressions and evaluation order (``pyfds/acoustics.py:27-38, 89-109, 178-203``) and hands
them to the device; ``sim_step``/``simulate`` run the matrix-free HIP step kernels.  The
``a_*`` sparse matrices stay available as attributes, built on first access from the same
coefficient arrays.  ``AcousticAS`` is an alias of ``Acoustic3DAxi``.
"""
import numpy as np

from . import _native as nat
from . import fields as fld

__all__ = [
    'Acoustic1D', 'Acoustic2D', 'Acoustic3DAxi', 'AcousticAS', 'AcousticMaterial',
]

_KX, _KY, _BX, _BY, _MX, _MY, _MC, _MT = range(8)


class _DeviceModel:
    """Mixin: coefficient bookkeeping, lazy ``a_*`` matrices and the device ``sim_step``."""

    _coef_version = 0
    _mats = None

    def _store(self, mats, coefs, scalars=None):
        self._mats = mats
        self._coefs = coefs
        self._scalars = scalars or {}
        self._coef_version = getattr(self, '_coef_version', 0) + 1
        self.__dict__.pop('_matrix_cache', None)
        self.matrices_assembled = True

    def _device_coefficients(self):
        return self._coefs, dict(self._scalars)

    def _matrix(self, name):
        if self._mats is None:
            return None
        cache = self.__dict__.setdefault('_matrix_cache', {})
        if name not in cache:
            cache[name] = self._build_matrix(name)
        return cache[name]

    def _device_sim_step(self):
        """One time step on the device at ``self.step`` (the step counter is advanced by
        ``simulate``, as in the reference)."""
        if not self.matrices_assembled:
            raise AttributeError("'NoneType' object has no attribute 'dot' "
                                 "(assemble_matrices() has not been called)")
        self._runner().single_step(self)

    sim_step = _device_sim_step


def _alias(coefs, x_slot, y_slot, x, y):
    """Store y only if it differs bitwise from x (the device then reads x for both)."""
    coefs[x_slot] = x
    if not np.array_equal(x, y):
        coefs[y_slot] = y


class Acoustic1D(_DeviceModel, fld.Field1D):
    """One-dimensional acoustics (``pyfds/acoustics.py:9-63``)."""

    _model_id = nat.ACOUSTIC1D
    _component_names = ('pressure', 'velocity', None)

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.pressure = fld.FieldComponent(self.num_points)
        self.velocity = fld.FieldComponent(self.num_points)

    def _device_grid(self):
        return self.num_points, 1

    def assemble_matrices(self):
        dt, dx = self.t.increment, self.x.increment
        rho = self.material_vector('density')
        k = dt / dx * self.material_vector('sound_velocity') ** 2 * rho
        b = dt / dx / rho
        m = dt / dx ** 2 * self.material_vector('absorption_coef') / rho
        coefs = {_KX: k, _BX: b}
        if np.any(m != 0):
            coefs[_MX] = m
        self._store(dict(k=k, b=b, m=m), coefs)

    def _build_matrix(self, name):
        m = self._mats
        if name == 'a_p_v':
            return self.d_x(factors=m['k'])
        if name == 'a_v_p':
            return self.d_x(factors=m['b'], variant='backward')
        return self.d_x2(factors=m['m'])

    a_p_v = property(lambda self: self._matrix('a_p_v'))
    a_v_p = property(lambda self: self._matrix('a_v_p'))
    a_v_v = property(lambda self: self._matrix('a_v_v'))

    def is_stable(self):
        """CFL check with 1 % headroom (``pyfds/acoustics.py:54-63``)."""
        return np.all(self.material_vector('sound_velocity')
                      < 0.99 * self.x.increment / self.t.increment)


class Acoustic2D(_DeviceModel, fld.Field2D):
    """Two-dimensional acoustics (``pyfds/acoustics.py:66-139``)."""

    _model_id = nat.ACOUSTIC2D
    _component_names = ('pressure', 'velocity_x', 'velocity_y')
    _axi = False

    def __init__(self, *args, **kwargs):
        super().__init__(*args, **kwargs)
        self.pressure = fld.FieldComponent(self.num_points)
        self.velocity_x = fld.FieldComponent(self.num_points)
        self.velocity_y = fld.FieldComponent(self.num_points)

    def _device_grid(self):
        return self.x.samples, self.y.samples

    def _radii(self):
        """Distance of each velocity point from the axis (``pyfds/acoustics.py:166-176``)."""
        return np.tile(self.x.vector, self.y.samples) + self.x.increment / 2

    def assemble_matrices(self):
        dt, dx, dy = self.t.increment, self.x.increment, self.y.increment
        c = self.material_vector('sound_velocity')
        rho = self.material_vector('density')
        mu = self.material_vector('absorption_coef')
        kx = dt / dx * c ** 2 * rho
        ky = dt / dy * c ** 2 * rho
        bx = dt / dx / rho
        by = dt / dy / rho
        mx = dt / dx ** 2 * mu / rho
        my = dt / dy ** 2 * mu / rho
        coefs = {}
        _alias(coefs, _KX, _KY, kx, ky)
        _alias(coefs, _BX, _BY, bx, by)
        mats = dict(kx=kx, ky=ky, bx=bx, by=by, mx=mx, my=my)
        lossy = bool(np.any(mx != 0) or np.any(my != 0))
        if self._axi:
            mats['mc'] = dt / dx * mu / rho
            lossy = lossy or bool(np.any(mats['mc'] != 0))
        if lossy:
            if self.x.samples == 1:
                # x and y neighbours coincide: scipy merges the two diagonals into one
                coefs[_MX] = mx + my
                coefs[_MY] = np.zeros_like(my)
            else:
                _alias(coefs, _MX, _MY, mx, my)
            if self._axi:
                coefs[_MC] = mats['mc']
                coefs[_MT] = dt * mu / rho
        scalars = {0: dx, 1: dx / 2}
        self._store(mats, coefs, scalars)

    def _build_matrix(self, name):
        m = self._mats
        r = self._radii() if self._axi else 1
        if name == 'a_p_vx':
            return self.d_x(factors=m['kx'] / r if self._axi else m['kx'])
        if name == 'a_p_vy':
            return self.d_y(factors=m['ky'])
        if name == 'a_vx_p':
            return self.d_x(factors=m['bx'], variant='backward')
        if name == 'a_vy_p':
            return self.d_y(factors=m['by'], variant='backward')
        a = self.d_x2(factors=m['mx']) + self.d_y2(factors=m['my'])
        if self._axi:
            a = a + self.d_x(factors=m['mc'] / r, variant='central')
        return a.todia()

    a_p_vx = property(lambda self: self._matrix('a_p_vx'))
    a_p_vy = property(lambda self: self._matrix('a_p_vy'))
    a_vx_p = property(lambda self: self._matrix('a_vx_p'))
    a_vy_p = property(lambda self: self._matrix('a_vy_p'))
    a_vx_vx = property(lambda self: self._matrix('a_vv'))
    a_vy_vy = property(lambda self: self._matrix('a_vv'))

    def is_stable(self):
        """CFL check with 1 % headroom (``pyfds/acoustics.py:130-139``)."""
        return np.all(self.material_vector('sound_velocity')
                      < 0.99 * min(self.x.increment, self.y.increment) / self.t.increment)


class Acoustic3DAxi(Acoustic2D):
    """Axial-symmetric acoustics, x is the radius and y the axis
    (``pyfds/acoustics.py:142-236``).  The device reproduces the per-step ``vx/r**2`` loss
    term with ``dt*mu/rho`` taken from the materials at assembly time."""

    _model_id = nat.ACOUSTIC_AXI
    _axi = True


AcousticAS = Acoustic3DAxi


class AcousticMaterial:
    """Acoustic material parameters (``pyfds/acoustics.py:239-286``)."""

    def __init__(self, sound_velocity, density, shear_viscosity=0, bulk_viscosity=0,
                 thermal_conductivity=0, isobaric_heat_cap=1, isochoric_heat_cap=1,
                 absorption_coef=None):
        self.sound_velocity = sound_velocity
        self.density = density
        self.shear_viscosity = shear_viscosity
        self.bulk_viscosity = bulk_viscosity
        self.thermal_conductivity = thermal_conductivity
        self.isobaric_heat_cap = isobaric_heat_cap
        self.isochoric_heat_cap = isochoric_heat_cap
        self._absorption_coef = absorption_coef

    @property
    def absorption_coef(self):
        """Sum of losses mu = 4/3 eta_s + eta_b + kappa (c_p - c_v) / (c_p c_v), unless set
        explicitly (a falsy explicit value counts as unset)."""
        if not self._absorption_coef:
            return (4 / 3 * self.shear_viscosity + self.bulk_viscosity
                    + self.thermal_conductivity
                    * (self.isobaric_heat_cap - self.isochoric_heat_cap)
                    / (self.isobaric_heat_cap * self.isochoric_heat_cap))
        return self._absorption_coef

    @absorption_coef.setter
    def absorption_coef(self, value):
        self._absorption_coef = value
