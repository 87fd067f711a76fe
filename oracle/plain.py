"""Plain FP64 restatement of pyfds' explicit staggered-grid solvers (numpy).

TEST INFRASTRUCTURE ONLY -- the oracle.  Imported by tests/, by
``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of bench.py, and
there only as the checker or the timed CPU baseline.  The product package
``pyfds_amd`` never imports it.

What is restated (reference = emtpb/pyfds, files cited as ``pyfds/<file>:<line>``):

* the scipy DIA operator algebra the reference relies on (scipy 1.15.3,
  ``scipy/sparse/_dia.py:181-234``; its compiled ``dia_matvec`` loop is the
  published BSD algorithm ``for each stored diagonal: y[i] += data[j] * x[j]``
  with ``y`` starting at zero) -- :class:`Dia`;
* the difference operators ``d_x/d_y/d_x2/d_y2`` (``pyfds/fields.py:158-202,
  273-365``) -- :func:`d_first`, :func:`d_second`;
* ``material_vector`` (``pyfds/fields.py:34-57``), boundary application
  (``pyfds/fields.py:591-600`` + ``pyfds/regions.py:125-145``), output recording
  (``pyfds/fields.py:602-611``), ``simulate`` (``pyfds/fields.py:67-95``) and
  ``reset`` (``pyfds/fields.py:121-127``);
* the per-model coefficient expressions and step orders of Acoustic1D/2D/3DAxi
  (``pyfds/acoustics.py:27-52, 89-128, 166-225``), Thermal1D/2D/3DAxi
  (``pyfds/thermal.py:29-51, 75-107, 132-176``), IdealGas1D and
  Acoustic2ndOrder1D (``pyfds/nonlinear_acoustics.py:38-89, 133-171``).

Coefficient expressions keep the reference's evaluation order so FP64 results
are bit-identical to the reference wherever the reference itself is
deterministic (all arithmetic here is IEEE add/sub/mul/div except ``**gamma``
in IdealGas1D).

The reference's ``**gamma`` is numpy's power, whose bits depend on the host: on
a CPU with AVX-512 numpy 2.2 computes it with its bundled SVML routine (which
rounds about 5 % of these calls differently from the nearest double), elsewhere
with glibc's ``pow`` (about 0.07 %).  :func:`pow_cr` is the correctly rounded
power, written independently of both (long double power, decided by Python's
``decimal`` wherever the long double result lies near a rounding midpoint);
``PlainIdealGas1D.power`` selects which one the oracle's equation of state uses
(numpy's by default: the reference on this host).

How each function is pinned (checked by tests/test_oracle.py):

* :class:`Dia`, :func:`d_first`, :func:`d_second`, the summed viscous
  operators and every model's ``a_*`` matrices: the reference's own known-answer
  matrices (``test/test_fields.py:57-107``, ``test/test_acoustics.py:19-66``,
  ``test/test_coupling.py:47-84``), scipy's ``dia_matrix.dot`` on random
  diagonals, and dense brute force.
* time-evolved runs of every model, boundary/output semantics, segmented
  ``simulate`` and ``reset``: golden fixtures written by running the reference
  itself (``tests/golden/make_golden.py``), plus the closed forms in
  tests/test_oracle.py (telescoping sums, the 1D magic time step, linearity).
"""
import numpy as np

__all__ = [
    'Dia', 'd_first', 'd_second', 'material_vector', 'apply_boundaries',
    'record_outputs', 'PlainField', 'PlainAcoustic1D', 'PlainAcoustic2D',
    'PlainAcoustic3DAxi', 'PlainThermal1D', 'PlainThermal2D', 'PlainThermal3DAxi',
    'PlainIdealGas1D', 'PlainAcoustic2ndOrder1D', 'absorption_coef', 'MODELS', 'pow_cr',
    'numpy_power_is_libm',
]


# --------------------------------------------------------------------------------------
# correctly rounded power
# --------------------------------------------------------------------------------------
def pow_cr(x, y):
    """x**y rounded to the nearest double (ties to even), elementwise, for finite x > 0.

    The long double (x87, 64-bit significand) power is within a few long double ulps of the
    exact value, i.e. within 2^-9 of a double ulp.  Where it lies farther than 2^-8 double
    ulps from the midpoint between the two doubles around it, rounding it gives the
    correctly rounded double; the remaining ~0.8 % of inputs are decided by ``decimal`` at
    80 significant digits (exp(y ln x)), which is far from any midpoint for these inputs.
    Pinned against ``decimal`` on random inputs and against glibc's pow where glibc is
    correctly rounded (tests/test_crmath.py).
    """
    from decimal import Decimal, localcontext
    x = np.asarray(x, np.float64)
    y = np.broadcast_to(np.asarray(y, np.float64), x.shape)
    if np.any(~np.isfinite(x) | (x <= 0) | ~np.isfinite(y)):
        raise ValueError('pow_cr: finite x > 0 and finite y only')
    ld = np.power(x.astype(np.longdouble), y.astype(np.longdouble))
    d = ld.astype(np.float64)
    other = np.nextafter(d, np.where(ld > d, np.inf, -np.inf))
    dl, ol = d.astype(np.longdouble), other.astype(np.longdouble)
    mid = (dl + ol) / 2
    ulp = np.abs(ol - dl)
    near = ~(np.abs(ld - mid) > ulp * np.longdouble(2.0) ** -8) | ~np.isfinite(d)
    out = d.copy()
    with localcontext() as ctx:
        ctx.prec = 80
        for i in np.flatnonzero(near):
            out.flat[i] = float((Decimal(float(y.flat[i])) * Decimal(float(x.flat[i])).ln()).exp())
    return out


def numpy_power_is_libm():
    """True when this host's numpy power rounds like glibc's pow (no SVML dispatch)."""
    import math
    probe = 1.0 + np.linspace(-1e-3, 1e-3, 4001)
    return bool(np.array_equal(probe ** 1.4, np.array([math.pow(v, 1.4) for v in probe])))


# --------------------------------------------------------------------------------------
# DIA operator algebra
# --------------------------------------------------------------------------------------
class Dia:
    """Square ``n x n`` operator stored like scipy's DIA format.

    ``diags`` is a list of ``(offset, data)`` in *stored* order; ``data`` has
    length ``n`` and is indexed by *column*: ``A[i, i + off] = data[i + off]``
    (scipy ``dia_matrix((data, offsets))`` semantics).
    """

    def __init__(self, n, diags):
        self.n = int(n)
        self.diags = [(int(o), np.asarray(d, dtype=float)) for o, d in diags]

    @property
    def offsets(self):
        return [o for o, _ in self.diags]

    def matvec(self, x):
        """``y = A x`` in scipy ``dia_matvec`` order (``_dia.py:221-234``):
        ``y = 0``; for each stored diagonal ``k``: ``y[i] += data[k, j] * x[j]``
        with ``j = i + off`` over the in-range rows.  The product is rounded
        before the add (no fused multiply-add)."""
        n = self.n
        x = np.asarray(x, dtype=float)
        y = np.zeros(n)
        for off, data in self.diags:
            if off >= n or -off >= n:
                continue
            if off >= 0:
                y[:n - off] += data[off:] * x[off:]
            else:
                y[-off:] += data[:n + off] * x[:n + off]
        return y

    # When True, ``dot`` multiplies with scipy's own ``dia_matrix`` built from the same stored
    # diagonals (scipy's compiled ``dia_matvec``, the reference's CPU kernel; same bits as
    # ``matvec``): the timed CPU baseline uses it.
    use_scipy = False

    def dot(self, x):
        if Dia.use_scipy:
            m = self.__dict__.get('_scipy')
            if m is None:
                import scipy.sparse as sp
                m = sp.dia_matrix((np.array([d for _, d in self.diags]), self.offsets),
                                  shape=(self.n, self.n))
                self._scipy = m
            return m.dot(np.asarray(x, dtype=float))
        return self.matvec(x)

    def __add__(self, other):
        """Sum of two DIA operators as scipy 1.15.3 forms it
        (``_dia.py:181-213``): offsets become the sorted union, a diagonal
        present in both operands holds ``self + other``, the others are
        copied."""
        mine = dict(self.diags)
        theirs = dict(other.diags)
        if self.offsets == other.offsets:
            return Dia(self.n, [(o, mine[o] + theirs[o]) for o in self.offsets])
        out = []
        for off in sorted(set(mine) | set(theirs)):
            if off in mine and off in theirs:
                out.append((off, mine[off] + theirs[off]))
            elif off in mine:
                out.append((off, mine[off].copy()))
            else:
                out.append((off, theirs[off].copy()))
        return Dia(self.n, out)

    def toarray(self):
        a = np.zeros((self.n, self.n))
        for off, data in self.diags:
            for i in range(self.n):
                j = i + off
                if 0 <= j < self.n:
                    a[i, j] += data[j]
        return a


def d_first(n, stride, factors=None, variant='forward'):
    """First difference with column-indexed factors; ``stride`` is 1 for x and
    ``x_samples`` for y (``pyfds/fields.py:158-184, 273-328``)."""
    if factors is None:
        factors = np.array(1).repeat(n)
    f = np.asarray(factors)
    if variant == 'forward':
        return Dia(n, [(0, -f), (stride, f)])
    if variant == 'central':
        return Dia(n, [(-stride, -f / 2), (stride, f / 2)])
    if variant == 'backward':
        return Dia(n, [(-stride, -f), (0, f)])
    raise ValueError('Unknown difference quotient variant {}.'.format(variant))


def d_second(n, stride, factors=None):
    """Three-point second difference (``pyfds/fields.py:186-202, 330-365``)."""
    if factors is None:
        factors = np.array(1).repeat(n)
    f = np.asarray(factors)
    return Dia(n, [(-stride, f), (0, -2 * f), (stride, f)])


# --------------------------------------------------------------------------------------
# regions, boundaries, outputs
# --------------------------------------------------------------------------------------
def absorption_coef(mat):
    """``AcousticMaterial.absorption_coef`` (``pyfds/acoustics.py:271-282``): the
    explicit value unless it is falsy, else 4/3 eta_s + eta_b + kappa (cp-cv)/(cp cv)."""
    explicit = mat.get('absorption_coef')
    if not explicit:
        return (4 / 3 * mat.get('shear_viscosity', 0) + mat.get('bulk_viscosity', 0)
                + mat.get('thermal_conductivity', 0)
                * (mat.get('isobaric_heat_cap', 1) - mat.get('isochoric_heat_cap', 1))
                / (mat.get('isobaric_heat_cap', 1) * mat.get('isochoric_heat_cap', 1)))
    return explicit


def material_vector(n, material_regions, parameter):
    """``Field.material_vector`` (``pyfds/fields.py:34-57``): zeros, then each
    region in list order assigns its material's value (later regions win);
    ``KeyError`` when no material has the parameter.

    ``material_regions`` is a list of ``(indices, params_dict)``."""
    vec = np.zeros(n)
    found = False
    for indices, params in material_regions:
        if parameter in params:
            vec[list(indices)] = params[parameter]
            found = True
    if not found:
        raise KeyError('Material parameter {} not found in set materials.'.format(parameter))
    return vec


def apply_boundaries(values, boundaries, step):
    """``FieldComponent.apply_bounds`` + ``Boundary.apply``
    (``pyfds/fields.py:591-600``, ``pyfds/regions.py:125-145``): boundaries in
    list order; each gathers its old values, computes ``additive*old + value``
    (value: scalar, ``signal[step]``, or one signal per point) and scatters
    (for repeated indices the last write wins)."""
    for b in boundaries:
        idx = list(b['indices'])
        old = values[idx]
        value, additive = b['value'], b['additive']
        if np.ndim(value) == 0:
            new = additive * old + value
        elif isinstance(value, np.ndarray):
            new = additive * old + value[step]
        else:
            new = [additive * old[ii] + signal[step] for ii, signal in enumerate(value)]
        values[idx] = new


def record_outputs(values, outputs):
    """``FieldComponent.write_outputs`` (``pyfds/fields.py:602-611``): each
    output keeps one list per index and appends the current value."""
    for out in outputs:
        if not out['signals']:
            out['signals'] = [[values[i]] for i in out['indices']]
        else:
            for i, sig in zip(out['indices'], out['signals']):
                sig.append(values[i])


# --------------------------------------------------------------------------------------
# fields
# --------------------------------------------------------------------------------------
class PlainField:
    """State container + time loop shared by all models.

    ``components``: name -> float64 array; ``boundaries``: name -> list of
    ``{'indices', 'value', 'additive'}``; ``outputs``: name -> list of
    ``{'indices', 'signals'}``; ``materials``: list of ``(indices, params)``.
    """
    component_names = ()
    dims = 1

    def __init__(self, nx, dx, nt, dt, materials, ny=1, dy=1.0):
        self.nx, self.ny = int(nx), int(ny)
        self.dx, self.dy, self.dt, self.nt = dx, dy, dt, int(nt)
        self.n = self.nx * self.ny
        self.materials = list(materials)
        self.components = {c: np.zeros(self.n) for c in self.component_names}
        self.boundaries = {c: [] for c in self.component_names}
        self.outputs = {c: [] for c in self.component_names}
        self.step = 0
        self.assembled = False

    # helpers
    def mv(self, parameter):
        return material_vector(self.n, self.materials, parameter)

    def x_vector(self):
        """``Dimension.vector`` (``pyfds/fields.py:551-555``)."""
        return np.arange(start=0, stop=self.nx) * self.dx

    def radii(self):
        """``Acoustic3DAxi._radii`` / ``Thermal3DAxi._radii``
        (``pyfds/acoustics.py:166-176``, ``pyfds/thermal.py:132-142``)."""
        return np.tile(self.x_vector(), self.ny) + self.dx / 2

    def bounds(self, comp):
        apply_boundaries(self.components[comp], self.boundaries[comp], self.step)

    def record(self, comp):
        record_outputs(self.components[comp], self.outputs[comp])

    def bounds_and_record(self, comp):
        self.bounds(comp)
        self.record(comp)

    def assemble(self):
        raise NotImplementedError

    def sim_step(self):
        raise NotImplementedError

    def simulate(self, num_steps=None):
        """``Field.simulate`` (``pyfds/fields.py:67-95``): falsy ``num_steps``
        runs ``t.samples`` steps; assembles once."""
        if not num_steps:
            num_steps = self.nt
        if not self.assembled:
            self.assemble()
        start = self.step
        while self.step < start + num_steps:
            self.sim_step()
            self.step += 1

    def reset(self):
        """``Field.reset`` (``pyfds/fields.py:121-127``): zero every component,
        step 0, outputs kept."""
        for c in self.components:
            self.components[c] = np.zeros_like(self.components[c])
        self.step = 0

    def signals(self, comp):
        return [o['signals'] for o in self.outputs[comp]]


class PlainAcoustic1D(PlainField):
    """``Acoustic1D`` (``pyfds/acoustics.py:9-52``)."""
    component_names = ('pressure', 'velocity')

    def assemble(self):
        dt, dx, n = self.dt, self.dx, self.n
        self.K = dt / dx * self.mv('sound_velocity') ** 2 * self.mv('density')
        self.B = dt / dx / self.mv('density')
        self.M = dt / dx ** 2 * self.mv('absorption_coef') / self.mv('density')
        self.a_p_v = d_first(n, 1, self.K)
        self.a_v_p = d_first(n, 1, self.B, 'backward')
        self.a_v_v = d_second(n, 1, self.M)
        self.assembled = True

    def sim_step(self):
        c = self.components
        self.bounds_and_record('pressure')
        c['velocity'] -= (self.a_v_p.dot(c['pressure']) - self.a_v_v.dot(c['velocity']))
        self.bounds_and_record('velocity')
        c['pressure'] -= self.a_p_v.dot(c['velocity'])


class PlainAcoustic2D(PlainField):
    """``Acoustic2D`` (``pyfds/acoustics.py:66-128``)."""
    component_names = ('pressure', 'velocity_x', 'velocity_y')
    dims = 2

    def coefficients(self):
        dt, dx, dy = self.dt, self.dx, self.dy
        c, rho, mu = self.mv('sound_velocity'), self.mv('density'), self.mv('absorption_coef')
        self.Kx = dt / dx * c ** 2 * rho
        self.Ky = dt / dy * c ** 2 * rho
        self.Bx = dt / dx / rho
        self.By = dt / dy / rho
        self.Mx = dt / dx ** 2 * mu / rho
        self.My = dt / dy ** 2 * mu / rho

    def assemble(self):
        self.coefficients()
        n, nx = self.n, self.nx
        self.a_p_vx = d_first(n, 1, self.Kx)
        self.a_p_vy = d_first(n, nx, self.Ky)
        self.a_vx_p = d_first(n, 1, self.Bx, 'backward')
        self.a_vy_p = d_first(n, nx, self.By, 'backward')
        self.a_vx_vx = d_second(n, 1, self.Mx) + d_second(n, nx, self.My)
        self.a_vy_vy = self.a_vx_vx
        self.assembled = True

    def sim_step(self):
        c = self.components
        self.bounds_and_record('pressure')
        c['velocity_x'] -= (self.a_vx_p.dot(c['pressure']) - self.a_vx_vx.dot(c['velocity_x']))
        c['velocity_y'] -= (self.a_vy_p.dot(c['pressure']) - self.a_vy_vy.dot(c['velocity_y']))
        self.bounds_and_record('velocity_x')
        self.bounds_and_record('velocity_y')
        c['pressure'] -= (self.a_p_vx.dot(c['velocity_x']) + self.a_p_vy.dot(c['velocity_y']))


class PlainAcoustic3DAxi(PlainField):
    """``Acoustic3DAxi`` (``pyfds/acoustics.py:142-225``); x is the radius."""
    component_names = ('pressure', 'velocity_x', 'velocity_y')
    dims = 2

    def assemble(self):
        dt, dx, dy, n, nx = self.dt, self.dx, self.dy, self.n, self.nx
        c, rho, mu = self.mv('sound_velocity'), self.mv('density'), self.mv('absorption_coef')
        r = self.radii()
        self.a_p_vx = d_first(n, 1, dt / dx * c ** 2 * rho / r)
        self.a_p_vy = d_first(n, nx, dt / dy * c ** 2 * rho)
        self.a_vx_p = d_first(n, 1, dt / dx / rho, 'backward')
        self.a_vy_p = d_first(n, nx, dt / dy / rho, 'backward')
        self.a_vx_vx = (d_second(n, 1, dt / dx ** 2 * mu / rho)
                        + d_second(n, nx, dt / dy ** 2 * mu / rho)
                        + d_first(n, 1, dt / dx * mu / rho / r, 'central'))
        self.a_vy_vy = self.a_vx_vx
        self.assembled = True

    def sim_step(self):
        c = self.components
        self.bounds_and_record('pressure')
        # material vectors and radii are re-evaluated every step (pyfds/acoustics.py:213-215)
        c['velocity_x'] -= (self.a_vx_p.dot(c['pressure'])
                            - self.a_vx_vx.dot(c['velocity_x'])
                            + self.dt * self.mv('absorption_coef') / self.mv('density')
                            * c['velocity_x'] / self.radii() ** 2)
        c['velocity_y'] -= (self.a_vy_p.dot(c['pressure']) - self.a_vy_vy.dot(c['velocity_y']))
        self.bounds_and_record('velocity_x')
        self.bounds_and_record('velocity_y')
        c['pressure'] -= (self.a_p_vx.dot(c['velocity_x'] * self.radii())
                          + self.a_p_vy.dot(c['velocity_y']))


class PlainThermal1D(PlainField):
    """``Thermal1D`` (``pyfds/thermal.py:12-51``)."""
    component_names = ('temperature', 'heat_flux')

    def assemble(self):
        dt, dx, n = self.dt, self.dx, self.n
        self.a_t_q = d_first(n, 1, dt / dx / self.mv('density') / self.mv('heat_capacity'))
        self.a_q_t = d_first(n, 1, 1 / dx * self.mv('thermal_conductivity_x'), 'backward')
        self.assembled = True

    def sim_step(self):
        c = self.components
        self.bounds_and_record('temperature')
        c['heat_flux'] = -self.a_q_t.dot(c['temperature'])
        self.bounds_and_record('heat_flux')
        c['temperature'] -= self.a_t_q.dot(c['heat_flux'])


class PlainThermal2D(PlainField):
    """``Thermal2D`` (``pyfds/thermal.py:54-107``)."""
    component_names = ('temperature', 'heat_flux_x', 'heat_flux_y')
    dims = 2
    axi = False

    def assemble(self):
        dt, dx, dy, n, nx = self.dt, self.dx, self.dy, self.n, self.nx
        rho, cp = self.mv('density'), self.mv('heat_capacity')
        gx = dt / dx / rho / cp
        if self.axi:
            gx = gx / self.radii()
        self.a_t_qx = d_first(n, 1, gx)
        self.a_t_qy = d_first(n, nx, dt / dy / rho / cp)
        self.a_qx_t = d_first(n, 1, 1 / dx * self.mv('thermal_conductivity_x'), 'backward')
        self.a_qy_t = d_first(n, nx, 1 / dy * self.mv('thermal_conductivity_y'), 'backward')
        self.assembled = True

    def sim_step(self):
        c = self.components
        self.bounds_and_record('temperature')
        c['heat_flux_x'] = -self.a_qx_t.dot(c['temperature'])
        c['heat_flux_y'] = -self.a_qy_t.dot(c['temperature'])
        self.bounds_and_record('heat_flux_x')
        self.bounds_and_record('heat_flux_y')
        qx = c['heat_flux_x'] * self.radii() if self.axi else c['heat_flux_x']
        c['temperature'] -= (self.a_t_qx.dot(qx) + self.a_t_qy.dot(c['heat_flux_y']))


class PlainThermal3DAxi(PlainThermal2D):
    """``Thermal3DAxi`` (``pyfds/thermal.py:110-176``): ``a_t_qx`` carries
    ``/r`` and is applied to ``q_x * r``."""
    axi = True


class PlainIdealGas1D(PlainField):
    """``IdealGas1D`` (``pyfds/nonlinear_acoustics.py:10-89``).

    ``power`` is the ``**`` of the nonlinear equation of state: ``'numpy'`` (the reference's
    numpy power on this host) or ``'cr'`` (:func:`pow_cr`)."""
    component_names = ('pressure', 'velocity', 'density')
    power = 'numpy'

    def __init__(self, *args, convective=True, nl_state=True, **kwargs):
        super().__init__(*args, **kwargs)
        self.convective, self.nl_state = convective, nl_state

    def assemble_common(self):
        dt, dx, n = self.dt, self.dx, self.n
        self.a_d_v = d_first(n, 1, dt / dx * np.ones(self.nx))
        self.a_v_p = d_first(n, 1, (dt / dx) * np.ones(self.nx), 'backward')
        self.a_v_v = d_second(n, 1, dt / dx ** 2 * self.mv('absorption_coef'))
        self.a_v_v2 = d_first(n, 1, (dt / dx / 2) * np.ones(self.nx), 'central')
        self.stat_density = self.mv('density')

    def assemble(self):
        self.assemble_common()
        self.sound_velocity = self.mv('sound_velocity')
        self.heat_cap_ratio = self.mv('isobaric_heat_cap') / self.mv('isochoric_heat_cap')
        self.assembled = True

    def velocity_update(self, convective):
        c, rho0 = self.components, self.stat_density
        if convective:
            c['velocity'] -= (self.a_v_p.dot(c['pressure']) / (rho0 + c['density'])
                              + self.a_v_v2.dot(c['velocity'])
                              - self.a_v_v.dot(c['velocity']) / (rho0 + c['density']))
        else:
            c['velocity'] -= (self.a_v_p.dot(c['pressure']) / (rho0 + c['density'])
                              - self.a_v_v.dot(c['velocity']) / (rho0 + c['density']))

    def density_update(self):
        c = self.components
        c['density'] -= self.a_d_v.dot((c['density'] + self.stat_density) * c['velocity'])

    def equation_of_state(self):
        c, rho0 = self.components, self.stat_density
        if self.nl_state:
            base = (rho0 + c['density']) / rho0
            pw = (pow_cr(base, self.heat_cap_ratio) if self.power == 'cr'
                  else base ** self.heat_cap_ratio)
            c['pressure'] = (rho0 * self.sound_velocity ** 2 / self.heat_cap_ratio
                             * (pw - 1))
        else:
            c['pressure'] = self.sound_velocity ** 2 * c['density']

    def sim_step(self):
        self.bounds_and_record('pressure')
        self.velocity_update(self.convective)
        self.bounds_and_record('velocity')
        self.density_update()
        self.bounds_and_record('density')
        self.equation_of_state()


class PlainAcoustic2ndOrder1D(PlainIdealGas1D):
    """``Acoustic2ndOrder1D`` (``pyfds/nonlinear_acoustics.py:103-171``)."""

    def assemble(self):
        self.assemble_common()
        self.d_rho_p = self.mv('d_rho_p')
        self.d_rho2_p = self.mv('d_rho2_p')
        self.assembled = True

    def equation_of_state(self):
        c = self.components
        c['pressure'] = self.d_rho_p * c['density'] + self.d_rho2_p / 2 * c['density'] ** 2

    def sim_step(self):
        self.bounds_and_record('pressure')
        self.velocity_update(True)
        self.bounds_and_record('velocity')
        self.density_update()
        self.bounds_and_record('density')
        self.equation_of_state()


MODELS = {
    'Acoustic1D': PlainAcoustic1D,
    'Acoustic2D': PlainAcoustic2D,
    'Acoustic3DAxi': PlainAcoustic3DAxi,
    'Thermal1D': PlainThermal1D,
    'Thermal2D': PlainThermal2D,
    'Thermal3DAxi': PlainThermal3DAxi,
    'IdealGas1D': PlainIdealGas1D,
    'Acoustic2ndOrder1D': PlainAcoustic2ndOrder1D,
}
