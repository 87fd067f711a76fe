"""What the FP32-Kahan form buys, measured against the FP64 form (bit-identical to the
reference) on the GPU:

* Thermal2D, 20 000 steps of diffusion: each step changes T by a small increment, so plain FP32
  loses the increments' low bits every step and the error grows with the step count; the
  compensated sum keeps them (FP32 vs Kahan: orders of magnitude).
* Acoustic2D, wave propagation: the state changes by O(1) relative each period, and the error
  floor is the rounding of the coefficients K, B to float32 (a wave-speed error that grows with
  the distance travelled).  An FP64 run whose coefficients are rounded to float32 shows that
  floor; FP32-Kahan sits at it, plain FP32 a little above.
"""
import numpy as np
import pytest

import pyfds_amd
import scenes

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def thermal_scene(n=256, steps=20000, seed=3):
    dx = 1e-3
    rho, cp, kmax = 1000.0, 4000.0, 50.0
    dt = 0.2 * dx ** 2 * rho * cp / kmax
    rng = np.random.default_rng(seed)
    nb = n // 16
    hot = np.kron(rng.random((nb, nb)) < 0.5, np.ones((16, 16), bool)).ravel()
    iy, ix = np.divmod(np.arange(n * n), n)
    return dict(
        name='thermal_long', model='Thermal2D',
        kwargs=dict(x_samples=n, x_delta=dx, y_samples=n, y_delta=dx, t_samples=steps,
                    t_delta=dt),
        material=('ThermalMaterial', dict(heat_capacity=cp, density=rho,
                                          thermal_conductivity=0.6)),
        regions=[(np.flatnonzero(hot), ('ThermalMaterial',
                                        dict(heat_capacity=cp, density=rho,
                                             thermal_conductivity=kmax)))],
        init={'temperature': np.exp(-0.5 * ((ix - n / 2) ** 2 + (iy - n / 2) ** 2) / 16.0 ** 2)},
        boundaries={}, outputs={'temperature': [[n // 2 + n * (n // 2), 5 + 7 * n]]},
        plan=[('simulate', None)])


def _run(scene, form, coef32=False):
    fld = scenes.build_field(pyfds_amd, pyfds_amd.regions, scene)
    fld.form = form
    if coef32:
        fld.assemble_matrices()
        for k, v in fld._coefs.items():
            if v is not None:
                fld._coefs[k] = np.asarray(v).astype(np.float32).astype(np.float64)
        fld._coef_version += 1
    scenes.run_plan(fld, scene)
    return scenes.collect(fld, scene)


def test_kahan_keeps_the_small_increments_of_a_long_thermal_run():
    scene = thermal_scene()
    ref = _run(scene, 'fp64')
    e32 = rel(_run(scene, 'fp32')['values/temperature'], ref['values/temperature'])
    ek = rel(_run(scene, 'fp32_kahan')['values/temperature'], ref['values/temperature'])
    print('Thermal2D 256^2 x 20000 steps: rel L2 vs FP64  fp32 %.3e  fp32_kahan %.3e'
          % (e32, ek))
    assert ek < e32 / 20


def test_acoustic_fp32_error_floor_is_the_coefficient_rounding():
    scene = dict(scenes.SCENES['a2d_main'](), plan=[('simulate', None)])
    ref = _run(scene, 'fp64')
    floor = _run(scene, 'fp64', coef32=True)
    e_floor = rel(floor['values/pressure'], ref['values/pressure'])
    e32 = rel(_run(scene, 'fp32')['values/pressure'], ref['values/pressure'])
    ek = rel(_run(scene, 'fp32_kahan')['values/pressure'], ref['values/pressure'])
    print('Acoustic2D 48x40 x 150 steps: rel L2 vs FP64  coefficients rounded to float32 '
          '%.3e  fp32 %.3e  fp32_kahan %.3e' % (e_floor, e32, ek))
    assert e_floor <= ek <= 2.5 * e_floor
    assert ek <= e32
