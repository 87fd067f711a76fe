"""Seeded simulation scenes shared by the golden-fixture generator, the oracle
tests and the device parity tests.

A scene is a plain dict: model name + constructor kwargs, main material,
extra material regions given as index lists, initial component values,
boundaries, outputs and a run plan (``simulate(n)`` / ``sim_step`` / ``reset``
actions).  :func:`build_field` constructs it through any namespace that offers
pyfds' public API (the reference ``pyfds`` when writing fixtures, ``pyfds_amd``
for the device path); :func:`build_oracle` constructs the same scene on the
CPU oracle.
"""
import hashlib
import json

import numpy as np

COMPONENTS = {
    'Acoustic1D': ('pressure', 'velocity'),
    'Acoustic2D': ('pressure', 'velocity_x', 'velocity_y'),
    'Acoustic3DAxi': ('pressure', 'velocity_x', 'velocity_y'),
    'Thermal1D': ('temperature', 'heat_flux'),
    'Thermal2D': ('temperature', 'heat_flux_x', 'heat_flux_y'),
    'Thermal3DAxi': ('temperature', 'heat_flux_x', 'heat_flux_y'),
    'IdealGas1D': ('pressure', 'velocity', 'density'),
    'Acoustic2ndOrder1D': ('pressure', 'velocity', 'density'),
}

WATER = dict(sound_velocity=1500.0, density=1000.0)
STEEL = dict(sound_velocity=5900.0, density=7850.0)


def _grid(s):
    k = s['kwargs']
    return k['x_samples'], k.get('y_samples', 1)


def rect_indices(nx, x0, y0, wx, wy):
    """Flat indices of an axis-aligned block, x fastest."""
    return [x + nx * y for y in range(y0, y0 + wy) for x in range(x0, x0 + wx)]


# --------------------------------------------------------------------------------------
# scene catalogue
# --------------------------------------------------------------------------------------
def scene_a1d_rigid():
    n, dx = 64, 1e-4
    dt = 0.5 * dx / 1500.0
    i = np.arange(n)
    return dict(
        name='a1d_rigid', model='Acoustic1D',
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=200, t_delta=dt),
        material=('AcousticMaterial', dict(WATER)), regions=[],
        init={'pressure': np.exp(-0.5 * ((i - 32) / 4.0) ** 2)},
        boundaries={'velocity': [dict(indices=[0], value=0, additive=False),
                                 dict(indices=[n - 1], value=0, additive=False)]},
        outputs={'pressure': [[10, 32, 50]], 'velocity': [[20]]},
        plan=[('simulate', None)])


def scene_a1d_lossy_src():
    n, dx = 50, 1e-4
    dt = 0.4 * dx / 1500.0
    t = np.arange(130) * dt
    sig = np.sin(2 * np.pi * 2e6 * t) * (t < 2e-6)
    return dict(
        name='a1d_lossy_src', model='Acoustic1D',
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=130, t_delta=dt),
        material=('AcousticMaterial', dict(WATER, shear_viscosity=1e-3, bulk_viscosity=2e-3)),
        regions=[(list(range(30, 40)), ('AcousticMaterial',
                                        dict(sound_velocity=2500.0, density=1800.0,
                                             absorption_coef=5e-3)))],
        init={},
        boundaries={'pressure': [dict(indices=[10], value=sig, additive=True),
                                 dict(indices=[n - 1], value=0, additive=False)]},
        outputs={'pressure': [[5, 25, 45]], 'velocity': [[11, 12]]},
        plan=[('simulate', 50), ('simulate', 70)])


def scene_a2d_main():
    nx, ny, dx = 48, 40, 1e-4
    dt = 0.5 * dx / 5900.0
    t = np.arange(150) * dt
    sig = np.sin(2 * np.pi * 1e6 * t) * (t < 5e-6)
    rng = np.random.default_rng(11)
    outs = [int(v) for v in rng.choice(nx * ny, 4, replace=False)]
    return dict(
        name='a2d_main', model='Acoustic2D',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=150, t_delta=dt),
        material=('AcousticMaterial', dict(WATER)),
        regions=[(rect_indices(nx, 28, 6, 12, 9), ('AcousticMaterial', dict(STEEL)))],
        init={},
        boundaries={'pressure': [dict(indices=[20 + nx * 20], value=sig, additive=True)],
                    'velocity_x': [dict(indices=[nx * y for y in range(ny)], value=0,
                                        additive=False)]},
        outputs={'pressure': [outs], 'velocity_x': [[nx * 20 + 5, nx * 21 + 47]],
                 'velocity_y': [[nx * 3 + 4], [nx * 39 + 10]]},
        plan=[('simulate', None)])


def scene_a2d_lossy():
    nx, ny, dx, dy = 24, 20, 1e-4, 1.5e-4
    dt = 0.3 * dx / 2500.0
    rng = np.random.default_rng(12)
    n = nx * ny
    return dict(
        name='a2d_lossy', model='Acoustic2D',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dy,
                    t_samples=80, t_delta=dt),
        material=('AcousticMaterial', dict(WATER, shear_viscosity=2e-3)),
        regions=[(rect_indices(nx, 3, 4, 8, 6), ('AcousticMaterial',
                                                 dict(sound_velocity=2500.0, density=1200.0,
                                                      bulk_viscosity=4e-3)))],
        init={'pressure': rng.standard_normal(n), 'velocity_x': 1e-3 * rng.standard_normal(n),
              'velocity_y': 1e-3 * rng.standard_normal(n)},
        boundaries={},
        outputs={'pressure': [[5, 100, 300]], 'velocity_y': [[7, 8]]},
        plan=[('simulate', 30), ('simulate', 50)])


def scene_axi_lossless():
    nx, ny, dx = 32, 24, 1e-4
    dt = 0.5 * dx / 1500.0
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
    p0 = np.exp(-0.5 * (((ix - 4) / 3.0) ** 2 + ((iy - 12) / 3.0) ** 2)).ravel()
    return dict(
        name='axi_lossless', model='Acoustic3DAxi',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=100, t_delta=dt),
        material=('AcousticMaterial', dict(WATER)), regions=[],
        init={'pressure': p0}, boundaries={},
        outputs={'pressure': [[0, 12 * nx + 10, 20 * nx + 31]],
                 'velocity_x': [[12 * nx + 3]]},
        plan=[('simulate', None)])


def scene_axi_lossy():
    nx, ny, dx = 20, 16, 1e-4
    dt = 0.4 * dx / 2000.0
    t = np.arange(60) * dt
    sig = np.sin(2 * np.pi * 2e6 * t)
    return dict(
        name='axi_lossy', model='Acoustic3DAxi',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=60, t_delta=dt),
        material=('AcousticMaterial', dict(WATER, shear_viscosity=5e-3)),
        regions=[(rect_indices(nx, 8, 2, 6, 5), ('AcousticMaterial',
                                                 dict(sound_velocity=2000.0, density=1500.0,
                                                      absorption_coef=1e-2)))],
        init={},
        boundaries={'pressure': [dict(indices=[nx * y + 2 for y in range(6, 9)], value=sig,
                                      additive=True)],
                    'velocity_x': [dict(indices=[nx * y + nx - 1 for y in range(ny)],
                                        value=0, additive=False)]},
        outputs={'pressure': [[nx * 12 + 4, nx * 3 + 15]], 'velocity_x': [[nx * 7 + 3]],
                 'velocity_y': [[nx * 7 + 3]]},
        plan=[('simulate', None)])


def _thermal_blocks(nx, ny, block, seed):
    rng = np.random.default_rng(seed)
    idx = []
    for by in range(0, ny, block):
        for bx in range(0, nx, block):
            if rng.random() < 0.5:
                idx += rect_indices(nx, bx, by, min(block, nx - bx), min(block, ny - by))
    return idx


def scene_t2d_blocks():
    nx, ny, dx = 64, 48, 1e-3
    rho, cp, kmax = 1000.0, 4000.0, 50.0
    dt = 0.2 * dx ** 2 * rho * cp / kmax
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
    t0 = np.exp(-0.5 * (((ix - 32) ** 2 + (iy - 24) ** 2) / 8.0 ** 2)).ravel()
    return dict(
        name='t2d_blocks', model='Thermal2D',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=300, t_delta=dt),
        material=('ThermalMaterial', dict(heat_capacity=cp, density=rho,
                                          thermal_conductivity=0.6)),
        regions=[(_thermal_blocks(nx, ny, 8, 3),
                  ('ThermalMaterial', dict(heat_capacity=cp, density=rho,
                                           thermal_conductivity=50.0)))],
        init={'temperature': t0},
        boundaries={'temperature': [dict(indices=list(range(nx)), value=0.0, additive=False)]},
        outputs={'temperature': [[24 * nx + 32, 30 * nx + 40]], 'heat_flux_x': [[24 * nx + 36]],
                 'heat_flux_y': [[28 * nx + 32]]},
        plan=[('simulate', None)])


def scene_t2d_aniso():
    nx, ny, dx, dy = 20, 16, 1e-3, 2e-3
    dt = 0.1 * dx ** 2 * 2e6 / 40.0
    rng = np.random.default_rng(13)
    t = np.arange(100) * dt
    return dict(
        name='t2d_aniso', model='Thermal2D',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dy,
                    t_samples=100, t_delta=dt),
        material=('ThermalMaterial', dict(heat_capacity=1000.0, density=2000.0,
                                          thermal_conductivity=(40.0, 5.0))),
        regions=[(rect_indices(nx, 5, 5, 6, 4), ('ThermalMaterial',
                                                 dict(heat_capacity=500.0, density=8000.0,
                                                      thermal_conductivity=15.0)))],
        init={'temperature': rng.random(nx * ny)},
        boundaries={'heat_flux_x': [dict(indices=[nx * y + 10 for y in range(3, 8)],
                                         value=1e3 * np.cos(2e2 * t), additive=False)],
                    'heat_flux_y': [dict(indices=[4 * nx + 2, 4 * nx + 3], value=50.0,
                                         additive=True)]},
        outputs={'temperature': [[nx * 5 + 5, nx * 10 + 15]],
                 'heat_flux_x': [[nx * 4 + 10]], 'heat_flux_y': [[4 * nx + 2]]},
        plan=[('simulate', 40), ('simulate', 60)])


def scene_t1d():
    n, dx = 40, 1e-3
    dt = 0.2 * dx ** 2 * 1000.0 * 4000.0 / 50.0
    i = np.arange(n)
    return dict(
        name='t1d', model='Thermal1D',
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=200, t_delta=dt),
        material=('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                          thermal_conductivity=0.6)),
        regions=[(list(range(15, 25)), ('ThermalMaterial',
                                        dict(heat_capacity=4000.0, density=1000.0,
                                             thermal_conductivity=50.0)))],
        init={'temperature': np.exp(-0.5 * ((i - 18) / 3.0) ** 2)},
        boundaries={'temperature': [dict(indices=[0], value=1.0, additive=False)]},
        outputs={'temperature': [[5, 20]], 'heat_flux': [[20]]},
        plan=[('simulate', None)])


def scene_taxi():
    nx, ny, dx = 24, 20, 1e-3
    dt = 0.2 * dx ** 2 * 1000.0 * 4000.0 / 50.0
    ix, iy = np.meshgrid(np.arange(nx), np.arange(ny))
    t0 = np.exp(-0.5 * ((ix ** 2 + (iy - 10) ** 2) / 4.0 ** 2)).ravel()
    return dict(
        name='taxi', model='Thermal3DAxi',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=150, t_delta=dt),
        material=('ThermalMaterial', dict(heat_capacity=4000.0, density=1000.0,
                                          thermal_conductivity=50.0)),
        regions=[(rect_indices(nx, 10, 0, 4, ny), ('ThermalMaterial',
                                                   dict(heat_capacity=4000.0, density=1000.0,
                                                        thermal_conductivity=(5.0, 20.0))))],
        init={'temperature': t0}, boundaries={},
        outputs={'temperature': [[10 * nx, 10 * nx + 12]], 'heat_flux_x': [[10 * nx + 3]]},
        plan=[('simulate', None)])


def _ig_signal(n):
    t = np.arange(n)
    return 200.0 * np.sin(2 * np.pi * t / 25.0) * (t < 50)


def scene_ig1d(convective, nl_state):
    n, dx = 60, 1e-3
    dt = 0.5 * dx / 323.0
    tag = ('c' if convective else 'n') + ('nl' if nl_state else 'lin')
    return dict(
        name='ig1d_' + tag, model='IdealGas1D',
        model_kwargs=dict(convective=convective, nl_state=nl_state),
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=150, t_delta=dt),
        material=('AcousticMaterial', dict(sound_velocity=323.0, density=1.66,
                                           shear_viscosity=2.2e-5, isobaric_heat_cap=520.0,
                                           isochoric_heat_cap=312.0)),
        regions=[], init={},
        boundaries={'pressure': [dict(indices=[5], value=_ig_signal(150), additive=True)],
                    'velocity': [dict(indices=[0], value=0, additive=False)]},
        outputs={'pressure': [[20, 40]], 'velocity': [[30]], 'density': [[25]]},
        plan=[('simulate', None)])


def scene_ig1d_long(amplitude=200.0):
    """IdealGas1D, nonlinear state: 2048 cells x 2000 steps, a 400-step sine burst."""
    n, dx, nt = 2048, 1e-3, 2000
    dt = 0.5 * dx / 323.0
    t = np.arange(nt)
    sig = amplitude * np.sin(2 * np.pi * t / 40.0) * (t < 400)
    return dict(
        name='ig1d_long', model='IdealGas1D',
        model_kwargs=dict(convective=True, nl_state=True),
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=nt, t_delta=dt),
        material=('AcousticMaterial', dict(sound_velocity=323.0, density=1.66,
                                           shear_viscosity=2.2e-5, isobaric_heat_cap=520.0,
                                           isochoric_heat_cap=312.0)),
        regions=[], init={},
        boundaries={'pressure': [dict(indices=[5], value=sig, additive=True)],
                    'velocity': [dict(indices=[0], value=0, additive=False)]},
        outputs={'pressure': [[500, 1500]], 'velocity': [[1000]], 'density': [[700]]},
        plan=[('simulate', None)])


def scene_a2o1d():
    n, dx = 60, 1e-3
    dt = 0.5 * dx / 1500.0
    return dict(
        name='a2o1d', model='Acoustic2ndOrder1D',
        kwargs=dict(x_samples=n, x_delta=dx, t_samples=150, t_delta=dt),
        material=('AcousticMaterial2ndOrder', dict(sound_velocity=1500.0, density=1000.0,
                                                   shear_viscosity=1e-3,
                                                   d_rho_p=1500.0 ** 2, d_rho2_p=2.0e7)),
        regions=[], init={},
        boundaries={'pressure': [dict(indices=[5], value=1e5 * _ig_signal(150) / 200.0,
                                      additive=True)]},
        outputs={'pressure': [[20, 40]], 'velocity': [[30]], 'density': [[25]]},
        plan=[('simulate', None)])


def scene_bounds_semantics():
    nx, ny, dx = 12, 10, 1e-4
    dt = 0.5 * dx / 5900.0
    rng = np.random.default_rng(14)
    nt = 40
    per_point = [rng.standard_normal(nt + 10) for _ in range(3)]
    return dict(
        name='bounds_semantics', model='Acoustic2D',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=nt, t_delta=dt),
        material=('AcousticMaterial', dict(WATER)),
        regions=[(rect_indices(nx, 2, 2, 4, 3), ('AcousticMaterial', dict(STEEL))),
                 (rect_indices(nx, 4, 3, 4, 3), ('AcousticMaterial',
                                                 dict(sound_velocity=800.0, density=900.0)))],
        init={'pressure': rng.standard_normal(nx * ny)},
        boundaries={
            'pressure': [
                dict(indices=[13, 14, 15], value=per_point, additive=True),
                dict(indices=[14, 40, 14, 41], value=rng.standard_normal(nt + 10),
                     additive=True),
                dict(indices=[15, 41], value=0.25, additive=True),
                dict(indices=[70, 71, 70], value=-1.5, additive=False)],
            'velocity_x': [dict(indices=[nx * y + nx - 1 for y in range(ny)], value=0,
                                additive=False),
                           dict(indices=[nx * 5 + 11, 30],
                                value=rng.standard_normal(nt + 10), additive=True)],
            'velocity_y': [dict(indices=list(range(nx)), value=0.0, additive=False)]},
        outputs={'pressure': [[13, 14, 40], [70, 119, 0]], 'velocity_x': [[30, nx * 5 + 11]],
                 'velocity_y': [[5, 50]]},
        plan=[('simulate', 7), ('sim_step',), ('simulate', 5), ('reset',), ('simulate', 0)])


def scene_a2d_wrap():
    nx, ny, dx = 7, 5, 1e-4
    dt = 0.5 * dx / 5900.0
    rng = np.random.default_rng(15)
    return dict(
        name='a2d_wrap', model='Acoustic2D',
        kwargs=dict(x_samples=nx, x_delta=dx, y_samples=ny, y_delta=dx,
                    t_samples=25, t_delta=dt),
        material=('AcousticMaterial', dict(WATER)),
        regions=[(rect_indices(nx, 4, 1, 3, 3), ('AcousticMaterial', dict(STEEL)))],
        init={'pressure': rng.standard_normal(nx * ny),
              'velocity_x': 1e-4 * rng.standard_normal(nx * ny),
              'velocity_y': 1e-4 * rng.standard_normal(nx * ny)},
        boundaries={}, outputs={'pressure': [list(range(nx * ny))]},
        plan=[('simulate', None)])


SCENES = {f.__name__[len('scene_'):]: f for f in [
    scene_a1d_rigid, scene_a1d_lossy_src, scene_a2d_main, scene_a2d_lossy,
    scene_axi_lossless, scene_axi_lossy, scene_t2d_blocks, scene_t2d_aniso, scene_t1d,
    scene_taxi, scene_a2o1d, scene_bounds_semantics, scene_a2d_wrap]}
SCENES['ig1d_cnl'] = lambda: scene_ig1d(True, True)
SCENES['ig1d_nlin'] = lambda: scene_ig1d(False, False)

# scenes whose reference results depend on the host's numpy power (IdealGas1D's nonlinear
# equation of state): their fixtures exist twice, <name>.npz as numpy computed it on an
# AVX-512 host (SVML) and <name>_libm.npz with numpy's AVX-512 dispatch disabled (glibc pow)
POW_SCENES = ('ig1d_cnl',)


def all_scenes():
    return {name: make() for name, make in SCENES.items()}


# --------------------------------------------------------------------------------------
# fingerprint (guards fixtures against scene drift)
# --------------------------------------------------------------------------------------
def _canon(obj):
    if isinstance(obj, np.ndarray):
        return {'nd': hashlib.sha256(np.ascontiguousarray(obj, dtype=float).tobytes())
                .hexdigest(), 'shape': list(obj.shape)}
    if isinstance(obj, dict):
        return {str(k): _canon(v) for k, v in sorted(obj.items())}
    if isinstance(obj, (list, tuple)):
        return [_canon(v) for v in obj]
    if isinstance(obj, (np.floating, np.integer)):
        return obj.item()
    return obj


def fingerprint(scene):
    return hashlib.sha256(json.dumps(_canon(scene), sort_keys=True).encode()).hexdigest()[:16]


# --------------------------------------------------------------------------------------
# construction through a pyfds-compatible namespace
# --------------------------------------------------------------------------------------
def make_material(ns, spec):
    cls, kw = spec
    return getattr(ns, cls)(**kw)


def build_field(ns, regions_mod, scene):
    """Build ``scene`` through the public API of ``ns`` (``pyfds`` or
    ``pyfds_amd``); ``regions_mod`` is that package's ``regions`` module."""
    kw = dict(scene['kwargs'])
    kw['material'] = make_material(ns, scene['material'])
    fld = getattr(ns, scene['model'])(**scene.get('model_kwargs', {}), **kw)
    for idx, spec in scene['regions']:
        fld.material_regions.append(regions_mod.MaterialRegion(
            regions_mod.Region(list(idx), 'inclusion'), make_material(ns, spec)))
    for comp, vals in scene['init'].items():
        getattr(fld, comp).values = np.array(vals, dtype=float)
    for comp, bl in scene['boundaries'].items():
        for b in bl:
            getattr(fld, comp).boundaries.append(regions_mod.Boundary(
                regions_mod.Region(list(b['indices']), 'b'), value=b['value'],
                additive=b['additive']))
    for comp, ol in scene['outputs'].items():
        for idx in ol:
            getattr(fld, comp).add_output(regions_mod.Region(list(idx), 'o'))
    return fld


def run_plan(fld, scene):
    for action in scene['plan']:
        if action[0] == 'simulate':
            fld.simulate(action[1])
        elif action[0] == 'sim_step':
            if not fld.matrices_assembled:
                fld.assemble_matrices()
            fld.sim_step()
            fld.step += 1
        elif action[0] == 'reset':
            fld.reset()
        else:
            raise ValueError(action)
    return fld


def collect(fld, scene):
    """Final component values and every output's signals, as arrays."""
    res = {'step': np.array(fld.step)}
    for comp in COMPONENTS[scene['model']]:
        res['values/' + comp] = np.array(getattr(fld, comp).values, dtype=float)
        for k, out in enumerate(getattr(fld, comp).outputs):
            res['signals/%s/%d' % (comp, k)] = np.array(out.signals, dtype=float)
    return res


# --------------------------------------------------------------------------------------
# the same scene on the oracle
# --------------------------------------------------------------------------------------
def material_params(spec):
    """Attribute values the reference material classes expose
    (``pyfds/acoustics.py:239-286``, ``pyfds/thermal.py:179-211``,
    ``pyfds/nonlinear_acoustics.py:174-190``) as a plain dict."""
    from oracle import absorption_coef
    cls, kw = spec
    if cls in ('AcousticMaterial', 'AcousticMaterial2ndOrder'):
        p = dict(shear_viscosity=0, bulk_viscosity=0, thermal_conductivity=0,
                 isobaric_heat_cap=1, isochoric_heat_cap=1)
        p.update(kw)
        p['absorption_coef'] = absorption_coef(p)
        if cls == 'AcousticMaterial2ndOrder':
            if not kw.get('d_rho_p'):
                p['d_rho_p'] = p['sound_velocity'] ** 2
            p.setdefault('d_rho2_p', 0)
        return p
    if cls == 'ThermalMaterial':
        k = kw['thermal_conductivity']
        kx, ky = (k if isinstance(k, (list, tuple)) else (k, k))
        return dict(heat_capacity=kw['heat_capacity'], density=kw['density'],
                    thermal_conductivity_x=kx, thermal_conductivity_y=ky)
    raise ValueError(cls)


def build_oracle(scene, power='numpy'):
    import oracle
    kw = scene['kwargs']
    nx, ny = _grid(scene)
    mats = [(list(range(nx * ny)), material_params(scene['material']))]
    mats += [(list(idx), material_params(spec)) for idx, spec in scene['regions']]
    cls = oracle.MODELS[scene['model']]
    extra = scene.get('model_kwargs', {})
    fld = cls(nx, kw['x_delta'], kw['t_samples'], kw['t_delta'], mats,
              ny=ny, dy=kw.get('y_delta', 1.0), **extra)
    if hasattr(fld, 'power'):
        fld.power = power
    for comp, vals in scene['init'].items():
        fld.components[comp] = np.array(vals, dtype=float)
    for comp, bl in scene['boundaries'].items():
        fld.boundaries[comp] = [dict(b) for b in bl]
    for comp, ol in scene['outputs'].items():
        fld.outputs[comp] = [dict(indices=list(idx), signals=[]) for idx in ol]
    return fld


def run_oracle(scene, power='numpy'):
    """Run a scene on the oracle; ``power`` picks IdealGas1D's ``**`` ('numpy' or 'cr')."""
    fld = build_oracle(scene, power)
    for action in scene['plan']:
        if action[0] == 'simulate':
            fld.simulate(action[1])
        elif action[0] == 'sim_step':
            if not fld.assembled:
                fld.assemble()
            fld.sim_step()
            fld.step += 1
        elif action[0] == 'reset':
            fld.reset()
    res = {'step': np.array(fld.step)}
    for comp in COMPONENTS[scene['model']]:
        res['values/' + comp] = np.array(fld.components[comp], dtype=float)
        for k, out in enumerate(fld.outputs[comp]):
            res['signals/%s/%d' % (comp, k)] = np.array(out['signals'], dtype=float)
    return res
