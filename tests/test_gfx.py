"""Animators (pyfds/gfx.py) over device frames.

CPU: construction, component lookup, scale handling and the metric prefixes.  GPU: a 1D and a
2D animation run with the Agg backend; the frames are the observed component after every
``steps_per_frame`` steps (the reference's ``_sim_function`` loop, ``pyfds/gfx.py:71-86``:
``simulate(steps_per_frame)`` then a copy of the component), bit-identical to the oracle run
in the same segments, and the recorded outputs equal the oracle's.
"""
import matplotlib

matplotlib.use('Agg')

import numpy as np  # noqa: E402
import pytest  # noqa: E402

import oracle  # noqa: E402
import pyfds_amd as fds  # noqa: E402
from pyfds_amd import gfx  # noqa: E402


def _oracle_of(f):
    """The oracle of an Acoustic1D / Acoustic2D field built through the API: its material
    regions, current component values, boundaries and outputs (index lists)."""
    mats = [(np.asarray(mr.region.indices),
             dict(sound_velocity=mr.materials[0].sound_velocity,
                  density=mr.materials[0].density,
                  absorption_coef=mr.materials[0].absorption_coef))
            for mr in f.material_regions]
    two_d = hasattr(f, 'y')
    cls = oracle.PlainAcoustic2D if two_d else oracle.PlainAcoustic1D
    kw = dict(ny=f.y.samples, dy=f.y.increment) if two_d else {}
    o = cls(f.x.samples, f.x.increment, f.t.samples, f.t.increment, mats, **kw)
    for name in o.components:
        comp = getattr(f, name)
        o.components[name] = np.array(comp.values, dtype=float)
        o.boundaries[name] = [dict(indices=list(b.region.indices), value=b.value,
                                   additive=b.additive) for b in comp.boundaries]
        o.outputs[name] = [dict(indices=list(out.region.indices), signals=[])
                           for out in comp.outputs]
    return o


def _a1d(steps=60):
    f = fds.Acoustic1D(x_samples=200, x_delta=1e-4, t_samples=steps, t_delta=3e-8,
                       material=fds.AcousticMaterial(1500.0, 1000.0))
    f.pressure.values = np.exp(-0.5 * ((np.arange(200) - 100) / 6.0) ** 2)
    f.velocity.add_boundary(fds.regions.PointRegion([0], 0.0, 'wall'), value=0, additive=False)
    f.pressure.add_output(f.get_point_region(5e-3, 'p'))
    f.pressure.add_output(f.get_line_region((1e-2, 1.5e-2), 'p line'))
    return f


def _a2d(steps=30):
    f = fds.Acoustic2D(x_samples=40, x_delta=1e-4, y_samples=30, y_delta=1e-4, t_samples=steps,
                       t_delta=3e-8, material=fds.AcousticMaterial(1500.0, 1000.0))
    ix, iy = np.meshgrid(np.arange(40), np.arange(30))
    f.pressure.values = np.exp(-0.5 * ((ix - 20) ** 2 + (iy - 15) ** 2) / 9.0).ravel()
    f.add_material_region(f.get_rect_region((1e-3, 1e-3, 5e-4, 5e-4)),
                          fds.AcousticMaterial(5900.0, 7850.0))
    f.pressure.add_boundary(f.get_point_region((3e-3, 2e-3)), value=0.0, additive=False)
    return f


def test_animator_setup_and_prefixes():
    f = _a1d()
    a = gfx.Animator1D(f, steps_per_frame=7.9, scale=2)
    assert a.observed_component in ('pressure', 'velocity')
    assert a.steps_per_frame == 7 and a.scale == (-2, 2)
    assert gfx.Animator1D(f, 'velocity', scale=(0, 1)).scale == (0, 1)
    with pytest.raises(KeyError):
        gfx.Animator1D(f, 'temperature')
    with pytest.raises(ValueError):
        gfx.Animator1D(f, scale=(1, 2, 3))
    assert gfx.get_prefix(2e-3) == ('m', 1e-3)
    assert gfx.get_prefix(5e-8) == ('n', 1e-9)
    assert gfx.get_prefix(12.0) == ('', 1.0)
    assert fds.Animator2D is gfx.Animator2D


@pytest.mark.gpu
@pytest.mark.parametrize('dim', [1, 2])
def test_animation_frames_match_plain_runs(dim, monkeypatch):
    make = _a1d if dim == 1 else _a2d
    cls = gfx.Animator1D if dim == 1 else gfx.Animator2D
    f = make()
    anim = cls(f, 'pressure', steps_per_frame=7, scale=1)
    shown = []
    real = anim._show_frames

    def spy(set_data):
        def record(data):
            shown.append(np.array(data))
            set_data(data)
        return real(record)

    monkeypatch.setattr(anim, '_show_frames', spy)
    monkeypatch.setattr(gfx.pp, 'pause', lambda t: None)
    monkeypatch.setattr(gfx.pp, 'show', lambda *a, **k: None)
    orc = _oracle_of(make())
    anim.start_simulation()
    nframes = f.t.samples // 7
    assert len(shown) == nframes and f.step == nframes * 7
    for k, frame in enumerate(shown):
        orc.simulate(7)                      # the oracle in the animator's segments
        np.testing.assert_array_equal(frame, orc.components['pressure'],
                                      err_msg='frame %d' % k)
    for out, rout in zip(f.pressure.outputs, orc.outputs['pressure']):
        np.testing.assert_array_equal(np.array(out.signals), np.array(rout['signals']))
    gfx.pp.close('all')
