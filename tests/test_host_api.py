"""Host-side checks that need no GPU: the C-ABI library loads and exports every entry point
of include/pyfds_amd.h; the pyfds-compatible API reproduces the reference's own tests
(test/test_fields.py, test/test_acoustics.py, test/test_regions.py, test/test_coupling.py
matrix checks), restated here against pyfds_amd; boundary compilation follows the
reference's semantics; simulation without a GPU fails loudly."""
import os
import re

import numpy as np
import pytest

import oracle
import pyfds_amd as fds
import pyfds_amd.fields as fls
import pyfds_amd.regions as reg
from pyfds_amd import _native, device

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


# ------------------------------------------------------------------ C-ABI library
def test_library_exports_every_declared_symbol():
    lib = _native.load()
    header = open(os.path.join(ROOT, 'include', 'pyfds_amd.h')).read()
    declared = set(re.findall(r'^\s*(?:int|const char \*)\s*(fds_\w+)\s*\(', header, re.M))
    assert declared, 'no declarations parsed'
    for name in declared:
        assert hasattr(lib, name), name
    assert set(_native.EXPORTS) == declared
    assert lib.fds_version().startswith(b'pyfds_amd')


def test_pattern_bench_rejects_bad_arguments():
    """The calibration entry point checks its arguments before touching the GPU."""
    import ctypes as ct
    lib = _native.load()
    ms, b = ct.c_double(), ct.c_double()
    good = dict(nx=512, ny=256, elem=8, stages=3, ghost=2, strip=12, ns=3, reps=1)
    for bad in (dict(elem=6), dict(stages=1), dict(ns=2), dict(nx=511), dict(strip=0),
                dict(ghost=-1), dict(reps=0), dict(nx=1 << 15, ny=1 << 14), dict(elem=4, nx=514)):
        a = dict(good, **bad)
        rc = lib.fds_stream_pattern_bench(a['nx'], a['ny'], a['elem'], a['stages'], a['ghost'],
                                          a['strip'], a['ns'], a['reps'], ct.byref(ms),
                                          ct.byref(b))
        assert rc == -1, bad


def test_tile_table():
    shapes, engines = _native.tile_shapes(), _native.tile_engines()
    assert (64, 16) in shapes and (256, 1) in shapes
    assert len(engines) == len(shapes) and 'stream' in engines
    # LDS-tile shapes come first, the streaming shapes after them
    first = engines.index('stream')
    assert all(e == 'tile' for e in engines[:first]) and all(e == 'stream' for e in engines[first:])
    assert all(ty > 1 for _, ty in shapes[first:])
    lib = _native.load()
    assert lib.fds_tile_engine(len(shapes)) < 0 and lib.fds_tile_engine(-1) < 0
    # the streaming table the GPU tests pick slots from: (tx, ty, steps per launch, nt,
    # prefetch depth, LDS ring)
    stream = [(d['tx'], d['ty'], d['steps'], d['nontemporal'], d['prefetch'], d['lds_ring'])
              for d in _native.tile_table() if d['engine'] == 'stream']
    F, T = False, True
    assert stream == [(256, 16, 1, F, 1, F), (256, 16, 2, F, 1, F), (128, 32, 2, F, 1, F),
                      (256, 32, 3, F, 1, F), (256, 16, 1, T, 1, F), (256, 16, 2, T, 1, F),
                      (256, 32, 3, T, 1, F), (64, 16, 1, F, 1, F), (64, 16, 2, F, 1, F),
                      (64, 32, 3, F, 1, F), (64, 16, 1, T, 1, F), (64, 16, 2, T, 1, F),
                      (64, 32, 3, F, 2, T), (64, 16, 2, F, 2, T), (64, 32, 3, F, 1, T)]
    assert [d['single_only'] for d in _native.tile_table() if d['engine'] == 'stream'] == \
        [F] * 12 + [T, T, F]
    assert all(d['steps'] == 1 and not d['nontemporal'] and d['prefetch'] == 0
               and not d['lds_ring'] for d in _native.tile_table() if d['engine'] == 'tile')


def test_simulation_without_gpu_fails_loudly():
    if os.environ.get('HIP_VISIBLE_DEVICES') is None and _gpu_present():
        pytest.skip('a GPU is present')
    fld = fds.Acoustic2D(x_samples=4, x_delta=1, y_samples=3, y_delta=1, t_samples=2,
                         t_delta=0.1, material=fds.AcousticMaterial(1, 1))
    with pytest.raises(_native.NativeError):
        fld.simulate()


def _gpu_present():
    import ctypes as ct
    lib = _native.load()
    ctx = ct.c_void_p()
    rc = lib.fds_create(ct.byref(ctx), 2, 0, 2, 2, 0)
    if rc == 0:
        lib.fds_destroy(ctx)
        return True
    return False


# ------------------------------------------------------------------ reference tests, restated
def test_acoustic_material():
    water = fds.AcousticMaterial(1500, 1000)
    water.bulk_viscosity = 1e-3
    water.shear_viscosity = 1e-3
    assert np.isclose(water.absorption_coef, 7e-3 / 3)
    water.absorption_coef = 3e-3
    assert np.isclose(water.absorption_coef, 3e-3)
    assert np.isclose(fds.AcousticMaterial(1500, 1000, absorption_coef=2e-3).absorption_coef,
                      2e-3)


def test_acoustic1d_create_matrices():
    fld = fds.Acoustic1D(t_delta=1, t_samples=1, x_delta=1, x_samples=3,
                         material=fds.AcousticMaterial(700, 0.01, bulk_viscosity=1))
    assert fld.a_p_v is None
    fld.assemble_matrices()
    assert np.allclose(fld.a_p_v.toarray(), [[-4900, 4900, 0], [0, -4900, 4900], [0, 0, -4900]])
    assert np.allclose(fld.a_v_p.toarray(), [[100, 0, 0], [-100, 100, 0], [0, -100, 100]])
    assert np.allclose(fld.a_v_v.toarray(), [[-200, 100, 0], [100, -200, 100], [0, 100, -200]])


def test_acoustic2d_create_matrices():
    fld = fds.Acoustic2D(t_delta=1, t_samples=1, x_delta=1, x_samples=2, y_delta=1, y_samples=2,
                         material=fds.AcousticMaterial(700, 0.01, bulk_viscosity=1))
    fld.assemble_matrices()
    assert np.allclose(fld.a_p_vx.toarray(), [[-4900, 4900, 0, 0], [0, -4900, 4900, 0],
                                              [0, 0, -4900, 4900], [0, 0, 0, -4900]])
    assert np.allclose(fld.a_p_vy.toarray(), [[-4900, 0, 4900, 0], [0, -4900, 0, 4900],
                                              [0, 0, -4900, 0], [0, 0, 0, -4900]])
    assert np.allclose(fld.a_vx_p.toarray(), [[100, 0, 0, 0], [-100, 100, 0, 0],
                                              [0, -100, 100, 0], [0, 0, -100, 100]])
    assert np.allclose(fld.a_vy_p.toarray(), [[100, 0, 0, 0], [0, 100, 0, 0],
                                              [-100, 0, 100, 0], [0, -100, 0, 100]])
    v = [[-400, 100, 100, 0], [100, -400, 100, 100], [100, 100, -400, 100], [0, 100, 100, -400]]
    assert np.allclose(fld.a_vx_vx.toarray(), v)
    assert fld.a_vy_vy is fld.a_vx_vx


def test_acoustic3d_axi_create_matrices():
    fld = fds.Acoustic3DAxi(t_delta=1, t_samples=1, x_delta=1, x_samples=2, y_delta=1,
                            y_samples=2, material=fds.AcousticMaterial(1, 1, bulk_viscosity=1))
    fld.assemble_matrices()
    assert np.allclose(fld.a_p_vx.toarray(), [[-2, 2 / 3, 0, 0], [0, -2 / 3, 2, 0],
                                              [0, 0, -2, 2 / 3], [0, 0, 0, -2 / 3]])
    assert np.allclose(fld.a_p_vy.toarray(), [[-1, 0, 1, 0], [0, -1, 0, 1],
                                              [0, 0, -1, 0], [0, 0, 0, -1]])
    assert np.allclose(fld.a_vx_vx.toarray(), [[-4, 4 / 3, 1, 0], [0, -4, 2, 1],
                                               [1, 2 / 3, -4, 4 / 3], [0, 1, 0, -4]])
    assert fds.AcousticAS is fds.Acoustic3DAxi


def test_thermal1d_matrix():
    ths = fds.Thermal1D(t_delta=1, t_samples=1, x_delta=1, x_samples=3,
                        material=fds.ThermalMaterial(1, 1, 1))
    ths.assemble_matrices()
    assert np.allclose(ths.a_t_q.toarray(), [[-1, 1, 0], [0, -1, 1], [0, 0, -1]])


def test_dimension():
    dim = fls.Dimension(3, 0.1)
    assert np.allclose(dim.vector, [0, 0.1, 0.2])
    assert dim.get_index(0.1) == 1
    with pytest.raises(AssertionError):
        dim.get_index(0.15)


def test_field_component_boundaries():
    fc = fls.FieldComponent(100)
    fc.values = np.random.default_rng(0).random(100)
    fc.boundaries = [reg.Boundary(reg.LineRegion([5, 6, 7], [0, 0.2], 'test boundary'))]
    fc.boundaries[0].value = 23
    fc.apply_bounds(step=0)
    assert np.allclose(fc.values[[5, 6, 7]], [23, 23, 23])
    fc = fls.FieldComponent(100)
    fc.values = np.ones(100)
    fc.boundaries = [reg.Boundary(reg.LineRegion([5, 6, 7], [0, 0.2], 'b'))]
    fc.boundaries[0].value = [np.arange(0, 3) * 23, np.arange(0, 3) * 42, np.arange(0, 3) * 23]
    fc.boundaries[0].additive = True
    fc.apply_bounds(step=1)
    assert np.allclose(fc.values[[5, 6, 7]], [24, 43, 24])
    fc = fls.FieldComponent(100)
    fc.values = np.ones(100)
    fc.boundaries = [reg.Boundary(reg.LineRegion([5, 6, 7], [0, 0.2], 'b'),
                                  value=np.arange(0, 3) * 23, additive=True)]
    fc.apply_bounds(step=2)
    assert np.allclose(fc.values[[5, 6, 7]], [47, 47, 47])


def test_field_component_output():
    fc = fls.FieldComponent(100)
    fc.outputs = [reg.Output(reg.LineRegion([0, 1, 2], [0, 0.2], 'test output'))]
    fc.write_outputs()
    fc.write_outputs()
    assert np.allclose(fc.outputs[0].signals, [[0, 0], [0, 0], [0, 0]])
    assert np.allclose(fc.outputs[0].mean_signal, np.zeros(2))
    out = reg.Output(reg.LineRegion([0, 1, 2], [0, 0.2], 'test output'))
    out.signals = [np.linspace(0, 1) for _ in range(3)]
    assert np.allclose(out.mean_signal, np.linspace(0, 1))


def test_fields_init_and_operators():
    fld = fls.Field1D(100, 0.1, 100, 0.1, int(5))
    assert np.allclose(fld.material_vector('real'), 5)
    fld = fls.Field1D(3, 1, 3, 1, int(5))
    assert np.allclose(fld.d_x().toarray(), [[-1, 1, 0], [0, -1, 1], [0, 0, -1]])
    assert np.allclose(fld.d_x(variant='backward').toarray(), [[1, 0, 0], [-1, 1, 0], [0, -1, 1]])
    assert np.allclose(fld.d_x(variant='central').toarray(),
                       [[0, 0.5, 0], [-0.5, 0, 0.5], [0, -0.5, 0]])
    assert np.allclose(fld.d_x2().toarray(), [[-2, 1, 0], [1, -2, 1], [0, 1, -2]])
    with pytest.raises(ValueError):
        fld.d_x(variant='sideways')
    fld = fls.Field2D(100, 0.1, 100, 0.1, 100, 0.1, int(5))
    assert np.allclose(fld.material_vector('real'), 5)
    assert np.size(fld.material_vector('real')) == 10000
    with pytest.raises(KeyError):
        fld.material_vector('sound_velocity')
    fld = fls.Field2D(2, 1, 2, 1, 10, 1, int(5))
    assert np.allclose(fld.d_y().toarray(), [[-1, 0, 1, 0], [0, -1, 0, 1],
                                             [0, 0, -1, 0], [0, 0, 0, -1]])
    assert np.allclose(fld.d_y(variant='backward').toarray(), [[1, 0, 0, 0], [0, 1, 0, 0],
                                                               [-1, 0, 1, 0], [0, -1, 0, 1]])
    assert np.allclose(fld.d_y(variant='central').toarray(),
                       [[0, 0, 0.5, 0], [0, 0, 0, 0.5], [-0.5, 0, 0, 0], [0, -0.5, 0, 0]])
    assert np.allclose(fld.d_y2().toarray(), [[-2, 0, 1, 0], [0, -2, 0, 1],
                                              [1, 0, -2, 0], [0, 1, 0, -2]])
    assert np.allclose(fld.d_x2().toarray(), [[-2, 1, 0, 0], [1, -2, 1, 0],
                                              [0, 1, -2, 1], [0, 0, 1, -2]])


def test_index_and_position():
    fld = fls.Field2D(4, 0.1, 3, 0.1, 1, 1, int(5))
    assert fld.get_index((0.2, 0.1)) == 6
    assert np.allclose(fld.get_position(fld.get_index((0.2, 0.1))), (0.2, 0.1))
    fld = fls.Field1D(4, 0.1, 1, 1, int(5))
    assert np.allclose(fld.get_position(fld.get_index(0.1)), 0.1)


def test_regions():
    fld = fls.Field1D(4, 0.1, 1, 1, int(5))
    fld.material_regions.append(reg.MaterialRegion(fld.get_line_region((0.1, 0.2)), int(23)))
    assert np.allclose(fld.material_vector('real'), [5, 23, 23, 5])
    fld = fls.Field2D(3, 1, 4, 0.5, 1, 1, int(5))
    assert np.allclose(fld.get_line_region((1, 0, 1, 1.5)).indices, [1, 4, 7, 10])
    assert np.allclose(fld.get_line_region((0, 0, 2, 0)).indices, [0, 1, 2])
    assert np.allclose(fld.get_line_region((0, 0, 2, 1.5)).indices, [0, 4, 7, 11])
    assert np.allclose(fld.get_line_region((0, 1.5, 2, 0)).indices, [9, 7, 4, 2])
    assert np.allclose(fld.get_rect_region((0, 0, 1, 1)).indices, [0, 3, 6, 1, 4, 7])
    assert np.allclose(fld.get_rect_region((2, 1.5, -1, -1)).indices, [4, 7, 10, 5, 8, 11])


def _tri_loop(fld, position):
    """The reference's scalar loop (pyfds/fields.py:443-502) on a small grid."""
    if (position[2] - position[0]) * (position[5] - position[1]) - \
            (position[3] - position[1]) * (position[4] - position[0]) > 0:
        position = (position[0], position[1], position[4], position[5], position[2], position[3])
    p = position

    def e(c, a, b):
        return (c[0] - p[a]) * (p[b + 1] - p[a + 1]) - (c[1] - p[a + 1]) * (p[b] - p[a])
    corners = (fld.get_index(p[:2]), fld.get_index(p[2:4]), fld.get_index(p[4:]))
    out = []
    for ii in range(min(corners), max(corners) + 1):
        pt = fld.get_position(ii)
        if e(pt, 0, 2) >= 0 and e(pt, 2, 4) >= 0 and e(pt, 4, 0) >= 0:
            out.append(ii)
    return out


def test_tri_and_ellipse_regions_match_scalar_loops():
    fld = fls.Field2D(21, 0.5, 17, 0.5, 1, 1, int(5))
    for pos in [(0, 0, 5, 0, 0, 5), (1, 1, 9, 3, 4, 7.5), (9, 3, 1, 1, 4, 7.5)]:
        assert list(fld.get_tri_region(pos).indices) == _tri_loop(fld, pos)
    centre, radii = (5, 4), (3, 2.5)
    expect = []
    lo = fld.get_index((centre[0], centre[1] - radii[1]))
    hi = fld.get_index((centre[0], centre[1] + radii[1]))
    for ii in range(lo, hi):
        x, y = fld.get_position(ii)
        if (centre[0] - x) ** 2 / radii[0] ** 2 + (centre[1] - y) ** 2 / radii[1] ** 2 <= 1:
            expect.append(ii)
    assert list(fld.get_ellipse_region(centre, radii).indices) == expect


def test_thermal_material():
    m = fds.ThermalMaterial(1, 2, (3, 4))
    assert (m.thermal_conductivity_x, m.thermal_conductivity_y) == (3, 4)
    m.thermal_conductivity = 5
    assert m.thermal_conductivity == (5, 5)
    with pytest.raises(ValueError):
        m.thermal_conductivity = 'x'


def test_second_order_material_default():
    m = fds.AcousticMaterial2ndOrder(sound_velocity=1500.0, density=1000.0)
    assert m.d_rho_p == 1500.0 ** 2 and m.d_rho2_p == 0
    with pytest.raises(RuntimeError):
        fds.Acoustic2ndOrder1D(x_samples=4, x_delta=1, t_samples=1, t_delta=1,
                               material=fds.AcousticMaterial(1, 1))


# ------------------------------------------------------------------ boundary compiler
@pytest.mark.parametrize('seed', range(8))
def test_boundary_programs_follow_reference_semantics(seed):
    """Evaluating the compiled per-cell programs gives exactly what the reference's
    sequential gather/compute/scatter gives (oracle.apply_boundaries)."""
    rng = np.random.default_rng(seed)
    n, steps = 40, 12
    bl = []
    for _ in range(int(rng.integers(1, 6))):
        k = int(rng.integers(1, 6))
        idx = [int(v) for v in rng.integers(-n, n, size=k)]
        kind = rng.integers(0, 4)
        if kind == 0:
            value = float(rng.standard_normal())
        elif kind == 1:
            value = rng.standard_normal(steps)
        elif kind == 2:
            value = [rng.standard_normal(steps) for _ in range(k)]
        else:
            value = rng.standard_normal((steps, k))
        bl.append(reg.Boundary(reg.Region(idx), value=value, additive=bool(rng.integers(0, 2))))
    prog = device.compile_boundaries(bl, n)
    vals = rng.standard_normal(n)
    for step in range(steps):
        want = vals.copy()
        oracle.apply_boundaries(want, [dict(indices=b.region.indices, value=b.value,
                                            additive=b.additive) for b in bl], step)
        got = vals.copy()
        for u, cell in enumerate(prog.cells):
            x = got[cell]
            for e in range(prog.ptr[u], prog.ptr[u + 1]):
                x = prog.a[e] * x + prog.table[prog.off[e] + step * prog.stride[e]]
            got[cell] = x
        np.testing.assert_array_equal(got, want)


def test_upload_change_detection_sees_every_edit():
    """A run re-uploads a boundary program or output list when anything the reference reads
    from them at run time changed since the last upload (pyfds/fields.py:591-611 reads the
    region's indices and the boundary's value and additive flag every step): edits made in
    place to an index list or a signal array count, values equal to the old ones do not."""
    idx = [5, 6, 7]
    sig = np.arange(4.0)
    b = reg.Boundary(reg.LineRegion(idx, [0, 0.2], 'b'), value=sig)
    o = reg.Output(reg.LineRegion(list(range(64)), [0, 0.2], 'o'))
    bs, os_ = device._snapshot([b], True), device._snapshot([o])
    assert device._unchanged([b], bs, True) and device._unchanged([o], os_, False)
    # an index list edited in place, grown, or holding the same values as numpy ints
    idx[1] = 9
    assert not device._unchanged([b], bs, True)
    idx[1] = 6
    assert device._unchanged([b], bs, True)
    idx[1] = np.int64(6)
    assert device._unchanged([b], bs, True)
    idx.append(8)
    assert not device._unchanged([b], bs, True)
    idx.pop()
    o.region.indices[63] = 0
    assert not device._unchanged([o], os_, False)
    o.region.indices[63] = 63
    assert device._unchanged([o], os_, False)
    # indices held as an array instead of the list
    o.region.indices = np.arange(64)
    assert device._unchanged([o], os_, False)
    o.region.indices[0] = 1
    assert not device._unchanged([o], os_, False)
    # a signal edited in place, replaced by an equal copy, by one of another dtype or length
    sig[2] = 7.0
    assert not device._unchanged([b], bs, True)
    sig[2] = 2.0
    assert device._unchanged([b], bs, True)
    b.value = sig.copy()
    assert device._unchanged([b], bs, True)
    b.value = sig.astype(np.float32)
    assert not device._unchanged([b], bs, True)
    b.value = sig[:3].copy()
    assert not device._unchanged([b], bs, True)
    b.value = sig
    b.additive = True
    assert not device._unchanged([b], bs, True)
    b.additive = False
    assert device._unchanged([b], bs, True)
    # another region object, or another list of boundaries
    b.region = reg.LineRegion([5, 6, 7], [0, 0.2], 'b')
    assert not device._unchanged([b], bs, True)
    assert not device._unchanged([], bs, True)


def test_boundary_errors():
    with pytest.raises(IndexError):
        device.compile_boundaries([reg.Boundary(reg.Region([5]), 1.0)], 5)
    with pytest.raises(ValueError):
        device.compile_boundaries([reg.Boundary(reg.Region([1, 2]), [np.zeros(3)])], 5)
    prog = device.compile_boundaries([reg.Boundary(reg.Region([1]), np.zeros(7)),
                                      reg.Boundary(reg.Region([2]), 3.0)], 5)
    assert prog.valid_steps == 7


def test_material_fill_of_rectangles_equals_scatter():
    """The 2D-slice fill of column-by-column rectangles writes exactly the cells of a plain
    scatter, for rectangles, reversed and perturbed index lists and sorted lists alike."""
    rng = np.random.default_rng(0)
    for trial in range(120):
        nx, ny = int(rng.integers(260, 600)), int(rng.integers(260, 600))
        g = fds.Acoustic2D(x_samples=nx, x_delta=1.0, y_samples=ny, y_delta=1.0,
                           t_samples=2, t_delta=1e-3,
                           material=fds.AcousticMaterial(1, 1))
        x0, y0 = int(rng.integers(0, nx // 2)), int(rng.integers(0, ny // 2))
        w, h = int(rng.integers(nx // 3, nx - x0 + 1)), int(rng.integers(ny // 3, ny - y0 + 1))
        idx = (x0 + np.arange(w)[:, None] + (y0 + np.arange(h)[None, :]) * nx).ravel()
        kind = trial % 4
        if kind == 1:
            idx = idx[::-1].copy()
        elif kind == 2:
            idx = idx.copy()
            idx[int(rng.integers(0, idx.size))] = int(rng.integers(0, nx * ny))
        elif kind == 3:
            idx = np.sort(idx)
        a, b = np.zeros(nx * ny), np.zeros(nx * ny)
        g._fill(a, idx, 2.5)
        b[idx] = 2.5
        assert np.array_equal(a, b), (trial, kind)


def test_pipeline_sizes():
    """Segments of a pipelined run: whole graph chunks rounded to whole fused launches, about
    eight of them, no more than the caller's segment or 16 MiB of outputs each."""
    from pyfds_amd import device
    assert device.pipeline_sizes(2000, 2000, 50, 64) == [250] * 8
    assert device.pipeline_sizes(1000, 1000, 50, 4096) == [150] * 6 + [100]
    assert device.pipeline_sizes(150, 150, 5, 3) == [20] * 7 + [10]
    assert device.pipeline_sizes(150, 150, 50, 3) is None          # under four graph chunks
    assert device.pipeline_sizes(5000, 5000, 50, 0) is None        # nothing to overlap
    assert device.pipeline_sizes(5000, 300, 50, 8) == [300] * 16 + [200]   # caller's segment
    # the staging cap: 16 MiB of outputs per run
    assert device.pipeline_sizes(100000, 100000, 1, 1 << 16) == [32] * 3125
    # whole fused launches: chunks of 50 with three steps per launch are 48 steps
    assert device.pipeline_sizes(2000, 2000, 50, 64, 3) == [288] * 6 + [272]
    assert device.pipeline_sizes(210, 210, 50, 64, 3) == [48] * 4 + [18]
    assert device.pipeline_sizes(150, 150, 0, 4, 2) == [20] * 7 + [10]


def test_output_signals_merge_device_blocks_in_order():
    """Device runs queue (steps x points) blocks; reading ``signals`` turns them into the
    per-point lists the reference builds step by step (pyfds/fields.py:602-611), in order with
    host-appended samples, as numpy float64 scalars."""
    from pyfds_amd import regions as reg
    out = reg.Output(reg.Region([3, 5], 'o'))
    a = np.arange(6.0).reshape(3, 2)            # 3 steps, 2 points
    b = 10 + np.arange(4.0).reshape(2, 2)
    out._record(a)
    out._record(b)
    assert out.signals == [[0.0, 2.0, 4.0, 10.0, 12.0], [1.0, 3.0, 5.0, 11.0, 13.0]]
    assert all(type(v) is np.float64 for s in out.signals for v in s)
    out.signals[0].append(99.0)                  # host step, as write_outputs does
    out.signals[1].append(98.0)
    out._record(np.array([[7.0, 8.0]]))
    assert out.signals == [[0.0, 2.0, 4.0, 10.0, 12.0, 99.0, 7.0],
                           [1.0, 3.0, 5.0, 11.0, 13.0, 98.0, 8.0]]
    np.testing.assert_array_equal(out.mean_signal, np.mean(np.asarray(out.signals), axis=0))
    out._record(np.array([[1.0, 1.0]]))
    out.signals = []                             # assignment drops queued blocks
    assert out.signals == []
    out._record(np.array([[2.0, 3.0]]))
    assert out.signals == [[2.0], [3.0]]
