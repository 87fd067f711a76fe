"""Write golden fixtures by running the reference pyfds on the scenes in
tests/scenes.py and tests/coupled_scenes.py.

Run in the development container only (it reads the reference checkout):

    python tests/golden/make_golden.py /root/reference [--only NAME ...]

The scenes in ``scenes.POW_SCENES`` (IdealGas1D's nonlinear equation of state, whose numpy
power rounds differently on AVX-512 hosts) are written a second time as ``<name>_libm.npz``
by a child process that starts numpy with its AVX-512 dispatch disabled
(``NPY_DISABLE_CPU_FEATURES``), so numpy's power is glibc's ``pow`` there.

What it pins: time-evolved FP64 results of every hot-path model (component
values after the run plan and every recorded output signal), including
segmented ``simulate`` calls, ``sim_step``, ``reset`` and the boundary corner
cases in ``scene_bounds_semantics``.  The reference's own tests pin only the
operator matrices, so these end-to-end results are pinned by nothing else.

Each fixture ``<scene>.npz`` holds the results plus ``fingerprint`` (a hash of
the scene inputs) so a changed scene is detected instead of silently compared
against stale data.  Fixtures are data only: inputs are rebuilt from
tests/scenes.py.
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))


AVX512 = 'AVX512F AVX512CD AVX512_SKX AVX512_CLX AVX512_CNL AVX512_ICL'


def main(reference_root, only=None, libm=False):
    import matplotlib
    matplotlib.use('Agg')
    sys.path.insert(0, reference_root)
    import pyfds
    import pyfds.regions as regions
    import scenes
    from oracle import numpy_power_is_libm

    if libm and not numpy_power_is_libm():
        raise SystemExit('numpy power is not glibc pow in this process')
    names = scenes.POW_SCENES if libm else list(scenes.SCENES)
    for name in names:
        if only and name not in only:
            continue
        scene = scenes.SCENES[name]()
        fld = scenes.build_field(pyfds, regions, scene)
        scenes.run_plan(fld, scene)
        res = scenes.collect(fld, scene)
        res['fingerprint'] = np.array(scenes.fingerprint(scene))
        path = os.path.join(HERE, name + ('_libm' if libm else '') + '.npz')
        np.savez_compressed(path, **res)
        print('wrote', os.path.relpath(path), 'step', int(res['step']))
    if libm:
        return
    if not only or any(n in only for n in scenes.POW_SCENES):
        env = dict(os.environ, NPY_DISABLE_CPU_FEATURES=AVX512)
        subprocess.run([sys.executable, os.path.abspath(__file__), reference_root, '--libm']
                       + (['--only'] + list(only) if only else []), env=env, check=True)
    if only and not any(n.startswith('coupled_') for n in only):
        return

    # coupled fields and flow (tests/coupled_scenes.py)
    import coupled_scenes as cs
    for name, make in cs.SCENES.items():
        fld, plan, comps, params = make(pyfds)
        cs.run(fld, plan)
        res = cs.collect(fld, comps)
        res['fingerprint'] = np.array(cs.fingerprint(params))
        path = os.path.join(HERE, 'coupled_' + name + '.npz')
        np.savez_compressed(path, **res)
        print('wrote', os.path.relpath(path), 'step', int(res['step']))


if __name__ == '__main__':
    args = sys.argv[1:]
    libm = '--libm' in args
    only = None
    if '--only' in args:
        only = args[args.index('--only') + 1:]
        args = args[:args.index('--only')]
    args = [a for a in args if a != '--libm']
    main(args[0] if args else '/root/reference', only, libm)
