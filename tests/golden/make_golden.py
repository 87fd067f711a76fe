"""Write golden fixtures by running the reference pyfds on the scenes in
tests/scenes.py and tests/coupled_scenes.py.

Run in the development container only (it reads the reference checkout):

    python tests/golden/make_golden.py /root/reference

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
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def main(reference_root):
    import matplotlib
    matplotlib.use('Agg')
    sys.path.insert(0, reference_root)
    import pyfds
    import pyfds.regions as regions
    import scenes

    for name, make in scenes.SCENES.items():
        scene = make()
        fld = scenes.build_field(pyfds, regions, scene)
        scenes.run_plan(fld, scene)
        res = scenes.collect(fld, scene)
        res['fingerprint'] = np.array(scenes.fingerprint(scene))
        path = os.path.join(HERE, name + '.npz')
        np.savez_compressed(path, **res)
        print('wrote', os.path.relpath(path), 'step', int(res['step']))

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
    main(sys.argv[1] if len(sys.argv) > 1 else '/root/reference')
