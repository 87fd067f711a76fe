"""pyfds_amd: MI355X-native time stepping behind pyfds' API.

Drop-in namespace for the reference's hot path (``pyfds/__init__.py:1-9``): the field
models, materials, regions, boundaries and outputs keep pyfds' names and signatures, and
``simulate()`` / ``sim_step()`` run hand-written gfx950 HIP stencil kernels through the
``libpyfds_amd.so`` C-ABI (``include/pyfds_amd.h``).
"""
from . import fields, regions  # noqa: F401
from .acoustic_flow import *  # noqa: F401,F403
from .acoustics import *  # noqa: F401,F403
from .coupled_fields import *  # noqa: F401,F403
from .coupling import *  # noqa: F401,F403
from .nonlinear_acoustics import *  # noqa: F401,F403
from .regions import *  # noqa: F401,F403
from .thermal import *  # noqa: F401,F403

__version__ = '0.1.0'


def __getattr__(name):
    # the animators (pyfds/gfx.py) pull in matplotlib.pyplot: imported on first use
    if name in ('Animator1D', 'Animator2D', 'gfx'):
        import importlib
        gfx = importlib.import_module(__name__ + '.gfx')
        return gfx if name == 'gfx' else getattr(gfx, name)
    raise AttributeError("module 'pyfds_amd' has no attribute {!r}".format(name))
