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
