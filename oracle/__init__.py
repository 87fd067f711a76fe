"""CPU oracle for the pyfds time-stepping hot path (test infrastructure).

This package is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it.  ``pyfds_amd`` never imports it and has no CPU fallback.

``oracle.plain`` restates pyfds' plain FP64 method (scipy DIA operators,
numpy element-wise updates, boundary/output semantics) in numpy; every
function cites the reference ``file:line`` it follows.
"""
from .plain import *  # noqa: F401,F403
