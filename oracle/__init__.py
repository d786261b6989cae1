"""CPU oracle for the Beam2D angular-spectrum path (test infrastructure only).

Nothing in ``lightprop2d_amd`` imports this package.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker or as the timed CPU restatement of the reference.
"""
from .angular_spectrum import *  # noqa: F401,F403
from . import profiles  # noqa: F401
