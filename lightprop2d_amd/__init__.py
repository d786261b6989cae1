"""lightprop2d_amd: the Beam2D angular-spectrum hot path of lightprop2d, rebuilt for
AMD Instinct MI355X (gfx950) as hand-written HIP kernels behind a C ABI.

Drop-in surface (lightprop2d/__init__.py:7-8): ``Beam2D``, units ``nm, um, mm, cm, m``
and the profile generators.  Tensor-level operations are in
``lightprop2d_amd.propagation``; differentiable ones in ``lightprop2d_amd.autograd``.
"""
from .beam2d import Beam2D, nm, um, mm, cm, m
from .profiles import *  # noqa: F401,F403
from ._lib import Lp2dError

__version__ = "1.0.0"
