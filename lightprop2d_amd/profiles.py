"""Field and mask generators with the reference's signatures (lightprop2d/profiles.py).

Every generator takes the grid vectors ``x`` (n,) and ``y`` (n, 1) -- numpy arrays or
torch tensors (``Beam2D.X`` / ``Beam2D.Y`` are device tensors) -- and returns an array
of the same kind on the same device.  The values are computed on the host in FP64
with numpy, so masks and random fields are identical to the reference's numpy path;
random generators draw from numpy's legacy global RNG exactly as the reference does
(``np.random.seed(s)`` reproduces its streams), or from ``rng`` when given.
"""
from __future__ import annotations

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

__all__ = ('plane_wave', 'random_wave', 'random_wave_bin', 'gaussian_beam',
           'round_hole', 'random_round_hole', 'random_round_hole_bin',
           'random_round_hole_phase', 'rectangle_hole', 'square_hole',
           'square_slits')


def _is_torch(a):
    return torch is not None and isinstance(a, torch.Tensor)


def _host(a):
    if _is_torch(a):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _like(result, ref):
    """Return `result` (numpy) as the array kind of `ref` (torch -> same device)."""
    if _is_torch(ref):
        return torch.from_numpy(np.ascontiguousarray(result)).to(ref.device)
    return result


def _rng(rng):
    return np.random if rng is None else rng


def binning(field, binning_order: int):
    """profiles.py:29-32: Kronecker up-sampling by binning_order."""
    out = np.kron(_host(field), np.ones((binning_order, binning_order)))
    return _like(out, field)


def plane_wave(x, y):
    """profiles.py:35-38"""
    return _like(np.ones((len(y), len(x))), x)


def random_wave_bin(x, y, binning_order: int = 1, rng=None):
    """profiles.py:41-50: binary {0, 1} speckle, binned."""
    r = _rng(rng).randint(0, 2, size=(int(len(y) // binning_order),
                                       int(len(x) // binning_order)))
    return _like(np.kron(r, np.ones((binning_order, binning_order))), x)


def random_wave(x, y, binning_order: int = 1, rng=None):
    """profiles.py:53-61: uniform [0, 1) speckle, binned."""
    r = _rng(rng).uniform(size=(int(len(y) // binning_order),
                                 int(len(x) // binning_order)))
    return _like(np.kron(r, np.ones((binning_order, binning_order))), x)


def gaussian_beam(x, y, A0: float, rho0: float, x0: float = 0, y0: float = 0):
    """profiles.py:64-71: A0 * exp(-(x^2 + y^2)/rho0^2).

    Like the reference it shifts ``x`` and ``y`` in place by (x0, y0)."""
    x -= x0
    y -= y0
    xh, yh = _host(x), _host(y)
    return _like(A0 * np.exp(- (xh ** 2 + yh ** 2) / rho0 ** 2), x)


def round_hole(x, y, R: float, x0: float = 0, y0: float = 0):
    """profiles.py:74-81: exp(-r^2/R^2) >= exp(-1/2), i.e. a disc of radius R/sqrt(2)."""
    xh, yh = _host(x), _host(y)
    d = np.exp(- ((xh - x0) ** 2 + (yh - y0) ** 2) / R ** 2)
    return _like(d >= 1 / np.exp(0.5), x)


def _round_hole_host(xh, yh, R, x0, y0):
    return np.exp(- ((xh - x0) ** 2 + (yh - y0) ** 2) / R ** 2) >= 1 / np.exp(0.5)


def random_round_hole_bin(x, y, R: float, x0: float = 0, y0: float = 0,
                          binning_order: int = 1, rng=None):
    """profiles.py:84-91"""
    xh, yh = _host(x), _host(y)
    r = _rng(rng).randint(0, 2, size=(int(len(yh) // binning_order),
                                       int(len(xh) // binning_order)))
    field = np.kron(r, np.ones((binning_order, binning_order)))
    field[_round_hole_host(xh, yh, R, x0, y0) == 0] = 0
    return _like(field, x)


def random_round_hole(x, y, R: float, x0: float = 0, y0: float = 0,
                      binning_order: int = 1, rng=None):
    """profiles.py:94-101"""
    xh, yh = _host(x), _host(y)
    r = _rng(rng).uniform(size=(int(len(yh) // binning_order),
                                 int(len(xh) // binning_order)))
    field = np.kron(r, np.ones((binning_order, binning_order)))
    field[_round_hole_host(xh, yh, R, x0, y0) == 0] = 0
    return _like(field, x)


def random_round_hole_phase(x, y, R: float, x0: float = 0, y0: float = 0,
                            binning_order: int = 1, rng=None):
    """profiles.py:104-113: unit amplitude inside round_hole(R), phase 2*pi*U[0,1)."""
    xh, yh = _host(x), _host(y)
    r = _rng(rng).uniform(size=(int(len(yh) // binning_order),
                                 int(len(xh) // binning_order)))
    u = np.kron(r, np.ones((binning_order, binning_order)))
    hole = _round_hole_host(xh, yh, R, x0, y0)
    u[hole == 0] = 0
    angle = u * 2 * np.pi
    phase = np.cos(angle) + 1j * np.sin(angle)
    return _like(hole * phase, x)


def rectangle_hole(x, y, dx: float, dy: float, x0=0, y0=0):
    """profiles.py:116-121"""
    xh, yh = _host(x), _host(y)
    return _like((np.abs(xh - x0) < (dx / 2)) & (np.abs(yh - y0) < (dy / 2)), x)


def square_hole(x, y, d: float, x0: float = 0, y0: float = 0):
    """profiles.py:124-128"""
    return rectangle_hole(x, y, d, d, x0, y0)


def square_slits(x, y, d: float, slits_distance: float, x0: float = 0, y0: float = 0):
    """profiles.py:131-136"""
    dl = slits_distance / 2
    a = square_hole(x, y, d, x0 - dl, y0)
    b = square_hole(x, y, d, x0 + dl, y0)
    return a | b
