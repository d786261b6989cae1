"""Field and mask generators with the reference's signatures (lightprop2d/profiles.py).

Every generator takes the grid vectors ``x`` (n,) and ``y`` (n, 1), the way Beam2D
calls them (beam2d.py:302, 338).

* With torch device tensors (``Beam2D.X`` / ``Beam2D.Y``), the values are computed on
  the GPU by ``lp2d_profile`` (one HIP kernel) and returned as device tensors: float64,
  bool, or complex128 for the random phase screen, as the reference's numpy dtypes.
* With numpy arrays (or host torch tensors), they are computed with numpy as in the
  reference (and returned in the input's kind).

Random generators draw from numpy's legacy global RNG exactly as the reference does
(``np.random.seed(s)`` reproduces its streams), or from ``rng`` when given.  The draw
(n/b x n/b values) is the only host work; binning, the pupil and the phase run on the
device.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

try:
    import torch
except ImportError:  # pragma: no cover
    torch = None

__all__ = ('plane_wave', 'random_wave', 'random_wave_bin', 'gaussian_beam',
           'round_hole', 'random_round_hole', 'random_round_hole_bin',
           'random_round_hole_phase', 'rectangle_hole', 'square_hole',
           'square_slits')

# 1 / exp(0.5), the round_hole threshold (profiles.py:80), with numpy's exp
_THRESH = float(1 / np.exp(0.5))


def _is_torch(a):
    return torch is not None and isinstance(a, torch.Tensor)


def _host(a):
    if _is_torch(a):
        return a.detach().cpu().numpy()
    return np.asarray(a)


def _like(result, ref):
    """numpy `result` as the array kind of `ref`: a host torch tensor for host torch
    inputs (device inputs never get here), numpy otherwise."""
    if _is_torch(ref):
        return torch.from_numpy(np.ascontiguousarray(result))
    return result


def _rng(rng):
    return np.random if rng is None else rng


def _on_device(x, y):
    return any(_is_torch(a) and a.device.type == "cuda" for a in (x, y))


def _grid_vectors(x, y):
    """(x, y) device tensors of shape (n,)/(1, n) and (n, 1)/(n,) -> contiguous float64
    vectors on x's device."""
    from ._lib import Lp2dError
    if not (_is_torch(x) and _is_torch(y)):
        raise Lp2dError("device generators take x and y as torch tensors on one device")
    if x.device.type != "cuda" or y.device != x.device:
        raise Lp2dError("device generators need x and y on the same GPU")
    if x.dim() > 2 or y.dim() > 2 or (x.dim() == 2 and x.shape[0] != 1) or \
            (y.dim() == 2 and y.shape[1] != 1):
        raise Lp2dError("device generators take the grid vectors x (n,) and y (n, 1)")
    xv = x.reshape(-1).to(torch.float64).contiguous()
    yv = y.reshape(-1).to(torch.float64).contiguous()
    return xv, yv


def _profile(kind, x, y, rand=None, binning_order=1, **kw):
    """Run lp2d_profile over the broadcast grid (len(y), len(x))."""
    from . import _lib as L
    from .propagation import _lib as lib_loaded
    lib = lib_loaded()
    xv, yv = _grid_vectors(x, y)
    nx, ny = xv.numel(), yv.numel()
    k = L.PROF[kind]
    code = (L.LP2D_MASK_U8 if kind in ("round_hole", "rect_hole", "slits") else
            L.LP2D_MASK_C128 if kind == "random_phase" else L.LP2D_MASK_F64)
    tdt = {L.LP2D_MASK_U8: torch.bool, L.LP2D_MASK_F64: torch.float64,
           L.LP2D_MASK_C128: torch.complex128}[code]
    out = torch.empty((ny, nx), dtype=tdt, device=xv.device)
    prm = L.lp2d_profile_params(kw.get("a0", 0.0), kw.get("rho0", 1.0), kw.get("r", 0.0),
                                _THRESH, kw.get("x0", 0.0), kw.get("y0", 0.0),
                                kw.get("dx", 0.0), kw.get("dy", 0.0), kw.get("slits", 0.0),
                                int(binning_order), 0)
    rd = None
    if rand is not None:
        rd = torch.from_numpy(np.ascontiguousarray(rand, dtype=np.float64)).to(xv.device)
    with torch.cuda.device(xv.device):
        stream = torch.cuda.current_stream(xv.device).cuda_stream
        L.check(lib.lp2d_profile(k, nx, ny, C.c_void_p(xv.data_ptr()), C.c_void_p(yv.data_ptr()),
                                 C.byref(prm), C.c_void_p(rd.data_ptr() if rd is not None else 0),
                                 C.c_void_p(out.data_ptr()), code, C.c_void_p(stream)),
                "profile")
    return out


def _draw(kind, ny, nx, b, rng):
    """The reference's host draw: (ny//b, nx//b) from the legacy RNG stream."""
    size = (int(ny // b), int(nx // b))
    if kind == "int":
        return _rng(rng).randint(0, 2, size=size).astype(np.float64)
    return _rng(rng).uniform(size=size)


def _binned_shape_check(x, y, b):
    from ._lib import Lp2dError
    n_x, n_y = x.numel(), y.numel()
    if b < 1 or n_x % b or n_y % b:
        raise Lp2dError(f"binning order {b} must divide the grid sizes ({n_y}, {n_x}): the "
                        "binned draw would not broadcast against the pupil")


def binning(field, binning_order: int):
    """profiles.py:29-32: Kronecker up-sampling by binning_order (kron with ones)."""
    b = int(binning_order)
    if _is_torch(field):
        return field.to(torch.float64).repeat_interleave(b, 0).repeat_interleave(b, 1)
    return np.kron(field, np.ones((b, b)))


def plane_wave(x, y):
    """profiles.py:35-38"""
    if _on_device(x, y):
        return _profile("plane", x, y)
    return _like(np.ones((len(y), len(x))), x)


def random_wave_bin(x, y, binning_order: int = 1, rng=None):
    """profiles.py:41-50: binary {0, 1} speckle, binned."""
    if _on_device(x, y):
        b = int(binning_order)
        r = _draw("int", len(y), len(x), b, rng)
        nxb, nyb = r.shape[1] * b, r.shape[0] * b
        return _profile("random", x.reshape(-1)[:nxb], y.reshape(-1)[:nyb], rand=r,
                        binning_order=b)
    r = _rng(rng).randint(0, 2, size=(int(len(y) // binning_order),
                                       int(len(x) // binning_order)))
    return _like(np.kron(r, np.ones((binning_order, binning_order))), x)


def random_wave(x, y, binning_order: int = 1, rng=None):
    """profiles.py:53-61: uniform [0, 1) speckle, binned."""
    if _on_device(x, y):
        b = int(binning_order)
        r = _draw("uniform", len(y), len(x), b, rng)
        nxb, nyb = r.shape[1] * b, r.shape[0] * b
        return _profile("random", x.reshape(-1)[:nxb], y.reshape(-1)[:nyb], rand=r,
                        binning_order=b)
    r = _rng(rng).uniform(size=(int(len(y) // binning_order),
                                 int(len(x) // binning_order)))
    return _like(np.kron(r, np.ones((binning_order, binning_order))), x)


def gaussian_beam(x, y, A0: float, rho0: float, x0: float = 0, y0: float = 0):
    """profiles.py:64-71: A0 * exp(-(x^2 + y^2)/rho0^2).

    Like the reference it shifts ``x`` and ``y`` in place by (x0, y0)."""
    x -= x0
    y -= y0
    if _on_device(x, y):
        return _profile("gaussian", x, y, a0=float(A0), rho0=float(rho0))
    xh, yh = _host(x), _host(y)
    return _like(A0 * np.exp(- (xh ** 2 + yh ** 2) / rho0 ** 2), x)


def round_hole(x, y, R: float, x0: float = 0, y0: float = 0):
    """profiles.py:74-81: exp(-r^2/R^2) >= exp(-1/2), i.e. a disc of radius R/sqrt(2)."""
    if _on_device(x, y):
        return _profile("round_hole", x, y, r=float(R), x0=float(x0), y0=float(y0))
    xh, yh = _host(x), _host(y)
    d = np.exp(- ((xh - x0) ** 2 + (yh - y0) ** 2) / R ** 2)
    return _like(d >= 1 / np.exp(0.5), x)


def _round_hole_host(xh, yh, R, x0, y0):
    return np.exp(- ((xh - x0) ** 2 + (yh - y0) ** 2) / R ** 2) >= 1 / np.exp(0.5)


def random_round_hole_bin(x, y, R: float, x0: float = 0, y0: float = 0,
                          binning_order: int = 1, rng=None):
    """profiles.py:84-91"""
    if _on_device(x, y):
        _binned_shape_check(x, y, int(binning_order))
        r = _draw("int", len(y), len(x), int(binning_order), rng)
        return _profile("random_hole", x, y, rand=r, binning_order=int(binning_order),
                        r=float(R), x0=float(x0), y0=float(y0))
    xh, yh = _host(x), _host(y)
    r = _rng(rng).randint(0, 2, size=(int(len(yh) // binning_order),
                                       int(len(xh) // binning_order)))
    field = np.kron(r, np.ones((binning_order, binning_order)))
    field[_round_hole_host(xh, yh, R, x0, y0) == 0] = 0
    return _like(field, x)


def random_round_hole(x, y, R: float, x0: float = 0, y0: float = 0,
                      binning_order: int = 1, rng=None):
    """profiles.py:94-101"""
    if _on_device(x, y):
        _binned_shape_check(x, y, int(binning_order))
        r = _draw("uniform", len(y), len(x), int(binning_order), rng)
        return _profile("random_hole", x, y, rand=r, binning_order=int(binning_order),
                        r=float(R), x0=float(x0), y0=float(y0))
    xh, yh = _host(x), _host(y)
    r = _rng(rng).uniform(size=(int(len(yh) // binning_order),
                                 int(len(xh) // binning_order)))
    field = np.kron(r, np.ones((binning_order, binning_order)))
    field[_round_hole_host(xh, yh, R, x0, y0) == 0] = 0
    return _like(field, x)


def random_round_hole_phase(x, y, R: float, x0: float = 0, y0: float = 0,
                            binning_order: int = 1, rng=None):
    """profiles.py:104-113: unit amplitude inside round_hole(R), phase 2*pi*U[0,1)."""
    if _on_device(x, y):
        _binned_shape_check(x, y, int(binning_order))
        r = _draw("uniform", len(y), len(x), int(binning_order), rng)
        return _profile("random_phase", x, y, rand=r, binning_order=int(binning_order),
                        r=float(R), x0=float(x0), y0=float(y0))
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
    if _on_device(x, y):
        return _profile("rect_hole", x, y, dx=float(dx), dy=float(dy), x0=float(x0),
                        y0=float(y0))
    xh, yh = _host(x), _host(y)
    return _like((np.abs(xh - x0) < (dx / 2)) & (np.abs(yh - y0) < (dy / 2)), x)


def square_hole(x, y, d: float, x0: float = 0, y0: float = 0):
    """profiles.py:124-128"""
    return rectangle_hole(x, y, d, d, x0, y0)


def square_slits(x, y, d: float, slits_distance: float, x0: float = 0, y0: float = 0):
    """profiles.py:131-136"""
    if _on_device(x, y):
        return _profile("slits", x, y, dx=float(d), slits=float(slits_distance), x0=float(x0),
                        y0=float(y0))
    dl = slits_distance / 2
    a = square_hole(x, y, d, x0 - dl, y0)
    b = square_hole(x, y, d, x0 + dl, y0)
    return a | b
