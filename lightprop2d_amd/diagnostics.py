"""Beam diagnostics on device (lightprop2d/beam2d.py:669-752): intensity-weighted
centroid and second moments from one native two-pass FP64 reduction (lp2d_moments)."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from . import propagation as P

__all__ = ["moments"]


def moments(field: torch.Tensor, dL: float):
    """(Xc, Yc, sigma_x, sigma_y) of |field|^2 with X = (ix - n/2)*dL (rows = y).

    For a (batch, n, n) field, returns a (batch, 4) float64 tensor on the host."""
    lib = P._lib()
    n, batch, dt = P._field_geom(field)
    out = torch.empty((batch, 8), dtype=torch.float64, device=field.device)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_moments(n, batch, dt, P._ptr(field), float(dL), P._ptr(out),
                                 P._stream(field)), "moments")
    r = out[:, 1:5].cpu()
    if field.dim() == 2:
        xc, yc, sx, sy = (float(v) for v in r[0])
        return xc, yc, sx, sy
    return r
