"""Differentiable propagation with the hand-written adjoint (HIP kernels).

``propagate(u, z, ...)`` is a ``torch.autograd.Function`` whose forward is one fused
``lp2d_propagate`` call and whose backward is one ``lp2d_propagate_adjoint`` call:

    forward   u_out = F^-1 [ H(z) A F [ u * exp(i*phase) * lens ] ]
    backward  g1 = F^-1 [ conj(H(z)) A F [ grad_out ] ]
              grad_u     = conj(exp(i*phase) * lens) * g1
              grad_phase = Im( conj(u * exp(i*phase) * lens) * g1 )   (summed over a
                           batch that shares one screen)

with PyTorch's complex-gradient convention (grad = dL/dRe + i dL/dIm).  The reference
has no gradients; the forward is the sequence coordinate_filter(exp(i*phase)) -> lens
-> propagate(z) -> spectral_filter(aperture) of lightprop2d/beam2d.py:339-342,
508-511, 481-490, 370-373.

``propagate_stack`` chains K phase screens with checkpointing of the intermediate
planes: only every s-th plane input is kept in the forward; the backward recomputes
each segment from its checkpoint before running the adjoints in reverse.  Forward and
recomputation run as fused stacks (lp2d_propagate_stack: 2K + 1 sweeps for K planes,
the checkpoints stored by the chained row passes), and the adjoint of a recomputed
segment as one fused stack adjoint (lp2d_propagate_stack_adjoint, 2m + 1 sweeps for
m planes).
"""
from __future__ import annotations

import math
from typing import Optional, Sequence

import torch

from . import propagation as P
from ._lib import Lp2dError

__all__ = ["propagate", "PropagateFn", "propagate_stack", "StackFn", "checkpoint_stride"]


class PropagateFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u, phase, z, area_size, wl, aperture_radius, lens_f, evanescent):
        u = u.contiguous()
        out = P.propagate(u, z, area_size, wl, phase=phase, aperture_radius=aperture_radius,
                          lens_f=lens_f, evanescent=evanescent)
        ctx.save_for_backward(u, phase if phase is not None else None)
        ctx.cfg = (z, area_size, wl, aperture_radius, lens_f, evanescent)
        ctx.has_phase = phase is not None
        return out

    @staticmethod
    def backward(ctx, grad_out):
        u, phase = ctx.saved_tensors
        z, area, wl, ap, lens, ev = ctx.cfg
        need_u, need_phase = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        gu, gphi = P.propagate_adjoint(grad_out.contiguous(), u, z, area, wl,
                                       phase=phase if ctx.has_phase else None,
                                       aperture_radius=ap, lens_f=lens, evanescent=ev,
                                       want_phase_grad=bool(need_phase and ctx.has_phase),
                                       want_field_grad=bool(need_u))
        return (gu if need_u else None, gphi if need_phase else None,
                None, None, None, None, None, None)


def propagate(u: torch.Tensor, z: float, area_size: float, wl: float,
              phase: Optional[torch.Tensor] = None, aperture_radius: float = 0.0,
              lens_f: float = 0.0, evanescent: str = "reference") -> torch.Tensor:
    """Differentiable field propagation (w.r.t. ``u`` and the float32 ``phase``)."""
    return PropagateFn.apply(u, phase, float(z), float(area_size), float(wl),
                             float(aperture_radius or 0.0), float(lens_f or 0.0), evanescent)


def checkpoint_stride(n_planes: int, plane_bytes: int, budget_bytes: Optional[int]):
    """Checkpoint schedule (stride s, recompute mode) for K = n_planes planes.

    The forward keeps every s-th plane input (ceil(K/s) planes).  The backward then
    either recomputes a whole segment at once ("segment": s more planes live, K -
    ceil(K/s) extra propagations) or recomputes each plane input from its checkpoint
    ("plane": one more plane live, s(s-1)/2 extra propagations per segment).  With no
    budget every plane is kept: (1, "segment")."""
    if budget_bytes is None:
        return 1, "segment"
    cap = budget_bytes // max(1, plane_bytes)
    best = None
    for s in range(1, n_planes + 1):
        nseg = math.ceil(n_planes / s)
        for mode, need, work in (("segment", nseg + s, n_planes - nseg),
                                 ("plane", nseg + 1,
                                  sum(m * (m - 1) // 2 for m in
                                      [min(s, n_planes - i * s) for i in range(nseg)]))):
            if need <= cap and (best is None or (work, need) < best[0]):
                best = ((work, need), s, mode)
    if best is None:
        raise Lp2dError(f"memory budget {budget_bytes} B holds fewer than the 2 planes "
                        f"({2 * plane_bytes} B) a checkpointed stack needs")
    return best[1], best[2]


class StackFn(torch.autograd.Function):
    """u_K = P(dz)[ ... P(dz)[ u_0 * exp(i*phi_0) ] ... * exp(i*phi_{K-1}) ]."""

    @staticmethod
    def forward(ctx, u0, screens, dz, area_size, wl, stride, recompute):
        K = screens.shape[0]
        u0 = u0.contiguous()
        screens_c = screens.contiguous()
        ckpts = {0: u0}
        planes = [None] * K
        for k in range(stride, K, stride):
            planes[k] = torch.empty_like(u0)
            ckpts[k] = planes[k]
        # one fused stack call: the chained row passes store the checkpointed inputs
        u = P.propagate_stack(u0, screens_c, dz, area_size, wl, planes=planes)
        ctx.ckpts = ckpts
        ctx.cfg = (dz, area_size, wl, stride, K, recompute)
        ctx.save_for_backward(screens)
        ctx.peak_planes = len(ckpts)
        return u

    @staticmethod
    def backward(ctx, grad_out):
        (screens,) = ctx.saved_tensors
        dz, area, wl, stride, K, recompute = ctx.cfg
        g = grad_out.contiguous()
        gscreens = (torch.zeros(screens.shape, dtype=screens.dtype, device=screens.device)
                    if ctx.needs_input_grad[1] else None)

        screens_c = screens.contiguous()

        def advance(u, k0, k1):  # plane input k1 from plane input k0
            if k1 <= k0:
                return u
            return P.propagate_stack(u, screens_c[k0:k1], dz, area, wl)

        for seg0 in reversed(range(0, K, stride)):
            seg1 = min(K, seg0 + stride)
            ck = ctx.ckpts[seg0]
            planes = None
            if recompute == "segment":
                m = seg1 - seg0
                planes = [ck]
                if m > 1:  # the segment's plane inputs from one stack call
                    pl = [None] + [torch.empty_like(ck) for _ in range(m - 2)]
                    last = P.propagate_stack(ck, screens_c[seg0:seg1 - 1], dz, area, wl,
                                             planes=pl)
                    planes += pl[1:] + [last]
            if planes is not None:
                # the segment's adjoint as one fused stack adjoint (2m + 1 sweeps)
                need_g = seg0 > 0 or ctx.needs_input_grad[0]
                g, _ = P.propagate_stack_adjoint(
                    g, planes, screens_c[seg0:seg1], dz, area, wl,
                    want_screen_grad=gscreens is not None, want_field_grad=need_g,
                    out=g if g is not grad_out else None,
                    grad_screens=gscreens[seg0:seg1] if gscreens is not None else None)
                del planes
                continue
            for k in reversed(range(seg0, seg1)):
                uk = advance(ck, seg0, k)
                need_g = k > 0 or ctx.needs_input_grad[0]
                g, gphi = P.propagate_adjoint(g, uk, dz, area, wl, phase=screens[k],
                                              want_phase_grad=gscreens is not None,
                                              want_field_grad=need_g)
                if gscreens is not None:
                    gscreens[k].copy_(gphi)
                del uk
        ctx.ckpts = None
        return (g if ctx.needs_input_grad[0] else None, gscreens, None, None, None, None, None)


def propagate_stack(u0: torch.Tensor, screens: torch.Tensor, dz: float, area_size: float,
                    wl: float, memory_budget: Optional[int] = None,
                    stride: Optional[int] = None, recompute: str = "segment") -> torch.Tensor:
    """Multi-plane stack with checkpointed intermediate planes.

    u0: (batch, n, n) or (n, n) complex field; screens: (K, n, n) float32 phase
    screens shared by the batch.  ``stride`` + ``recompute`` ("segment" or "plane"),
    or a ``memory_budget`` in bytes for the planes held at once, set the checkpoint
    schedule (see ``checkpoint_stride``); by default every plane input is kept.
    """
    if recompute not in ("segment", "plane"):
        raise ValueError("recompute must be 'segment' or 'plane'")
    if screens.dim() != 3 or screens.dtype != torch.float32:
        raise Lp2dError("screens must be a float32 (K, n, n) tensor")
    plane_bytes = u0.numel() * u0.element_size()
    if stride is None:
        stride, recompute = checkpoint_stride(screens.shape[0], plane_bytes, memory_budget)
    return StackFn.apply(u0, screens, float(dz), float(area_size), float(wl), int(stride),
                         recompute)
