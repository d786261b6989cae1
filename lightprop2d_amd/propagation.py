"""Tensor-level operations on PyTorch-ROCm device tensors, each one a call into liblp2d.

Fields are complex64 / complex128 tensors of shape (n, n) or (batch, n, n), rows = y,
as in the reference (lightprop2d/beam2d.py:299-308).  Every function runs on the
caller's current HIP stream; none of them has a CPU path.
"""
from __future__ import annotations

import ctypes as C
from typing import Iterable, Optional, Sequence

import numpy as np
import torch

from . import _lib as L
from ._lib import Lp2dError

__all__ = ["fft2", "ifft2", "kz_grid", "spatial_apply", "spectral_apply", "propagate",
           "propagate_adjoint", "propagate_adjoint_noalloc", "propagate_stack",
           "propagate_stack_adjoint", "Plan", "phase_grad", "intensity", "sum_planes", "UnevenBatch", "optics",
           "expand", "coarse", "mode_matrix", "modes_project", "modes_construct",
           "Lp2dError"]

_DT = {torch.complex64: L.LP2D_C64, torch.complex128: L.LP2D_C128}
_MASK_DT = {torch.bool: L.LP2D_MASK_U8, torch.uint8: L.LP2D_MASK_U8,
            torch.float32: L.LP2D_MASK_F32, torch.float64: L.LP2D_MASK_F64,
            torch.complex64: L.LP2D_MASK_C64, torch.complex128: L.LP2D_MASK_C128}


_tuned = False


def _lib():
    """Load liblp2d and, on first use, install the tuned kernel variants and
    initialise the library on every visible device (outside any graph capture: the
    initialisation allocates, so a library first used under capture would fail)."""
    global _tuned
    lib = L.lib()
    if not _tuned:
        _tuned = True
        from . import autotune
        try:
            autotune.apply_defaults(autotune.arch())
        except L.Lp2dError:
            pass
        if torch.cuda.is_available() and not torch.cuda.is_current_stream_capturing():
            for d in range(torch.cuda.device_count()):
                lib.lp2d_init(d)  # a failure resurfaces with its message at the first call
    return lib


def _small_path(n: int, dtype: str = "c64") -> bool:
    """True when size n runs every operation in one whole-field-in-LDS launch."""
    return L.lib().lp2d_variant_count(n, L.LP2D_C64 if dtype == "c64" else L.LP2D_C128, 2) >= 1


def _stream(t: torch.Tensor):
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _field_geom(t: torch.Tensor, name="field"):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise Lp2dError(f"{name} must live on a HIP device (got {t.device})")
    if t.dtype not in _DT:
        raise Lp2dError(f"{name} must be complex64 or complex128 (got {t.dtype})")
    if t.dim() not in (2, 3) or t.shape[-1] != t.shape[-2]:
        raise Lp2dError(f"{name} must have shape (n, n) or (batch, n, n), got {tuple(t.shape)}")
    if not t.is_contiguous():
        raise Lp2dError(f"{name} must be contiguous")
    n = t.shape[-1]
    batch = t.shape[0] if t.dim() == 3 else 1
    return n, batch, _DT[t.dtype]


def _check_like(t: torch.Tensor, ref: torch.Tensor, name: str) -> None:
    """A caller-supplied buffer must be exactly ``ref``'s shape, dtype and device, and
    contiguous: the kernels read or write batch * n * n elements of ref's dtype
    through it."""
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.shape != ref.shape or t.dtype != ref.dtype or t.device != ref.device:
        raise Lp2dError(f"{name} must be a {ref.dtype} tensor of shape {tuple(ref.shape)} "
                        f"on {ref.device}; got {t.dtype} {tuple(t.shape)} on {t.device}")
    if not t.is_contiguous():
        raise Lp2dError(f"{name} must be contiguous")


def optics(area_size: float, wl: float, z: float = 0.0, aperture_radius: float = 0.0,
           lens_f: float = 0.0, evanescent: str = "reference") -> L.lp2d_optics:
    if evanescent not in L.EVANESCENT:
        raise ValueError(f"evanescent must be one of {sorted(L.EVANESCENT)}")
    return L.lp2d_optics(float(area_size), float(wl), float(z),
                         float(aperture_radius or 0.0), float(lens_f or 0.0),
                         L.EVANESCENT[evanescent], 0)


def _as_device_tensor(m, device):
    if isinstance(m, torch.Tensor):
        return m.to(device)
    if isinstance(m, np.ndarray):
        return torch.from_numpy(np.ascontiguousarray(m)).to(device)
    if np.isscalar(m):
        return torch.tensor(m, device=device)
    raise TypeError(f"unsupported multiplier type {type(m)}")


def _masks(masks: Sequence, n: int, batch: int, device):
    """Build lp2d_mask descriptors (broadcast strides, no copies when possible).
    Returns (array, keepalive list)."""
    keep = []
    arr = (L.lp2d_mask * max(1, len(masks)))()
    for i, m in enumerate(masks):
        t = _as_device_tensor(m, device)
        if t.dtype not in _MASK_DT:
            t = t.to(torch.float64)
        if t.dim() > 3:
            raise Lp2dError(f"mask {i} has {t.dim()} dims; at most (batch, n, n)")
        try:
            e = torch.broadcast_to(t, (batch, n, n))
        except RuntimeError as exc:
            raise Lp2dError(f"mask {i} of shape {tuple(t.shape)} does not broadcast to "
                            f"({batch}, {n}, {n})") from exc
        keep.append(t)
        sb, sy, sx = e.stride()
        arr[i] = L.lp2d_mask(e.data_ptr(), _MASK_DT[t.dtype], 0, sb, sy, sx)
    return arr, keep


def kz_grid(n: int, area_size: float, wl: float, device=None) -> torch.Tensor:
    """Kz = nan_to_num(sqrt(k0^2 - Kx^2 - Ky^2)), FP64 (n, n) (beam2d.py:304-308)."""
    lib = _lib()
    dev = torch.device(device) if device is not None else torch.device("cuda")
    out = torch.empty((n, n), dtype=torch.float64, device=dev)
    o = optics(area_size, wl)
    with torch.cuda.device(out.device):
        L.check(lib.lp2d_kz_grid(n, C.byref(o), _ptr(out), _stream(out)), "kz_grid")
    return out


def fft2(field: torch.Tensor, out: Optional[torch.Tensor] = None, inverse: bool = False):
    """2-D FFT over the last two axes, numpy normalisation (beam2d.py:207-229)."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    if out is None:
        out = torch.empty_like(field)
    else:
        _check_like(out, field, "out")
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_fft2(n, batch, dt, _ptr(field), _ptr(out), int(bool(inverse)),
                              _stream(field)), "fft2")
    return out


def ifft2(spectrum: torch.Tensor, out: Optional[torch.Tensor] = None):
    return fft2(spectrum, out, inverse=True)


def spatial_apply(field: torch.Tensor, masks: Sequence = (), lens_f: float = 0.0,
                  area_size: float = 0.0, wl: float = 0.0, phase: Optional[torch.Tensor] = None,
                  spectrum_out: Optional[torch.Tensor] = None):
    """field *= masks * lens * exp(i*phase) in place; spectrum_out = fft2(field)."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    if len(masks) > 2:
        raise Lp2dError("at most two masks per call")
    marr, keep = _masks(masks, n, batch, field.device)
    o = optics(area_size, wl, lens_f=lens_f) if lens_f else None
    ph, psb = _phase(phase, n, batch, field.device)
    if spectrum_out is not None:
        _check_like(spectrum_out, field, "spectrum_out")
        if spectrum_out.data_ptr() == field.data_ptr():
            raise Lp2dError("spectrum_out must not alias field")
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_spatial_apply(n, batch, dt, _ptr(field), marr, len(masks),
                                       C.byref(o) if o is not None else None, _ptr(ph), psb,
                                       _ptr(spectrum_out), _stream(field)), "spatial_apply")
    del keep
    return spectrum_out


def spectral_apply(spectrum: torch.Tensor, masks: Sequence = (), z: float = 0.0,
                   area_size: float = 0.0, wl: float = 0.0, aperture_radius: float = 0.0,
                   evanescent: str = "reference", field_out: Optional[torch.Tensor] = None):
    """spectrum *= H(z) * aperture * masks in place; field_out = ifft2(spectrum)."""
    lib = _lib()
    n, batch, dt = _field_geom(spectrum, "spectrum")
    if len(masks) > 2:
        raise Lp2dError("at most two masks per call")
    marr, keep = _masks(masks, n, batch, spectrum.device)
    o = optics(area_size, wl, z=z, aperture_radius=aperture_radius,
               evanescent=evanescent) if (z or aperture_radius) else None
    if field_out is not None:
        _check_like(field_out, spectrum, "field_out")
        if field_out.data_ptr() == spectrum.data_ptr():
            raise Lp2dError("field_out must not alias spectrum")
    with torch.cuda.device(spectrum.device):
        L.check(lib.lp2d_spectral_apply(n, batch, dt, _ptr(spectrum), marr, len(masks),
                                        C.byref(o) if o is not None else None,
                                        _ptr(field_out), _stream(spectrum)), "spectral_apply")
    del keep
    return field_out


def _phase(phase, n, batch, device):
    if phase is None:
        return None, 0
    if not isinstance(phase, torch.Tensor):
        phase = torch.as_tensor(np.asarray(phase), device=device)
    if phase.dtype != torch.float32 or phase.device != device or not phase.is_contiguous():
        raise Lp2dError("phase must be a contiguous float32 tensor on the field's device")
    if phase.shape == (n, n):
        return phase, 0
    if phase.shape == (batch, n, n):
        return phase, n * n
    raise Lp2dError(f"phase must have shape ({n}, {n}) or ({batch}, {n}, {n})")


def propagate(field: torch.Tensor, z: float, area_size: float, wl: float,
              phase: Optional[torch.Tensor] = None, aperture_radius: float = 0.0,
              lens_f: float = 0.0, evanescent: str = "reference",
              out: Optional[torch.Tensor] = None, spectrum_out: Optional[torch.Tensor] = None):
    """Field-to-field angular-spectrum propagation (one fused call):

        out = ifft2(H(z) * aperture * fft2(field * exp(i*phase) * lens(lens_f)))

    H(z) = exp(-2*pi*i*frac(Kz*z)) with Kz*z in FP64 (beam2d.py:485-489), evanescent
    modes pass unchanged unless `evanescent` is "zero" or "decay".  `out` may be
    `field` (in place).  Not differentiable; see lightprop2d_amd.autograd.
    """
    lib = _lib()
    n, batch, dt = _field_geom(field)
    if out is None:
        out = torch.empty_like(field)
    else:
        _check_like(out, field, "out")
    if spectrum_out is not None:
        _check_like(spectrum_out, field, "spectrum_out")
        if spectrum_out.data_ptr() in (field.data_ptr(), out.data_ptr()):
            raise Lp2dError("spectrum_out must not alias the fields")
    ph, psb = _phase(phase, n, batch, field.device)
    o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
               evanescent=evanescent)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_propagate(n, batch, dt, _ptr(field), _ptr(out), _ptr(ph), psb,
                                   C.byref(o), _ptr(spectrum_out), _stream(field)), "propagate")
    return out


def propagate_stack(field: torch.Tensor, screens: torch.Tensor, z: float, area_size: float,
                    wl: float, aperture_radius: float = 0.0, lens_f: float = 0.0,
                    evanescent: str = "reference", out: Optional[torch.Tensor] = None,
                    planes: Optional[Sequence[Optional[torch.Tensor]]] = None):
    """Multi-plane stack u_{k+1} = propagate(u_k, z, phase=screens[k]) for the K float32
    screens (K, n, n) shared by the batch; returns u_K.  planes[k] (1 <= k < K), when
    given, receives u_k (checkpoints).  Complex64 sizes with a block-tiled intermediate
    run 2K + 1 sweeps: the inverse row pass of a plane, the next screen and the next
    forward row pass are one sweep."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    if not isinstance(screens, torch.Tensor) or screens.dtype != torch.float32 or \
            screens.dim() != 3 or screens.shape[1:] != (n, n) or \
            screens.device != field.device or not screens.is_contiguous():
        raise Lp2dError(f"screens must be a contiguous float32 (K, {n}, {n}) tensor on the "
                        "field's device")
    K = screens.shape[0]
    if K < 1:
        raise Lp2dError("at least one screen")
    if out is None:
        out = torch.empty_like(field)
    else:
        _check_like(out, field, "out")
    parr = None
    if planes is not None:
        planes = list(planes)
        if len(planes) != K:
            raise Lp2dError(f"planes must hold {K} entries (None where not stored)")
        parr = (C.c_void_p * K)()
        for k, pl in enumerate(planes):
            if pl is None or k == 0:
                continue
            _check_like(pl, field, f"planes[{k}]")
            if pl.data_ptr() in (field.data_ptr(), out.data_ptr()):
                raise Lp2dError(f"planes[{k}] must not alias the fields")
            parr[k] = pl.data_ptr()
    o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
               evanescent=evanescent)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_propagate_stack(n, batch, dt, _ptr(field), _ptr(out), _ptr(screens), K,
                                         C.byref(o), parr, _stream(field)), "propagate_stack")
    return out


def propagate_adjoint(grad_out: torch.Tensor, field_in: Optional[torch.Tensor], z: float,
                      area_size: float, wl: float, phase: Optional[torch.Tensor] = None,
                      aperture_radius: float = 0.0, lens_f: float = 0.0,
                      evanescent: str = "reference", want_phase_grad: bool = True,
                      want_field_grad: bool = True, out: Optional[torch.Tensor] = None,
                      work: Optional[torch.Tensor] = None, use_workspace: bool = True):
    """Adjoint of `propagate` (vector-Jacobian product, PyTorch complex convention).

    Returns (grad_field, grad_phase).  grad_phase has the phase's shape: per field for a
    (batch, n, n) phase, summed over the batch (inside the kernels) for a shared (n, n)
    screen; None when not requested.  grad_field is None when want_field_grad is False
    (its buffer is then only the passes' workspace).

    For complex64 fields that take the three-sweep path, the column pass writes into a
    workspace the size of grad_out (`work`, allocated here when not given), so the
    intermediate plane can stay block-tiled; use_workspace=False keeps every pass in
    `out` (less memory, row-major intermediate)."""
    lib = _lib()
    n, batch, dt = _field_geom(grad_out, "grad_out")
    if out is None:
        out = torch.empty_like(grad_out)
    else:
        _check_like(out, grad_out, "out")
    if field_in is not None:
        _check_like(field_in, grad_out, "field_in")
        if field_in.data_ptr() == out.data_ptr():
            raise Lp2dError("out must not alias field_in")
    if work is not None:
        _check_like(work, grad_out, "work")
        if work.data_ptr() in (out.data_ptr(), grad_out.data_ptr(),
                               field_in.data_ptr() if field_in is not None else -1):
            raise Lp2dError("work must not alias grad_out, field_in or out")
    elif use_workspace and grad_out.dtype == torch.complex64 and not _small_path(n, "c64"):
        work = torch.empty_like(grad_out)
    ph, psb = _phase(phase, n, batch, grad_out.device)
    if (ph is not None or lens_f) and field_in is None:
        raise Lp2dError("field_in is required with a phase screen or lens")
    gph = None
    if ph is not None and want_phase_grad:
        gph = torch.empty((batch, n, n), dtype=torch.float32, device=grad_out.device)
    o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
               evanescent=evanescent)
    flags = 0 if want_field_grad else L.LP2D_ADJ_NO_GRAD_IN
    with torch.cuda.device(grad_out.device):
        L.check(lib.lp2d_propagate_adjoint_ws(n, batch, dt, _ptr(grad_out), _ptr(field_in),
                                              _ptr(ph), psb, C.byref(o), _ptr(out), _ptr(gph),
                                              flags, _ptr(work), _stream(grad_out)),
                "propagate_adjoint")
    if gph is not None and psb == 0:
        gph = gph[0]
    return (out if want_field_grad else None), gph


def propagate_stack_adjoint(grad_out: torch.Tensor, planes: Sequence[torch.Tensor],
                            screens: torch.Tensor, z: float, area_size: float, wl: float,
                            aperture_radius: float = 0.0, lens_f: float = 0.0,
                            evanescent: str = "reference", want_screen_grad: bool = True,
                            want_field_grad: bool = True, out: Optional[torch.Tensor] = None,
                            grad_screens: Optional[torch.Tensor] = None):
    """Adjoint of `propagate_stack`: returns (grad_field, grad_screens).

    planes[k] = u_k, the input of plane k (planes[0] = the stack's input field), as
    `propagate_stack(..., planes=...)` stores them.  grad_screens: (K, n, n) float32,
    the batch-summed screen gradients (None when want_screen_grad is False).
    grad_field is None when want_field_grad is False.  out may be grad_out (in place).
    Complex64 sizes with a block-tiled intermediate run 2K + 1 sweeps: the inverse row
    pass of plane k, the fold and gradient of screen k and the forward row pass of
    plane k-1's adjoint are one sweep."""
    lib = _lib()
    n, batch, dt = _field_geom(grad_out, "grad_out")
    if not isinstance(screens, torch.Tensor) or screens.dtype != torch.float32 or \
            screens.dim() != 3 or screens.shape[1:] != (n, n) or \
            screens.device != grad_out.device or not screens.is_contiguous():
        raise Lp2dError(f"screens must be a contiguous float32 (K, {n}, {n}) tensor on the "
                        "field's device")
    K = screens.shape[0]
    if K < 1:
        raise Lp2dError("at least one screen")
    planes = list(planes)
    if len(planes) != K:
        raise Lp2dError(f"planes must hold the {K} plane inputs")
    if out is None:
        out = torch.empty_like(grad_out)
    else:
        _check_like(out, grad_out, "out")
    parr = (C.c_void_p * K)()
    for k, pl in enumerate(planes):
        _check_like(pl, grad_out, f"planes[{k}]")
        if pl.data_ptr() == out.data_ptr():
            raise Lp2dError(f"out must not alias planes[{k}]")
        parr[k] = pl.data_ptr()
    gs = None
    if want_screen_grad:
        if grad_screens is None:
            gs = torch.empty((K, n, n), dtype=torch.float32, device=grad_out.device)
        else:
            if grad_screens.dtype != torch.float32 or grad_screens.shape != (K, n, n) or \
                    grad_screens.device != grad_out.device or not grad_screens.is_contiguous():
                raise Lp2dError(f"grad_screens must be a contiguous float32 ({K}, {n}, {n}) "
                                "tensor on the field's device")
            gs = grad_screens
    need = int(lib.lp2d_propagate_stack_adjoint_ws_size(n, batch, dt))
    if need < 0:
        raise Lp2dError(f"invalid stack adjoint geometry n={n} batch={batch}")
    work = None
    if gs is not None and need > 0:
        work = torch.empty((need + 3) // 4, dtype=torch.float32, device=grad_out.device)
    o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
               evanescent=evanescent)
    flags = 0 if want_field_grad else L.LP2D_ADJ_NO_GRAD_IN
    with torch.cuda.device(grad_out.device):
        L.check(lib.lp2d_propagate_stack_adjoint(
            n, batch, dt, _ptr(grad_out), parr, _ptr(screens), K, C.byref(o), _ptr(out),
            _ptr(gs), flags, _ptr(work), need if work is not None else 0, _stream(grad_out)),
            "propagate_stack_adjoint")
    return (out if want_field_grad else None), gs


class Plan:
    """A plan for fields of one geometry: (batch, n, n) of one complex dtype on one
    device (lp2d_plan_*).  Creating it initialises the device and copies the kernel
    schedule selected for n at that moment (later lp2d_set_variant / autotuner
    changes do not affect it); its calls allocate nothing inside the library, and the
    adjoint takes a caller-owned workspace (``workspace(op)`` allocates one of
    ``workspace_size(op)`` bytes).  With a shared screen the plan's adjoint sums the
    phase gradient over the batch into one (n, n) plane."""

    def __init__(self, n: int, batch: int = 1, dtype=torch.complex64, device=None):
        lib = _lib()
        if dtype not in _DT:
            raise Lp2dError(f"dtype must be complex64 or complex128 (got {dtype})")
        self.n, self.batch, self.dtype = int(n), int(batch), dtype
        self.device = (torch.device(device) if device is not None
                       else torch.device("cuda", torch.cuda.current_device()))
        self._h = None
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(lib.lp2d_plan_create(self.n, self.batch, _DT[dtype], C.byref(h)),
                    "plan_create")
        self._h = h

    def _handle(self):
        if self._h is None:
            raise Lp2dError("plan is closed")
        return self._h

    def close(self):
        if self._h is not None and L._lib is not None:
            L._lib.lp2d_plan_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown
            pass

    def workspace_size(self, op: int) -> int:
        b = int(_lib().lp2d_plan_workspace_size(self._handle(), int(op)))
        if b < 0:
            raise Lp2dError(f"unknown plan operation {op}")
        return b

    def workspace(self, op: int) -> Optional[torch.Tensor]:
        b = self.workspace_size(op)
        return torch.empty(b, dtype=torch.uint8, device=self.device) if b else None

    def _field(self, t: torch.Tensor, name: str = "field"):
        n, batch, _ = _field_geom(t, name)
        if n != self.n or batch != self.batch or t.dtype != self.dtype or \
                t.device != self.device:
            raise Lp2dError(f"{name} must be {self.batch} x {self.n}^2 {self.dtype} on "
                            f"{self.device}; got {tuple(t.shape)} {t.dtype} on {t.device}")

    def fft2(self, field: torch.Tensor, out: Optional[torch.Tensor] = None,
             inverse: bool = False) -> torch.Tensor:
        self._field(field)
        out = torch.empty_like(field) if out is None else out
        _check_like(out, field, "out")
        with torch.cuda.device(self.device):
            L.check(_lib().lp2d_plan_fft2(self._handle(), _ptr(field), _ptr(out), int(inverse),
                                          _stream(field)), "plan_fft2")
        return out

    def propagate(self, field: torch.Tensor, z: float, area_size: float, wl: float,
                  phase: Optional[torch.Tensor] = None, aperture_radius: float = 0.0,
                  lens_f: float = 0.0, evanescent: str = "reference",
                  out: Optional[torch.Tensor] = None,
                  spectrum_out: Optional[torch.Tensor] = None) -> torch.Tensor:
        """`propagate` with the plan's schedule."""
        self._field(field)
        out = torch.empty_like(field) if out is None else out
        _check_like(out, field, "out")
        if spectrum_out is not None:
            _check_like(spectrum_out, field, "spectrum_out")
        ph, psb = _phase(phase, self.n, self.batch, field.device)
        o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
                   evanescent=evanescent)
        with torch.cuda.device(self.device):
            L.check(_lib().lp2d_plan_propagate(self._handle(), _ptr(field), _ptr(out), _ptr(ph),
                                               psb, C.byref(o), _ptr(spectrum_out),
                                               _stream(field)), "plan_propagate")
        return out

    def propagate_adjoint(self, grad_out: torch.Tensor, field_in: Optional[torch.Tensor],
                          z: float, area_size: float, wl: float,
                          phase: Optional[torch.Tensor] = None, aperture_radius: float = 0.0,
                          lens_f: float = 0.0, evanescent: str = "reference",
                          want_phase_grad: bool = True, want_field_grad: bool = True,
                          out: Optional[torch.Tensor] = None,
                          work: Optional[torch.Tensor] = None):
        """`propagate_adjoint` with the plan's schedule and a caller workspace (a
        device tensor of at least workspace_size(LP2D_OP_ADJOINT) bytes; allocated
        here when not given).  Returns (grad_field, grad_phase)."""
        self._field(grad_out, "grad_out")
        out = torch.empty_like(grad_out) if out is None else out
        _check_like(out, grad_out, "out")
        if field_in is not None:
            _check_like(field_in, grad_out, "field_in")
        ph, psb = _phase(phase, self.n, self.batch, grad_out.device)
        gph = None
        if ph is not None and want_phase_grad:
            shape = (self.n, self.n) if psb == 0 else (self.batch, self.n, self.n)
            gph = torch.empty(shape, dtype=torch.float32, device=self.device)
        need = self.workspace_size(L.LP2D_OP_ADJOINT)
        if work is None and need:
            work = torch.empty(need, dtype=torch.uint8, device=self.device)
        if work is not None and (work.device != self.device or not work.is_contiguous()):
            raise Lp2dError("work must be a contiguous tensor on the plan's device")
        wb = work.numel() * work.element_size() if work is not None else 0
        o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
                   evanescent=evanescent)
        flags = 0 if want_field_grad else L.LP2D_ADJ_NO_GRAD_IN
        with torch.cuda.device(self.device):
            L.check(_lib().lp2d_plan_propagate_adjoint(
                self._handle(), _ptr(grad_out), _ptr(field_in), _ptr(ph), psb, C.byref(o),
                _ptr(out), _ptr(gph), flags, _ptr(work), wb, _stream(grad_out)),
                "plan_propagate_adjoint")
        return (out if want_field_grad else None), gph


def propagate_adjoint_noalloc(grad_out: torch.Tensor, field_in: Optional[torch.Tensor],
                              z: float, area_size: float, wl: float, out: torch.Tensor,
                              phase: Optional[torch.Tensor] = None,
                              grad_phase: Optional[torch.Tensor] = None,
                              aperture_radius: float = 0.0, lens_f: float = 0.0,
                              evanescent: str = "reference", want_field_grad: bool = True):
    """`propagate_adjoint` without any device allocation (lp2d_propagate_adjoint_noalloc),
    for code under graph capture: the caller passes `out` (grad_out's shape) and, for a
    phase gradient, `grad_phase` of the phase's shape ((n, n) for a screen shared by the
    batch, summed over it in the library's own scratch array).  Calls on one device
    share that array, so they must not overlap in time."""
    lib = _lib()
    n, batch, dt = _field_geom(grad_out, "grad_out")
    _check_like(out, grad_out, "out")
    if field_in is not None:
        _check_like(field_in, grad_out, "field_in")
        if field_in.data_ptr() == out.data_ptr():
            raise Lp2dError("out must not alias field_in")
    ph, psb = _phase(phase, n, batch, grad_out.device)
    if (ph is not None or lens_f) and field_in is None:
        raise Lp2dError("field_in is required with a phase screen or lens")
    if grad_phase is not None:
        if ph is None:
            raise Lp2dError("grad_phase requires phase")
        if grad_phase.shape != ph.shape or grad_phase.dtype != torch.float32 or \
                grad_phase.device != grad_out.device or not grad_phase.is_contiguous():
            raise Lp2dError(f"grad_phase must be a contiguous float32 tensor of shape "
                            f"{tuple(ph.shape)} on {grad_out.device}")
    o = optics(area_size, wl, z=z, aperture_radius=aperture_radius, lens_f=lens_f,
               evanescent=evanescent)
    flags = 0 if want_field_grad else L.LP2D_ADJ_NO_GRAD_IN
    with torch.cuda.device(grad_out.device):
        L.check(lib.lp2d_propagate_adjoint_noalloc(n, batch, dt, _ptr(grad_out), _ptr(field_in),
                                                   _ptr(ph), psb, C.byref(o), _ptr(out),
                                                   _ptr(grad_phase), flags, _stream(grad_out)),
                "propagate_adjoint_noalloc")
    return out, grad_phase


def phase_grad(field: torch.Tensor, grad: torch.Tensor, phase: Optional[torch.Tensor] = None,
               sum_batch: bool = True, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Im(conj(field * exp(i*phase)) * grad) per field, or summed over the batch
    (sum_batch, one (n, n) plane): the phase-mask gradient as a separate pass
    (lp2d_phase_grad), for adjoints composed from fft2 / spectral_apply."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    _check_like(grad, field, "grad")
    ph, psb = _phase(phase, n, batch, field.device)
    shape = (n, n) if sum_batch or field.dim() == 2 else (batch, n, n)
    if out is None:
        out = torch.empty(shape, dtype=torch.float32, device=field.device)
    elif out.shape != shape or out.dtype != torch.float32 or out.device != field.device or \
            not out.is_contiguous():
        raise Lp2dError(f"out must be a contiguous float32 tensor of shape {shape}")
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_phase_grad(n, batch, dt, _ptr(field), _ptr(grad), _ptr(ph), psb,
                                    _ptr(out), int(bool(sum_batch)), _stream(field)),
                "phase_grad")
    return out


def sum_planes(planes: torch.Tensor, out: torch.Tensor):
    """out = planes.sum(0) for a float32 (batch, ...) tensor, in a HIP kernel."""
    lib = _lib()
    b = planes.shape[0]
    cnt = planes[0].numel()
    L.check(lib.lp2d_sum_planes(cnt, b, _ptr(planes), _ptr(out), _stream(planes)), "sum_planes")
    return out


def intensity(field: torch.Tensor, out: Optional[torch.Tensor] = None):
    """|field|^2 (beam2d.py:717): float32 for complex64, float64 for complex128."""
    lib = _lib()
    if field.dtype not in _DT or field.device.type != "cuda" or not field.is_contiguous():
        raise Lp2dError("intensity needs a contiguous complex tensor on a HIP device")
    rdt = torch.float32 if field.dtype == torch.complex64 else torch.float64
    if out is None:
        out = torch.empty(field.shape, dtype=rdt, device=field.device)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_intensity(field.numel(), _DT[field.dtype], _ptr(field), _ptr(out),
                                   _stream(field)), "intensity")
    return out


def expand(field: torch.Tensor, n_out: int) -> torch.Tensor:
    """Edge-replicating enlargement (n, n) -> (n_out, n_out), n_out - n even
    (Beam2D.expand, beam2d.py:389-401); batched over leading fields."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    n_out = int(n_out)
    out = torch.empty(field.shape[:-2] + (n_out, n_out), dtype=field.dtype, device=field.device)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_expand(n, n_out, batch, dt, _ptr(field), _ptr(out), _stream(field)),
                "expand")
    return out


def coarse(field: torch.Tensor, mean_order: int) -> torch.Tensor:
    """Block mean over mean_order x mean_order (Beam2D.coarse, beam2d.py:456-459)."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    m = int(mean_order)
    if m < 1 or n % m:
        raise ValueError(f"cannot reshape a {n} x {n} field into blocks of {m} x {m}")
    out = torch.empty(field.shape[:-2] + (n // m, n // m), dtype=field.dtype,
                      device=field.device)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_coarse(n, m, batch, dt, _ptr(field), _ptr(out), _stream(field)),
                "coarse")
    return out


def interp_linear(field: torch.Tensor, n_out: int, desc: "L.lp2d_interp_desc") -> torch.Tensor:
    """Bilinear resampling (..., n, n) -> (..., n_out, n_out): scipy's
    interp2d(kind='linear') on the regular grid ``desc`` describes, with its output
    scale and chirp (the interp2d calls of beam2d.py:428-434, 527-535).  Computed in
    FP64 and rounded once to the field's dtype."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    n_out = int(n_out)
    out = torch.empty(field.shape[:-2] + (n_out, n_out), dtype=field.dtype, device=field.device)
    with torch.cuda.device(field.device):
        L.check(lib.lp2d_interp_linear(n, n_out, batch, dt, _ptr(field), _ptr(out),
                                       C.byref(desc), _stream(field)), "interp_linear")
    return out


def crop_desc(n: int, area_size: float, area_new: float, npoints: int = 0):
    """Window and grids of Beam2D.crop (beam2d.py:419-434): the centred samples with
    |X| < area_new / 2 of the n-point grid of side area_size (an even count 2h around
    n // 2), taken at the new grid X' = arange(-m // 2, m // 2) * area_new / m
    (m = npoints, or n when 0).  Returns (desc, m)."""
    n = int(n)
    m = int(npoints) if npoints else n
    dL = area_size / n
    X = np.arange(-n // 2, n // 2, 1) * dL            # beam2d.py:300-301
    h = len(X[np.abs(X) < area_new / 2]) // 2
    if h < 1:
        raise ValueError(f"crop to {area_new} keeps {2 * h} samples per axis; interp2d "
                         "needs at least 2")
    if m < 1:
        raise ValueError("npoints must be positive")
    Nc = n // 2
    d = L.lp2d_interp_desc(win_off=Nc - h, win_len=2 * h, src_i0=Nc - h + (-n // 2),
                           dst_i0=-m // 2, dst_sign=1, grid_i0=0, src_dx=dL,
                           dst_dx=area_new / m, dst_scale=1.0, scale=1.0, phase_coef=0.0,
                           grid_dl=0.0)
    return d, m


def crop(field: torch.Tensor, area_size: float, area_new: float, npoints: int = 0):
    """Beam2D.crop's resampling (beam2d.py:406-435) of (..., n, n) fields of side
    area_size onto npoints (n when 0) samples of side area_new."""
    d, m = crop_desc(field.shape[-1], area_size, area_new, npoints)
    return interp_linear(field, m, d)


def lens_image_desc(n: int, area_size: float, wl: float, l1: float, l2: float):
    """Grids of Beam2D.lens_image (beam2d.py:525-535): the field on X = arange(-n // 2,
    n // 2) * dL taken at sorted(X * s), s = -l1 / l2, times s and the chirp
    exp(-i k0 (l1 + l2) / (2 l2^2) (X^2 + Y^2)), k0 = 1 / wl."""
    n = int(n)
    dL = area_size / n
    k0 = 1 / wl
    s = - l1 / l2
    i0 = -n // 2
    # interp2d sorts the queries X * s: for s < 0 that is X read backwards
    sign, q0 = (1, i0) if s >= 0 else (-1, n - 1 + i0)
    return L.lp2d_interp_desc(win_off=0, win_len=n, src_i0=i0, dst_i0=q0, dst_sign=sign,
                              grid_i0=i0, src_dx=dL, dst_dx=dL, dst_scale=s, scale=s,
                              phase_coef=k0 * (l1 + l2) / 2 / l2 ** 2, grid_dl=dL)


def lens_image(field: torch.Tensor, area_size: float, wl: float, l1: float, l2: float):
    """Beam2D.lens_image's imaging between conjugate planes (beam2d.py:513-536) of
    (..., n, n) fields of side area_size."""
    n = field.shape[-1]
    return interp_linear(field, n, lens_image_desc(n, area_size, wl, l1, l2))


def mode_matrix(modes, dtype, device) -> torch.Tensor:
    """(nmodes, nb*nb) contiguous device tensor of `dtype` from a list of flattened
    modes (pyMMF layout, beam2d.py:538-561), a 2-D array or a tensor."""
    if isinstance(modes, torch.Tensor):
        m = modes
    elif isinstance(modes, np.ndarray):
        m = torch.from_numpy(np.ascontiguousarray(modes))
    else:
        m = torch.from_numpy(np.ascontiguousarray(np.vstack([np.asarray(
            x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x).reshape(1, -1)
            for x in modes])))
    if m.dim() != 2:
        m = m.reshape(m.shape[0], -1)
    nb = int(round(m.shape[1] ** 0.5))
    if nb * nb != m.shape[1]:
        raise Lp2dError(f"modes of {m.shape[1]} points are not square")
    return m.to(device=device, dtype=dtype).contiguous()


# Batches from this size on project on the matrix cores (lp2d_modes_project_ws).
MFMA_MIN_BATCH = 8


def modes_project(field: torch.Tensor, modes: torch.Tensor,
                  matrix_cores: Optional[bool] = None) -> torch.Tensor:
    """coeffs[b, k] = sum_p modes[k, p] * field[b, window(p)] (unconjugated, as the
    reference's ``modes_matrix @ flatten_field``, beam2d.py:597); complex128 result,
    shape (nmodes,) for one field or (batch, nmodes).

    Batches of MFMA_MIN_BATCH fields or more run as a complex GEMM on the FP64 matrix
    cores with a workspace allocated here; matrix_cores=False / True forces the vector
    kernel / the matrix-core one."""
    lib = _lib()
    n, batch, dt = _field_geom(field)
    modes = mode_matrix(modes, field.dtype, field.device)
    nm = modes.shape[0]
    nb = int(round(modes.shape[1] ** 0.5))
    out = torch.empty((batch, nm), dtype=torch.complex128, device=field.device)
    if matrix_cores is None:
        matrix_cores = batch >= MFMA_MIN_BATCH
    with torch.cuda.device(field.device):
        if matrix_cores:
            wb = int(lib.lp2d_modes_project_ws_size(n, nb, nm, batch))
            work = torch.empty(max(wb, 16), dtype=torch.uint8, device=field.device)
            L.check(lib.lp2d_modes_project_ws(n, nb, nm, batch, dt, _ptr(modes), _ptr(field),
                                              _ptr(out), _ptr(work), wb, _stream(field)),
                    "modes_project")
        else:
            L.check(lib.lp2d_modes_project(n, nb, nm, batch, dt, _ptr(modes), _ptr(field),
                                           _ptr(out), _stream(field)), "modes_project")
    return out[0] if field.dim() == 2 else out


def modes_construct(modes: torch.Tensor, coeffs, n: int, dtype=torch.complex64,
                    device=None) -> torch.Tensor:
    """field = sum_k coeffs[k] * mode_k zero-padded into an n x n grid
    (beam2d.py:653-667); coeffs (nmodes,) or (batch, nmodes)."""
    lib = _lib()
    dev = device if device is not None else (modes.device if isinstance(modes, torch.Tensor)
                                            else torch.device("cuda", torch.cuda.current_device()))
    modes = mode_matrix(modes, dtype, dev)
    nm = modes.shape[0]
    nb = int(round(modes.shape[1] ** 0.5))
    c = coeffs if isinstance(coeffs, torch.Tensor) else torch.from_numpy(np.asarray(coeffs))
    c = c.to(device=dev, dtype=torch.complex128).contiguous()
    single = c.dim() == 1
    c = c.reshape(-1, nm)
    batch = c.shape[0]
    out = torch.empty((batch, n, n), dtype=dtype, device=dev)
    with torch.cuda.device(dev):
        L.check(lib.lp2d_modes_construct(n, nb, nm, batch, _DT[dtype], _ptr(modes), _ptr(c),
                                         _ptr(out), _stream(out)), "modes_construct")
    return out[0] if single else out


class UnevenBatch:
    """Many independent fields of different sizes (powers of two up to 8192), distances
    and optics, propagated per size by the tuned kernels reading per-field pointer and
    optics tables (one launch for sizes that fit in LDS, three sweeps otherwise).

    Fields are propagated src -> dst (dst may be src).  `run()` allocates nothing and
    can be captured in a CUDA/HIP graph (`capture_graph()`).
    """

    def __init__(self, srcs: Sequence[torch.Tensor], dsts: Sequence[torch.Tensor],
                 zs: Iterable[float], area_size, wl, aperture_radius=0.0):
        lib = _lib()
        srcs = list(srcs)
        dsts = list(dsts)
        zs = list(zs)
        m = len(srcs)
        if not (len(dsts) == m == len(zs)) or m == 0:
            raise Lp2dError("srcs, dsts and zs must have the same non-zero length")
        areas = _per_field(area_size, m)
        wls = _per_field(wl, m)
        aps = _per_field(aperture_radius, m)
        descs = (L.lp2d_field_desc * m)()
        for i, (s, d) in enumerate(zip(srcs, dsts)):
            n, b, dt = _field_geom(s, f"srcs[{i}]")
            if b != 1 or s.dim() != 2:
                raise Lp2dError("uneven batches take (n, n) fields")
            if i == 0:
                dtype0 = dt
            elif dt != dtype0:
                raise Lp2dError("all fields of an uneven batch share one dtype")
            if d.shape != s.shape or d.dtype != s.dtype or not d.is_contiguous():
                raise Lp2dError(f"dsts[{i}] must match srcs[{i}]")
            descs[i] = L.lp2d_field_desc(n, 0, s.data_ptr(), d.data_ptr(), float(areas[i]),
                                         float(wls[i]), float(zs[i]), float(aps[i]))
        check_disjoint_fields([(s.data_ptr(), s.numel() * s.element_size()) for s in srcs],
                              [(d.data_ptr(), d.numel() * d.element_size()) for d in dsts])
        self._keep = (srcs, dsts)
        self.device = srcs[0].device
        self._h = C.c_void_p()
        with torch.cuda.device(self.device):
            L.check(lib.lp2d_batch_create(descs, m, dtype0, C.byref(self._h)), "batch_create")
        self.count = m
        self.elements = sum(s.numel() for s in srcs)

    def run(self):
        with torch.cuda.device(self.device):
            L.check(L.lib().lp2d_batch_run(self._h, C.c_void_p(
                torch.cuda.current_stream(self.device).cuda_stream)), "batch_run")

    def capture_graph(self):
        """Capture run() into a torch.cuda.CUDAGraph (HIP graph) and return it.  The
        graph's kernels read this batch's device tables and fields, so the graph holds a
        reference to the batch (``graph.lp2d_batch``): replaying it never outlives them."""
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self.run()
        torch.cuda.current_stream(self.device).wait_stream(s)
        g.lp2d_batch = self
        return g

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and L._lib is not None:
            L._lib.lp2d_batch_destroy(h)
            self._h = None


def check_disjoint_fields(srcs, dsts) -> None:
    """Fields of one uneven batch run concurrently in one launch per sweep, so a
    destination may only share memory with its own source (in place, the same
    address).  srcs / dsts: (address, bytes) per field.  Raises Lp2dError for a
    destination that overlaps another field's source or destination, or its own
    source at a different address."""
    import bisect
    m = len(dsts)
    for i in range(m):
        (s0, sb), (d0, db) = srcs[i], dsts[i]
        if d0 != s0 and d0 < s0 + sb and s0 < d0 + db:
            raise Lp2dError(f"dsts[{i}] overlaps srcs[{i}] without being it (in place "
                            "needs the same tensor)")
    order = sorted(range(m), key=lambda i: dsts[i][0])
    for a, b in zip(order, order[1:]):
        if dsts[b][0] < dsts[a][0] + dsts[a][1]:
            raise Lp2dError(f"dsts[{a}] and dsts[{b}] overlap")
    # sources may be shared (read-only); merge them into disjoint runs with their owners
    runs = []  # [start, end, owners]
    for i in sorted(range(m), key=lambda i: srcs[i][0]):
        s0, sb = srcs[i]
        if runs and s0 < runs[-1][1]:
            runs[-1][1] = max(runs[-1][1], s0 + sb)
            runs[-1][2].add(i)
        else:
            runs.append([s0, s0 + sb, {i}])
    starts = [r[0] for r in runs]
    for i in range(m):
        d0, db = dsts[i]
        k = bisect.bisect_left(starts, d0 + db)  # runs starting before the end of dst i
        j = k - 1
        while j >= 0 and runs[j][1] > d0:
            if runs[j][2] - {i}:
                other = min(runs[j][2] - {i})
                raise Lp2dError(f"dsts[{i}] overlaps srcs[{other}] of another field")
            j -= 1


def _per_field(v, m):
    if np.isscalar(v):
        return [float(v)] * m
    v = list(v)
    if len(v) != m:
        raise Lp2dError("per-field parameter list has the wrong length")
    return v
