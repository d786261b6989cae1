"""Drop-in ``Beam2D`` on PyTorch-ROCm device tensors, backed by liblp2d (HIP, gfx950).

Same constructor fields, methods and state semantics as the reference
(lightprop2d/beam2d.py:44-757): ``field`` and ``spectrum`` are carried state, operators
mutate them in place or rebind them to new transform outputs, methods return None.
Differences, all deliberate:

* arrays are ``torch`` tensors on the current HIP device (``xp`` is ``torch``); the
  ``use_gpu`` flag is accepted and ignored (the HIP path always runs);
* ``complex_bits=64`` computes in complex64 kernels, 128 in complex128 kernels;
* a device tensor ``init_field`` (or ``init_spectrum``) of the beam's complex dtype is
  kept as the beam's state itself, as the reference keeps the array it is given, so
  ``coordinate_filter`` / ``lens`` multiply it in place as the reference's
  ``field *=`` does; a numpy array is copied to the device (as the reference's cupy
  path copies host arrays) and another dtype is converted;
* ``field`` is materialised lazily: ``propagate`` and ``spectral_filter`` only multiply
  the carried ``spectrum`` (one elementwise sweep, the transfer function of a
  ``propagate`` is folded into the next spectral operation), and the inverse FFT runs
  when ``field`` (or ``iprofile``, ``centroid``, ...) is next read.  Every value a
  caller can observe is the reference's: reading ``spectrum`` also freezes the field
  it stood for, and assigning either attribute keeps the other as the reference
  would.  Assigned arrays are stored in the beam's complex dtype;
* ``f_gen`` callables receive the device grids (``X``/``Y`` or ``Kx``/``Ky``) as torch
  tensors; one written for numpy arrays (``np.exp``, ``np.abs`` ...) that raises on
  them is called again on host copies of the grids, and its result is uploaded (any
  in-place shift of its grid arguments is copied back, as the reference's numpy path
  would keep it);
* assigning ``Kz`` replaces the transfer function's Kz as in the reference (propagate
  then multiplies by exp(-2*pi*i*frac(Kz*z)) built from the assigned grid in FP64);
* ``npoints`` is a power of two from 2 to 8192 (the device FFT); ``expand``,
  ``coarse`` and ``crop`` check the new grid and raise ``Lp2dError`` before they
  change any state;
* ``crop`` / ``lens_image`` resample with a HIP kernel that computes what scipy's
  removed ``interp2d(kind='linear')`` computed (``propagation.interp_linear``), and the
  field keeps the beam's precision (the reference's becomes complex128 there, since
  interp2d returns float64); ``lens_image`` leaves ``X`` / ``Y`` alone, as the
  reference's cupy path does (its numpy path scales them in place through an alias).
"""
from __future__ import annotations

from collections.abc import Iterable
from dataclasses import dataclass
from typing import Tuple, Union

import numpy as np
import torch

from . import propagation as P
from ._lib import Lp2dError

# CGS units, centimetre base (beam2d.py:37-41)
nm = 1e-7
um = 1e-4
mm = 1e-1
cm = 1
m = 100

ArrayLike = Union[np.ndarray, torch.Tensor]


def _check_grid(npoints: int) -> None:
    """The device FFT runs powers of two from 2 to 8192 points."""
    if npoints < 2 or npoints > 8192 or npoints & (npoints - 1):
        raise Lp2dError(f"npoints = {npoints}: the device FFT needs a power of two "
                        "from 2 to 8192")


@dataclass
class Beam2D:
    """Electromagnetic field propagation with the angular-spectrum method.

    Parameters mirror beam2d.py:79-91: area_size (cm), npoints, wl (cm), z, xp,
    init_field, init_spectrum, init_field_gen, init_gen_args, complex_bits,
    use_gpu, numpy_output.  ``device`` (extension) selects the HIP device.
    """

    area_size: float
    npoints: int
    wl: float
    z: float = 0.
    xp: object = torch
    init_field: object = None
    init_spectrum: object = None
    init_field_gen: object = None
    init_gen_args: tuple = ()
    complex_bits: int = 128

    use_gpu: bool = False
    numpy_output: bool = True
    device: object = None
    evanescent: str = "reference"

    def __post_init__(self):
        # beam2d.py:105-133
        self.xp = torch
        self._device = torch.device(self.device) if self.device is not None else \
            torch.device("cuda", torch.cuda.current_device())
        if self._device.type != "cuda":
            raise Lp2dError("Beam2D runs on a HIP device; got " + str(self._device))
        if self.complex_bits == 64:
            self.complex = torch.complex64
        elif self.complex_bits == 128:
            self.complex = torch.complex128
        self._field = None
        self._spectrum = None
        self._pending_z = 0.0
        self._construct_grids()
        if self.init_field_gen is not None:
            field = self._asxp(self._call_gen(self.init_field_gen, self.X, self.Y,
                                              self.init_gen_args))
        elif self.init_field is not None:
            field = self._asxp(self.init_field)
        else:
            raise ValueError(
                "Init field data is None : " +
                "'init_field_gen' must be a function or 'init_field' must be an array.")
        self._field = self._field_tensor(field)
        if self.init_spectrum is None:
            self._spectrum = self._fft2(self._field)
        else:
            self._spectrum = self._field_tensor(self._asxp(self.init_spectrum))

    # ------------------------------------------------------------- lazy state
    # Invariant: _spectrum is the reference's spectrum once the pending distance
    # _pending_z (a propagate not yet applied) is multiplied in; _field is the
    # reference's field, or None when it is ifft2 of that spectrum.
    def _flush(self, masks=(), field_out=None) -> None:
        """spectrum *= H(pending z) * masks in place (one elementwise sweep, or the
        column pass of an inverse FFT into field_out)."""
        z, self._pending_z = self._pending_z, 0.0
        masks = list(masks)
        if z != 0 and self._Kz_user:
            masks.insert(0, self._user_transfer(z))
            z = 0.0
        if len(masks) > 2:  # the kernels take two multipliers per call
            P.spectral_apply(self._spectrum, masks[:1], z=z, area_size=self.area_size,
                             wl=self.wl, evanescent=self.evanescent)
            masks, z = masks[1:], 0.0
        if z == 0 and not masks:
            if field_out is not None:
                P.ifft2(self._spectrum, out=field_out)
            return
        P.spectral_apply(self._spectrum, masks, z=z, area_size=self.area_size, wl=self.wl,
                         evanescent=self.evanescent, field_out=field_out)

    def _field_now(self) -> torch.Tensor:
        """The current field, computed from the spectrum if it is stale."""
        if self._field is None:
            out = torch.empty_like(self._spectrum)
            self._flush(field_out=out)
            self._field = out
        return self._field

    def _spectrum_now(self) -> torch.Tensor:
        """The current spectrum with any pending propagation applied."""
        if self._pending_z != 0:
            self._flush()
        return self._spectrum

    @property
    def field(self) -> torch.Tensor:
        return self._field_now()

    @field.setter
    def field(self, value):
        # the reference rebinds field and leaves spectrum as it was
        self._spectrum_now()
        self._field = self._field_tensor(self._asxp(value))

    @property
    def spectrum(self) -> torch.Tensor:
        # handed out, the spectrum may be changed in place by the caller; the field of
        # this state is the reference's ifft2 of it *now*, so it is fixed first
        self._field_now()
        return self._spectrum

    @spectrum.setter
    def spectrum(self, value):
        self._field_now()
        self._spectrum = self._field_tensor(self._asxp(value))

    def _call_gen(self, f, a: torch.Tensor, b: torch.Tensor, args):
        """f(a, b, *args) on the device grids, or, when f fails on torch tensors (numpy
        code), on host copies of them, with the result uploaded.  a / b are the grid
        vectors (b a view of a); an in-place shift of them by f is kept in both cases,
        as the reference's grids keep it (profiles.py:69-70)."""
        saved = a.clone()
        try:
            return f(a, b, *args)
        except (TypeError, RuntimeError, AttributeError):
            a.copy_(saved)  # undo a partial in-place shift before the host call
        ah = a.detach().cpu().numpy().copy()
        bh = ah.reshape(tuple(b.shape))
        r = f(ah, bh, *args)
        a.copy_(torch.from_numpy(ah))
        return r

    # ------------------------------------------------------------- array plumbing
    def _np(self, data) -> np.ndarray:
        """beam2d.py:135-152: host numpy copy."""
        if isinstance(data, torch.Tensor):
            return data.detach().cpu().numpy()
        return data

    def _asxp(self, data) -> torch.Tensor:
        """beam2d.py:154-187: device tensor from a numpy array or tensor."""
        if isinstance(data, torch.Tensor):
            return data.to(self._device)
        if isinstance(data, np.ndarray):
            return torch.from_numpy(np.ascontiguousarray(data)).to(self._device)
        raise ValueError("Unknown type of the given array, not numpy or torch array")

    def _k_grid(self, dL: float, npoints: int) -> torch.Tensor:
        """beam2d.py:189-205: fftfreq(npoints, d=dL) (same FP64 values as numpy)."""
        return torch.from_numpy(np.fft.fftfreq(npoints, d=dL)).to(self._device)

    def _fft2(self, data: torch.Tensor) -> torch.Tensor:
        """beam2d.py:207-217"""
        return P.fft2(self._field_tensor(data))

    def _ifft2(self, data: torch.Tensor) -> torch.Tensor:
        """beam2d.py:219-229"""
        return P.ifft2(self._field_tensor(data))

    def _field_tensor(self, t: torch.Tensor) -> torch.Tensor:
        if t.dtype != self.complex:
            t = t.to(self.complex)
        return t.contiguous()

    def _update_obj(self, field: ArrayLike, spectrum: ArrayLike = None):
        """beam2d.py:283-288: replace the state without (or with) an FFT; z = 0."""
        self.z = 0
        self._pending_z = 0.0
        self._field = self._field_tensor(self._asxp(field))
        if spectrum is None:
            self._spectrum = self._fft2(self._field)
        else:
            self._spectrum = self._field_tensor(self._asxp(spectrum))

    def _construct_grids(self):
        """beam2d.py:290-308.  Kz comes from the same kernel code that generates it on
        chip during propagation."""
        self.dL = self.area_size / self.npoints
        self.X = (torch.arange(-self.npoints // 2, self.npoints // 2, 1,
                               dtype=torch.float64) * self.dL).to(self._device)
        self.Y = self.X.reshape((-1, 1))
        self.k0 = 1 / self.wl
        self.Kx = self._k_grid(self.dL, self.npoints)
        self.Ky = self.Kx.reshape((-1, 1))
        self._Kz = None
        self._Kz_user = False

    @property
    def Kz(self) -> torch.Tensor:
        if self._Kz is None:
            self._Kz = P.kz_grid(self.npoints, self.area_size, self.wl, self._device)
        return self._Kz

    @Kz.setter
    def Kz(self, value):
        """An assigned Kz replaces the on-chip transfer function in propagate (the
        reference's propagate reads self.Kz, beam2d.py:485)."""
        self._spectrum_now()
        self._Kz = self._asxp(value).to(torch.float64)
        self._Kz_user = True

    def _user_transfer(self, z: float) -> torch.Tensor:
        """exp(-2*pi*i*frac(Kz*z)) from the assigned Kz, FP64 (beam2d.py:485-489)."""
        d = self._Kz * z
        ph = 2 * np.pi * (torch.trunc(d) - d)
        return torch.complex(torch.cos(ph), torch.sin(ph))

    # ------------------------------------------------------------- operators
    def coordinate_filter(self, f_init: ArrayLike = None, f_gen=None, fargs: tuple = ()) -> None:
        """beam2d.py:310-342: field *= f_gen(X, Y, *fargs) * f_init; spectrum = fft2(field)."""
        masks = []
        if f_gen is not None:
            assert callable(f_gen)
            assert isinstance(fargs, Iterable)
            masks.append(self._call_gen(f_gen, self.X, self.Y, fargs))
        if f_init is not None:
            assert isinstance(f_init, (np.ndarray, torch.Tensor))
            masks.append(f_init)
        field = self._field_now()
        spec = torch.empty_like(field)
        P.spatial_apply(field, masks, spectrum_out=spec)
        self._spectrum = spec

    def spectral_filter(self, f_init: ArrayLike = None, f_gen=None, fargs: tuple = ()) -> None:
        """beam2d.py:344-373: spectrum *= f_gen(Kx, Ky, *fargs) * f_init; field = ifft2.

        The multiply (with a pending propagate's transfer function) runs now; the
        inverse FFT when the field is next read."""
        masks = []
        if f_gen is not None:
            assert callable(f_gen)
            assert isinstance(fargs, Iterable)
            masks.append(self._call_gen(f_gen, self.Kx, self.Ky, fargs))
        if f_init is not None:
            assert isinstance(f_init, (np.ndarray, torch.Tensor))
            masks.append(f_init)
        self._flush(masks)
        self._field = None

    def expand(self, area_size: float):
        """beam2d.py:375-404: enlarge the area at constant dL, replicating the edge rows
        and columns of the field; spectrum = fft2(field)."""
        npoints = int(self.npoints * area_size / self.area_size)
        npoints -= npoints % 2
        _check_grid(npoints)
        field = P.expand(self._field_now(), npoints)
        self.npoints = npoints
        self.area_size = area_size
        self._construct_grids()
        self._set_state(field)

    def crop(self, area_size: float, npoints: int = 0):
        """beam2d.py:406-435: keep the centred samples with |X| < area_size/2 and
        interpolate them bilinearly onto npoints (kept when 0) samples of the new area;
        spectrum = fft2(field)."""
        _check_grid(npoints or self.npoints)
        field = P.crop(self._field_now(), self.area_size, area_size, npoints)
        if npoints:
            self.npoints = npoints
        self.area_size = area_size
        self._construct_grids()
        self._set_state(field)

    def lens_image(self, f: float, l1: float, l2: float):
        """beam2d.py:513-536: the image between conjugate planes l1 before and l2 after
        the lens -- the field resampled at X * (-l1/l2) (sorted, so never flipped),
        times -l1/l2 and the chirp exp(-i k0 (l1 + l2) r^2 / (2 l2^2)); f is unused,
        as in the reference.  spectrum = fft2(field)."""
        self._set_state(P.lens_image(self._field_now(), self.area_size, self.wl, l1, l2))

    def coarse(self, mean_order: int = 1):
        """beam2d.py:437-460: block-average the field over mean_order x mean_order
        (npoints //= mean_order); spectrum = fft2(field)."""
        if mean_order >= 1 and self.npoints % mean_order == 0:
            _check_grid(self.npoints // mean_order)
        field = P.coarse(self._field_now(), mean_order)  # ValueError for a non-divisor
        self.npoints //= mean_order
        self._construct_grids()
        self._set_state(field)

    def _set_state(self, field: torch.Tensor) -> None:
        """field = field; spectrum = fft2(field) (a new grid or a resampled field)."""
        self._pending_z = 0.0
        self._field = field
        self._spectrum = self._fft2(field)

    # ------------------------------------------------------------- mode bases
    def _expand_basis(self, modes_list) -> torch.Tensor:
        """beam2d.py:538-561: flattened Nb x Nb modes zero-padded into the centre of
        the npoints grid, (nmodes, npoints, npoints)."""
        eye = torch.eye(len(modes_list), dtype=torch.complex128, device=self._device)
        return P.modes_construct(modes_list, eye, self.npoints, self.complex, self._device)

    def deconstruct_by_modes(self, modes_list, mode: str = "scalar") -> torch.Tensor:
        """beam2d.py:563-603: 'scalar' = the unconjugated product of the mode matrix with
        the central window of the field (HIP kernel); 'lstsq' = least squares of
        modes^T c = field (torch.linalg on the device)."""
        if mode == "scalar":
            self.modes_coeffs = P.modes_project(self._field_now(), modes_list)
        elif mode == "lstsq":
            m = P.mode_matrix(modes_list, torch.complex128, self._device)
            nb = int(round(m.shape[1] ** 0.5))
            d = (self.npoints - nb) // 2
            f = self._field_now()[d:self.npoints - d, d:self.npoints - d].reshape(-1, 1)
            self.modes_coeffs = torch.linalg.lstsq(m.T, f.to(torch.complex128)).solution[:, 0]
        return self.modes_coeffs

    def fast_deconstruct_by_modes(self, modes_matrix_t, modes_matrix_dot_t) -> torch.Tensor:
        """beam2d.py:605-651: lstsq(Gram matrix, modes_matrix_t . field), the product by
        the projection kernel, the small solve with torch.linalg on the device."""
        b = P.modes_project(self._field_now(), modes_matrix_t)
        g = self._asxp(modes_matrix_dot_t).to(torch.complex128)
        self.modes_coeffs = torch.linalg.lstsq(g, b.reshape(-1, 1)).solution[:, 0]
        return self.modes_coeffs

    def construct_by_modes(self, modes_list, modes_coeffs):
        """beam2d.py:653-667: field = sum_i c_i M_i (zero-padded); spectrum = fft2."""
        self._set_state(P.modes_construct(modes_list, modes_coeffs, self.npoints, self.complex,
                                          self._device))

    def propagate(self, z: float):
        """beam2d.py:462-490: spectrum *= exp(-2*pi*i*frac(Kz*z)); field = ifft2(spectrum).

        The transfer function is generated on chip (FP64 Kz*z reduced to its fractional
        cycle); evanescent modes pass unchanged, as in the reference.  The multiply is
        folded into the next operation on the spectrum (a spectral filter, the next
        propagate, or the inverse FFT when the field is read)."""
        if z == 0:
            return
        self.z += z
        if self._pending_z != 0:
            self._flush()
        self._pending_z = z
        self._field = None

    def lens(self, f: float):
        """beam2d.py:492-511: field *= exp(i*(X^2+Y^2)*k0/(2f)); spectrum = fft2(field)."""
        field = self._field_now()
        spec = torch.empty_like(field)
        P.spatial_apply(field, (), lens_f=f, area_size=self.area_size, wl=self.wl,
                        spectrum_out=spec)
        self._spectrum = spec

    # ------------------------------------------------------------- observables
    @property
    def iprofile(self):
        """beam2d.py:706-721: |field|^2 (host numpy when numpy_output)."""
        profile = P.intensity(self._field_now())
        if self.numpy_output:
            return self._np(profile)
        return profile

    @property
    def phiprofile(self):
        """beam2d.py:723-738: arg(field)."""
        profile = torch.angle(self._field_now())
        if self.numpy_output:
            return self._np(profile)
        return profile

    @property
    def centroid(self) -> Tuple[float, float, int, int]:
        """beam2d.py:669-685: intensity-weighted centroid (Xc, Yc, nxc, nyc)."""
        from .diagnostics import moments
        n = self.npoints // 2
        Xc, Yc, _, _ = moments(self._field_now(), self.dL)
        return Xc, Yc, int(Xc / self.dL) + n, int(Yc / self.dL) + n

    @property
    def D4sigma(self) -> Tuple[float, float]:
        """beam2d.py:687-704: D = 4*sigma along x and y."""
        from .diagnostics import moments
        _, _, sx, sy = moments(self._field_now(), self.dL)
        return 4 * sx, 4 * sy

    @property
    def centroid_intensity(self) -> float:
        """beam2d.py:740-752: |A|^2 at the centroid pixel."""
        _, _, nxc, nyc = self.centroid
        v = self._field_now()[nyc, nxc]
        return float((v.real ** 2 + v.imag ** 2).item())

    def __repr__(self) -> str:
        return (f"Beam {self.npoints:d}x{self.npoints:d} points" +
                f"\n\t{self.area_size:.3g}x{self.area_size:.3g} cm " +
                f"\n\t<wl={self.wl / nm:.3g} nm, z={self.z:.3g} cm>")
