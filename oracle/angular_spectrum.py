"""Oracle: numpy complex128 restatement of lightprop2d's Beam2D hot path.

TEST INFRASTRUCTURE ONLY.  This module is the checker that the HIP path is
compared against; it is never the thing measured or shipped.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it.

Every function follows the reference file:line given in its docstring
(reference = vongostev/lightprop2d 1.0.4, ``lightprop2d/beam2d.py``).  The FFT is
``numpy.fft`` (pocketfft), which is what the reference itself calls when pyfftw is
not installed (beam2d.py:31-34, 217, 229).

How each function is pinned (see tests/test_oracle.py):

* ``construct_grids``       closed form + ``numpy.fft.fftfreq``; evanescent count
                             checked against a brute-force count.
* ``fft2`` / ``ifft2``      brute-force DFT matrices built from exact integer index
                             products (tests/golden fixtures), delta/tone KATs,
                             round trip.
* ``transfer_function``     |H| = 1, H = 1 on evanescent modes, H(-z) = conj H(z),
                             H(z1)H(z2) = H(z1+z2), explicit exp(-2*pi*i*Kz*z).
* ``OracleBeam2D.propagate``  analytic Gaussian-beam on-axis intensity, width and
                             Gouy phase (reference tests/gaussbeam_diffraction.py:51-79),
                             round/square hole on-axis formulas
                             (tests/hole_diffraction.py:44-47, 76-100), composition,
                             reversibility, Parseval, golden fixtures from the
                             brute-force DFT.
* ``lens``                  closed form; focal-plane KAT (a lens of effective focal
                             length 2*pi*f focuses a plane wave).
* ``coordinate_filter`` / ``spectral_filter``  golden fixtures (brute-force DFT).
* ``iprofile``              |u|^2.
* ``propagate_field_adjoint``  <A u, g> = <u, A^H g> dot-product test, central
                             finite differences on the phase gradient.
* ``interp2d_linear``       scipy's RectBivariateSpline(kx=1, ky=1, s=0), which runs the
                             same FITPACK fit (regrid_smth) and evaluation (bispev) that
                             interp2d(kind='linear') ran; exact on bilinear functions;
                             samples reproduced at the nodes; edge clamping.
* ``OracleBeam2D.crop``     the window copied unchanged when the new grid lands on the
                             old samples; ``lens_image``: l1 = l2 gives minus the field
                             shifted by one sample (interp2d's sorted queries) times the
                             closed-form chirp, l1 = -l2 the field unchanged.

No function is "parity unpinned".
"""
from __future__ import annotations

from dataclasses import dataclass
from collections.abc import Iterable

import numpy as np

__all__ = [
    "nm", "um", "mm", "cm", "m",
    "Grids", "construct_grids", "fft2", "ifft2", "transfer_function",
    "lens_phase", "circular_aperture", "OracleBeam2D",
    "propagate_field", "propagate_field_adjoint", "propagate_stack",
    "propagate_stack_grad", "interp2d_linear",
]

# Units, centimetre base (beam2d.py:37-41)
nm = 1e-7
um = 1e-4
mm = 1e-1
cm = 1
m = 100


@dataclass
class Grids:
    dL: float
    X: np.ndarray   # (n,)  spatial x, FP64
    Y: np.ndarray   # (n,1) spatial y, FP64
    k0: float       # 1/wl, cycles per cm
    Kx: np.ndarray  # (n,)  fftfreq order
    Ky: np.ndarray  # (n,1)
    Kz: np.ndarray  # (n,n) evanescent -> 0


def construct_grids(area_size: float, npoints: int, wl: float) -> Grids:
    """beam2d.py:290-308 (with _k_grid :189-205).

    dL = area/n; X = arange(-n//2, n//2)*dL; k0 = 1/wl (cycles per cm, not 2*pi/wl);
    Kx = fftfreq(n, dL) in unshifted order; Kz = nan_to_num(sqrt(k0^2-Kx^2-Ky^2)),
    so evanescent modes get Kz = 0.
    """
    dL = area_size / npoints
    X = np.arange(-npoints // 2, npoints // 2, 1) * dL
    Y = X.reshape((-1, 1))
    k0 = 1 / wl
    Kx = np.fft.fftfreq(npoints, d=dL)
    Ky = Kx.reshape((-1, 1))
    with np.errstate(invalid="ignore"):
        Kz = np.nan_to_num(np.sqrt(k0 ** 2 - Kx ** 2 - Ky ** 2))
    return Grids(dL, X, Y, k0, Kx, Ky, Kz)


def fft2(a: np.ndarray) -> np.ndarray:
    """beam2d.py:207-217, numpy branch: unnormalised forward 2-D FFT, last two axes."""
    return np.fft.fft2(a)


def ifft2(a: np.ndarray) -> np.ndarray:
    """beam2d.py:219-229, numpy branch: inverse 2-D FFT scaled by 1/n^2."""
    return np.fft.ifft2(a)


def transfer_function(Kz: np.ndarray, z: float) -> np.ndarray:
    """beam2d.py:485-489: H = cos(p) + i sin(p), p = 2*pi*(trunc(Kz*z) - Kz*z).

    Equal to exp(-2*pi*i*Kz*z) with the integer cycles dropped in FP64.  Where Kz = 0
    (evanescent) H = 1: the mode passes through unchanged.
    """
    _deltak = Kz * z
    phase = 2 * np.pi * (np.trunc(_deltak) - _deltak)
    return np.cos(phase) + 1j * np.sin(phase)


def lens_phase(X: np.ndarray, Y: np.ndarray, k0: float, f: float) -> np.ndarray:
    """beam2d.py:508: phase = (X^2 + Y^2) * k0 / (2 f) (no 2*pi factor, as in the reference)."""
    return (X ** 2 + Y ** 2) * k0 / (2 * f)


def circular_aperture(Kx: np.ndarray, Ky: np.ndarray, radius: float) -> np.ndarray:
    """Spectral pupil |f| <= radius (cycles/cm), f in fftfreq order.

    The on-chip aperture of the HIP path keeps Kx^2 + Ky^2 <= radius^2.  (The
    reference's own ``round_hole(Kx, Ky, R)`` keeps radius R/sqrt(2), see
    oracle.profiles.round_hole.)
    """
    return (Kx ** 2 + Ky ** 2) <= radius ** 2


def _fitpack_linear_taps(t, q):
    """FITPACK fpbisp / fpbspl for degree 1 on knots at the samples t (ascending): the
    argument clamped to [t_0, t_-1], the interval the last l <= len-2 with t_l <= q,
    basis values f*(t_{l+1} - q) and f*(q - t_l), f = 1/(t_{l+1} - t_l)."""
    q = np.clip(q, t[0], t[-1])
    l = np.minimum(np.searchsorted(t, q, side="right") - 1, len(t) - 2)
    f = 1.0 / (t[l + 1] - t[l])
    return l, f * (t[l + 1] - q), f * (q - t[l])


def interp2d_linear(x, y, z, xq, yq):
    """``scipy.interpolate.interp2d(x, y, z, kind='linear')(xq, yq)`` for data on a
    rectangular grid, as it ran before scipy 1.14 removed it (the calls of Beam2D.crop
    and Beam2D.lens_image, beam2d.py:428-429, 434, 527-528, 534).

    interp2d fitted FITPACK's regrid_smth with kx = ky = 1 and s = 0 -- the piecewise
    bilinear interpolant with a breakpoint at every sample -- and evaluated it with
    bisplev at the SORTED queries (mergesort; an image is never flipped), where a query
    outside [x_0, x_-1] takes the edge value.  z[i, j] is the value at (x[j], y[i]);
    the result is (len(yq), len(xq)).  Real and imaginary parts are interpolated
    separately by the reference (with real weights, the same as complex arithmetic),
    in FP64, summed in fpbisp's order (x tap outer, y tap inner).
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    z = np.asarray(z)
    if len(x) < 2 or len(y) < 2:
        raise ValueError("a linear spline needs at least 2 samples per axis")
    jx = np.argsort(x, kind="mergesort")
    jy = np.argsort(y, kind="mergesort")
    x, y, z = x[jx], y[jy], z[np.ix_(jy, jx)]
    xq = np.sort(np.atleast_1d(np.asarray(xq, dtype=np.float64)), kind="mergesort")
    yq = np.sort(np.atleast_1d(np.asarray(yq, dtype=np.float64)), kind="mergesort")
    lx, hx0, hx1 = _fitpack_linear_taps(x, xq)
    ly, hy0, hy1 = _fitpack_linear_taps(y, yq)
    zc = z.astype(np.result_type(z.dtype, np.float64))

    def tap(a, b):
        return zc[np.ix_(ly + b, lx + a)]

    out = (tap(0, 0) * hx0[None, :]) * hy0[:, None]
    out = out + (tap(0, 1) * hx0[None, :]) * hy1[:, None]
    out = out + (tap(1, 0) * hx1[None, :]) * hy0[:, None]
    out = out + (tap(1, 1) * hx1[None, :]) * hy1[:, None]
    return out


class OracleBeam2D:
    """Numpy-path restatement of ``Beam2D`` (beam2d.py:44-133, 231-511, 669-752).

    Field/spectrum state, in-place multiplications and FFT rebinding follow the
    reference line by line; ``use_gpu`` does not exist here (numpy only).
    """

    def __init__(self, area_size, npoints, wl, z=0., init_field=None,
                 init_spectrum=None, init_field_gen=None, init_gen_args=(),
                 complex_bits=128):
        # beam2d.py:93-133
        self.area_size = area_size
        self.npoints = npoints
        self.wl = wl
        self.z = z
        self.complex_bits = complex_bits
        if complex_bits == 64:
            self.complex = np.complex64
        elif complex_bits == 128:
            self.complex = np.complex128
        self._construct_grids()
        if init_field_gen is not None:
            self.field = init_field_gen(self.X, self.Y, *init_gen_args)
        elif init_field is not None:
            self.field = init_field
        else:
            raise ValueError(
                "Init field data is None : "
                "'init_field_gen' must be a function or 'init_field' must be an array.")
        if self.field.dtype != self.complex:
            self.field = self.field.astype(self.complex)
        if init_spectrum is None:
            self.spectrum = fft2(self.field)
        else:
            self.spectrum = init_spectrum

    def _construct_grids(self):
        # beam2d.py:290-308
        g = construct_grids(self.area_size, self.npoints, self.wl)
        self.dL, self.X, self.Y, self.k0 = g.dL, g.X, g.Y, g.k0
        self.Kx, self.Ky, self.Kz = g.Kx, g.Ky, g.Kz

    def _update_obj(self, field, spectrum=None):
        # beam2d.py:283-288
        self.z = 0
        self.field = field
        if spectrum is None:
            self.spectrum = fft2(self.field)
        else:
            self.spectrum = spectrum

    def coordinate_filter(self, f_init=None, f_gen=None, fargs=()):
        # beam2d.py:335-342
        if f_gen is not None:
            assert callable(f_gen)
            assert isinstance(fargs, Iterable)
            self.field *= f_gen(self.X, self.Y, *fargs)
        if f_init is not None:
            assert isinstance(f_init, np.ndarray)
            self.field *= f_init
        self.spectrum = fft2(self.field)

    def spectral_filter(self, f_init=None, f_gen=None, fargs=()):
        # beam2d.py:366-373
        if f_gen is not None:
            assert callable(f_gen)
            assert isinstance(fargs, Iterable)
            self.spectrum *= f_gen(self.Kx, self.Ky, *fargs)
        if f_init is not None:
            assert isinstance(f_init, np.ndarray)
            self.spectrum *= f_init
        self.field = ifft2(self.spectrum)

    def propagate(self, z):
        # beam2d.py:481-490
        if z == 0:
            return
        self.z += z
        self.spectrum *= transfer_function(self.Kz, z)
        self.field = ifft2(self.spectrum)

    def lens(self, f):
        # beam2d.py:508-511
        phase = lens_phase(self.X, self.Y, self.k0, f)
        self.field *= np.cos(phase) + 1j * np.sin(phase)
        self.spectrum = fft2(self.field)

    def expand(self, area_size):
        # beam2d.py:385-404: edge-replicate padding, dL kept
        old_npoints = self.npoints
        self.npoints = int(self.npoints * area_size / self.area_size)
        self.npoints -= self.npoints % 2
        self.area_size = area_size
        d_npoints = (self.npoints - old_npoints) // 2
        self._construct_grids()
        f = self.field
        f = np.hstack((np.repeat(f[:, 0].reshape((-1, 1)), d_npoints, axis=1), f,
                       np.repeat(f[:, -1].reshape((-1, 1)), d_npoints, axis=1)))
        f = np.vstack((np.repeat(f[0].reshape((1, -1)), d_npoints, axis=0), f,
                       np.repeat(f[-1].reshape((1, -1)), d_npoints, axis=0)))
        self.field = f
        self.spectrum = fft2(self.field)

    def crop(self, area_size, npoints=0):
        # beam2d.py:419-435: the centred samples with |X| < area_size/2 (an even count
        # 2n), interpolated onto the new grid (npoints kept when 0).  interp2d returns
        # float64, so the field becomes complex128 here, as in the reference.
        old_X = self.X[np.abs(self.X) < area_size / 2]
        Nc = self.npoints // 2
        n = len(old_X) // 2
        old_X = self.X[Nc - n:Nc + n]
        if npoints:
            self.npoints = npoints
        self.area_size = area_size
        cropped_profile = self.field[Nc - n:Nc + n, Nc - n:Nc + n]
        self._construct_grids()
        self.field = interp2d_linear(old_X, old_X, cropped_profile, self.X, self.X)
        self.spectrum = fft2(self.field)

    def lens_image(self, f, l1, l2):
        # beam2d.py:525-536 (f is unused there too).  The reference's numpy path scales
        # self.X in place through the alias X = self._np(self.X); its cupy path (.get())
        # copies.  This follows the copy, leaving the grids alone.
        scale_factor = - l1 / l2
        X = self.X.copy()
        phase = - (self.X ** 2 + self.Y ** 2) * self.k0 * (l1 + l2) / 2 / l2 ** 2
        Xq = X * scale_factor
        self.field = scale_factor * interp2d_linear(X, X, self.field, Xq, Xq) * \
            (np.cos(phase) + 1j * np.sin(phase))
        self.spectrum = fft2(self.field)

    def coarse(self, mean_order=1):
        # beam2d.py:454-460: block mean
        self.npoints //= mean_order
        self._construct_grids()
        self.field = self.field.reshape(
            (self.npoints, mean_order, self.npoints, mean_order)).mean((1, 3))
        self.spectrum = fft2(self.field)

    # ---------------------------------------------------------- mode bases
    def _expand_basis(self, modes_list):
        # beam2d.py:538-561: flattened Nb x Nb modes zero-padded into the centre of the
        # npoints grid (or just reshaped when Nb == npoints)
        Nb = int(np.sqrt(len(modes_list[0])))
        if Nb != self.npoints:
            dN = (self.npoints - Nb) // 2
            Nt = self.npoints - dN
            out = []
            for m in modes_list:
                me = np.zeros([self.npoints] * 2, dtype=np.complex128)
                me[dN:Nt, dN:Nt] = np.asarray(m).reshape((Nb, Nb))
                out.append(me)
            return np.array(out)
        return np.array([np.asarray(m).reshape((Nb, Nb)) for m in modes_list])

    def _central_field(self, Nb):
        dN = (self.npoints - Nb) // 2
        return self.field[dN:self.npoints - dN, dN:self.npoints - dN].flatten()

    def deconstruct_by_modes(self, modes_list, mode="scalar"):
        # beam2d.py:563-603: coefficients by the (unconjugated) product with the mode
        # matrix, or by least squares of modes^T c = field over the central Nb x Nb
        modes_matrix = np.vstack([np.asarray(m) for m in modes_list])
        Nb = int(np.sqrt(modes_matrix.shape[1]))
        f = self._central_field(Nb)
        if mode == "scalar":
            self.modes_coeffs = modes_matrix @ f
        elif mode == "lstsq":
            self.modes_coeffs = np.linalg.lstsq(modes_matrix.T, f, rcond=-1)[0]
        return self.modes_coeffs

    def fast_deconstruct_by_modes(self, modes_matrix_t, modes_matrix_dot_t):
        # beam2d.py:605-651: lstsq(Gram matrix, modes_matrix_t . field)
        Nb = int(np.sqrt(modes_matrix_t.shape[1]))
        f = self._central_field(Nb)
        self.modes_coeffs = np.linalg.lstsq(modes_matrix_dot_t, modes_matrix_t.dot(f),
                                            rcond=None)[0]
        return self.modes_coeffs

    def construct_by_modes(self, modes_list, modes_coeffs):
        # beam2d.py:653-667: field = sum_i c_i M_i (padded), spectrum = fft2(field)
        self.field = np.einsum('ijk,i->jk', self._expand_basis(modes_list), modes_coeffs)
        self.spectrum = fft2(self.field)

    @property
    def iprofile(self):
        # beam2d.py:717
        return np.abs(self.field) ** 2

    @property
    def phiprofile(self):
        # beam2d.py:734
        return np.angle(self.field)

    @property
    def centroid(self):
        # beam2d.py:680-685
        n = self.npoints // 2
        Y, X = np.mgrid[-n:n, -n:n] * self.dL
        intensity = self.iprofile
        Xc = np.average(X, weights=intensity)
        Yc = np.average(Y, weights=intensity)
        return Xc, Yc, int(Xc / self.dL) + n, int(Yc / self.dL) + n

    @property
    def D4sigma(self):
        # beam2d.py:696-704
        n = self.npoints // 2
        intensity = self.iprofile
        Y, X = np.mgrid[-n:n, -n:n] * self.dL
        Xc, Yc, _, _ = self.centroid

        def sigma(x, xc):
            return np.sqrt(np.average((x - xc) ** 2, weights=intensity)).tolist()

        return 4 * sigma(X, Xc), 4 * sigma(Y, Yc)

    @property
    def centroid_intensity(self):
        # beam2d.py:751-752
        _, _, nxc, nyc = self.centroid
        return self.iprofile[nyc, nxc].tolist()


# ---------------------------------------------------------------------------
# Field-to-field forms used by the batched / differentiable paths.  Each one is the
# composition of the Beam2D operations above, written out so that a test can state
# the reference sequence it stands for.
# ---------------------------------------------------------------------------

def _grids(area_size, n, wl):
    return construct_grids(area_size, n, wl)


def propagate_field(u, area_size, wl, z, phase_mask=None, aperture_radius=None,
                    lens_f=None):
    """u_out = ifft2(H(z) * M * fft2(u * exp(i*phi))), complex128.

    Reference sequence (beam2d.py): ``Beam2D(init_field=u)`` (:122-131) ->
    ``coordinate_filter(f_init=exp(i*phi))`` (:339-342) -> optional ``lens(f)``
    (:508-511) -> ``propagate(z)`` (:481-490) -> optional
    ``spectral_filter(f_init=M)`` (:370-373).  ``fft2`` after ``ifft2`` is the
    identity up to rounding, so the chain is one forward and one inverse FFT.
    """
    u = np.asarray(u, dtype=np.complex128)
    n = u.shape[-1]
    g = _grids(area_size, n, wl)
    if phase_mask is not None:
        u = u * np.exp(1j * np.asarray(phase_mask, dtype=np.float64))
    if lens_f is not None:
        ph = lens_phase(g.X, g.Y, g.k0, lens_f)
        u = u * (np.cos(ph) + 1j * np.sin(ph))
    S = fft2(u)
    if z != 0:
        S = S * transfer_function(g.Kz, z)
    if aperture_radius is not None:
        S = S * circular_aperture(g.Kx, g.Ky, aperture_radius)
    return ifft2(S)


def propagate_field_prepared(u, g, z, dtype=np.complex128, workers=None):
    """One ``Beam2D.propagate(z)`` on a field whose grids `g` (``construct_grids``) were
    built once, in the state precision `dtype` (beam2d.py:111-114): spectrum =
    fft2(field) (numpy >= 2 keeps complex64), H rebuilt from Kz in FP64 on every call
    and multiplied in place, so a complex64 state casts the complex128 product back
    (beam2d.py:484-489), then field = ifft2(spectrum) (:490).  `workers` runs scipy.fft
    on that many threads instead of numpy.fft (the pyfftw ``threads=-1`` stand-in,
    beam2d.py:27-29).  This is the per-field CPU cost of a reference user who reuses one
    Beam2D geometry for many fields."""
    u = np.asarray(u, dtype=dtype)
    if workers is None:
        S = np.fft.fft2(u)
    else:
        import scipy.fft as sfft
        S = sfft.fft2(u, workers=workers)
    if z != 0:
        S *= transfer_function(g.Kz, z)
    if workers is None:
        return np.fft.ifft2(S)
    import scipy.fft as sfft
    return sfft.ifft2(S, workers=workers)


def propagate_field_threaded(u, area_size, wl, z, workers):
    """``propagate_field`` (no mask, no lens) with scipy.fft on `workers` threads: the
    stand-in for the reference's pyfftw branch with ``threads=-1`` (beam2d.py:27-29,
    213-216), which is not installed.  Same grids and transfer function."""
    import scipy.fft as sfft
    u = np.asarray(u, dtype=np.complex128)
    g = _grids(area_size, u.shape[-1], wl)
    S = sfft.fft2(u, workers=workers)
    if z != 0:
        S = S * transfer_function(g.Kz, z)
    return sfft.ifft2(S, workers=workers)


def propagate_field_adjoint(grad_out, u_in, area_size, wl, z, phase_mask=None,
                            aperture_radius=None):
    """Adjoint of ``propagate_field`` for a real loss (PyTorch complex-gradient
    convention: grad = dL/dRe + i dL/dIm).

    A = F^-1 H M F D(phi) with F the unnormalised DFT; A^H = D(-phi) F^-1 conj(H) M F.
    Returns (grad_u, grad_phi) with grad_phi = Im(conj(u*exp(i*phi)) * g1),
    g1 = F^-1 conj(H) M F grad_out.
    """
    g = np.asarray(grad_out, dtype=np.complex128)
    n = g.shape[-1]
    gr = _grids(area_size, n, wl)
    G = fft2(g)
    if z != 0:
        G = G * np.conj(transfer_function(gr.Kz, z))
    if aperture_radius is not None:
        G = G * circular_aperture(gr.Kx, gr.Ky, aperture_radius)
    g1 = ifft2(G)
    u = np.asarray(u_in, dtype=np.complex128)
    if phase_mask is None:
        return g1, None
    e = np.exp(1j * np.asarray(phase_mask, dtype=np.float64))
    grad_u = np.conj(e) * g1
    grad_phi = np.imag(np.conj(u * e) * g1)
    return grad_u, grad_phi


def propagate_stack(u0, screens, area_size, wl, dz, lens_f=None):
    """Multi-plane stack: for each screen k, u <- P(dz)[u * exp(i*phi_k)].

    Reference sequence per plane: ``coordinate_filter(f_init=exp(i*phi_k))`` then
    ``propagate(dz)`` (beam2d.py:339-342, 481-490).  Returns the output field.
    """
    u = np.asarray(u0, dtype=np.complex128)
    for phi in screens:
        u = propagate_field(u, area_size, wl, dz, phase_mask=phi)
    return u


def propagate_stack_grad(u0, screens, area_size, wl, dz, grad_out):
    """Reverse-mode gradient of a real loss through ``propagate_stack`` w.r.t.
    every screen, given grad_out = dL/d(u_out) (PyTorch convention)."""
    planes = [np.asarray(u0, dtype=np.complex128)]
    for phi in screens:
        planes.append(propagate_field(planes[-1], area_size, wl, dz, phase_mask=phi))
    g = np.asarray(grad_out, dtype=np.complex128)
    grads = [None] * len(screens)
    for k in range(len(screens) - 1, -1, -1):
        g, grads[k] = propagate_field_adjoint(g, planes[k], area_size, wl, dz,
                                              phase_mask=screens[k])
    return g, grads
