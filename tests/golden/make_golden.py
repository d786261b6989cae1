"""Write the golden fixtures in tests/golden/*.npz.

The expected outputs are computed WITHOUT numpy.fft: every 2-D DFT is a product with
explicit DFT matrices W[k, j] = exp(-2*pi*i*((k*j) mod n)/n), and the transfer
function is the textbook exp(-2*pi*i*Kz*z) evaluated directly.  The operation
sequences are those of the reference Beam2D methods (lightprop2d/beam2d.py:
__post_init__ :118-133, coordinate_filter :335-342, spectral_filter :366-373,
propagate :481-490, lens :508-511).  Inputs are seeded.  The oracle and the HIP
path are both checked against these files.

Run:  python tests/golden/make_golden.py      (writes next to this script)
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import profiles as P  # noqa: E402  (input generators only)

nm, um, mm = 1e-7, 1e-4, 1e-1


def dft_matrix(n):
    k = np.arange(n)
    prod = np.outer(k, k) % n                 # exact integer index products
    return np.exp(-2j * np.pi * prod / n)


def bf_fft2(a):
    W = dft_matrix(a.shape[-1])
    return W @ a @ W.T


def bf_ifft2(A):
    n = A.shape[-1]
    Wc = np.conj(dft_matrix(n))
    return (Wc @ A @ Wc.T) / (n * n)


def freqs(n, dL):
    ks = np.where(np.arange(n) < (n + 1) // 2, np.arange(n), np.arange(n) - n)
    return ks / (n * dL)


def grids(area, n, wl):
    dL = area / n
    X = (np.arange(n) - n // 2) * dL
    f = freqs(n, dL)
    k0 = 1.0 / wl
    kz2 = k0 * k0 - f[None, :] ** 2 - f[:, None] ** 2
    Kz = np.where(kz2 > 0, np.sqrt(np.maximum(kz2, 0)), 0.0)
    return dL, X, X.reshape(-1, 1), k0, f, f.reshape(-1, 1), Kz


def H(Kz, z):
    return np.exp(-2j * np.pi * Kz * z)


class BFBeam:
    """Brute-force Beam2D: same op sequence, DFT matrices instead of FFT."""

    def __init__(self, area, n, wl, field):
        self.area, self.n, self.wl = area, n, wl
        self.dL, self.X, self.Y, self.k0, self.Kx, self.Ky, self.Kz = grids(area, n, wl)
        self.field = np.asarray(field, dtype=np.complex128).copy()
        self.spectrum = bf_fft2(self.field)

    def propagate(self, z):
        if z == 0:
            return
        self.spectrum = self.spectrum * H(self.Kz, z)
        self.field = bf_ifft2(self.spectrum)

    def lens(self, f):
        ph = (self.X ** 2 + self.Y ** 2) * self.k0 / (2 * f)
        self.field = self.field * np.exp(1j * ph)
        self.spectrum = bf_fft2(self.field)

    def coordinate_filter(self, mask):
        self.field = self.field * mask
        self.spectrum = bf_fft2(self.field)

    def spectral_filter(self, mask):
        self.spectrum = self.spectrum * mask
        self.field = bf_ifft2(self.spectrum)


def run_case(name, area, n, wl, field, ops, arrays):
    """ops: list of (op, arg) with arg either a float or a key into `arrays`."""
    b = BFBeam(area, n, wl, field)
    for op, arg in ops:
        val = arrays[arg] if isinstance(arg, str) else arg
        getattr(b, op)(val)
    out = dict(field_in=np.asarray(field, dtype=np.complex128),
               field_out=b.field, spectrum_out=b.spectrum)
    for k, v in arrays.items():
        out["arr_" + k] = v
    meta = dict(name=name, area=area, n=n, wl=wl, ops=ops)
    np.savez_compressed(os.path.join(HERE, name + ".npz"),
                        meta=np.array(json.dumps(meta)), **out)


def adjoint_case(name, area, n, wl, z, seed):
    rs = np.random.RandomState(seed)
    u = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    g = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    phi = rs.uniform(0, 2 * np.pi, (n, n))
    _, _, _, _, _, _, Kz = grids(area, n, wl)
    e = np.exp(1j * phi)
    fwd = bf_ifft2(H(Kz, z) * bf_fft2(u * e))
    g1 = bf_ifft2(np.conj(H(Kz, z)) * bf_fft2(g))
    grad_u = np.conj(e) * g1
    grad_phi = np.imag(np.conj(u * e) * g1)
    meta = dict(name=name, area=area, n=n, wl=wl, z=z)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), meta=np.array(json.dumps(meta)),
                        u=u, g=g, phi=phi, fwd=fwd, grad_u=grad_u, grad_phi=grad_phi)


def main():
    # 1. Gaussian beam, one propagate (the config of the reference CPU case, scaled down)
    area, n, wl = 200 * um, 32, 632 * nm
    dL = area / n
    X = (np.arange(n) - n // 2) * dL
    g0 = P.gaussian_beam(X.copy(), X.reshape(-1, 1).copy(), 1.0, 25 * um)
    run_case("gauss_propagate_n32", area, n, wl, g0, [("propagate", 100 * um)], {})

    # 2. Speckle field + NA aperture, 3 steps
    area, n, wl = 400 * um, 64, 632 * nm
    dL = area / n
    X = (np.arange(n) - n // 2) * dL
    sp = P.random_round_hole_phase(X, X.reshape(-1, 1), 100 * um,
                                   rng=np.random.RandomState(0))
    f = freqs(n, dL)
    ap = ((f[None, :] ** 2 + f[:, None] ** 2) <= (0.2 / wl) ** 2).astype(np.float64)
    run_case("speckle_aperture_n64", area, n, wl, sp,
             [("propagate", 50 * um), ("spectral_filter", "aperture")] * 3,
             {"aperture": ap})

    # 3. Evanescent modes present (dL < wl/2): H = 1 where Kz = 0
    area, n, wl = 8 * um, 32, 632 * nm
    rs = np.random.RandomState(1)
    fe = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    run_case("evanescent_n32", area, n, wl, fe, [("propagate", 1 * um), ("propagate", 3 * um)], {})

    # 4. Lens chain: propagate -> lens -> propagate (phase wraps many cycles)
    area, n, wl = 500 * um, 64, 1064 * nm
    dL = area / n
    X = (np.arange(n) - n // 2) * dL
    g1 = P.gaussian_beam(X.copy(), X.reshape(-1, 1).copy(), 1.0, 50 * um)
    run_case("lens_chain_n64", area, n, wl, g1,
             [("propagate", 1 * mm), ("lens", 5 * mm), ("propagate", 5 * mm)], {})

    # 5. Coordinate filter with a complex mask, then propagate; odd-looking z
    area, n, wl = 300 * um, 16, 532 * nm
    rs = np.random.RandomState(2)
    fc = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    mk = np.exp(1j * rs.uniform(0, 2 * np.pi, (n, n))) * (rs.uniform(size=(n, n)) > 0.3)
    run_case("coord_filter_n16", area, n, wl, fc,
             [("coordinate_filter", "mask"), ("propagate", 0.37 * mm), ("propagate", -0.11 * mm)],
             {"mask": mk})

    # 6. Adjoint of propagate with a phase mask (for the gradient path)
    adjoint_case("adjoint_n32", 200 * um, 32, 632 * nm, 0.3 * mm, seed=3)
    adjoint_case("adjoint_evanescent_n16", 6 * um, 16, 632 * nm, 2 * um, seed=4)


if __name__ == "__main__":
    main()
