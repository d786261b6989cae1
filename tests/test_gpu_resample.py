"""expand / coarse / crop / lens_image (Beam2D.expand, .coarse, .crop, .lens_image;
beam2d.py:375-460, 513-536) on the device against the oracle.  expand is bit-exact;
coarse agrees to the rounding of an m*m-term mean (relative 1e-7 in complex64, 1e-15
in complex128); the bilinear interpolation of crop / lens_image is computed in FP64
like the oracle's interp2d restatement and rounded once (relative L2 1e-7 in
complex64, 1e-13 in complex128); spectra and propagated fields to the FFT bar
(1e-5 / 1e-10)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle.angular_spectrum import OracleBeam2D

pytestmark = pytest.mark.gpu

DEV = "cuda"


def cplx(rs, shape, dt):
    return (rs.standard_normal(shape) + 1j * rs.standard_normal(shape)).astype(dt)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("n,n_out,batch", [(64, 128, 3), (256, 512, 2), (1024, 2048, 1),
                                           (16, 30, 2), (8, 8, 1)])
def test_expand_matches_edge_padding(dt, n, n_out, batch):
    from lightprop2d_amd import propagation as P
    u = cplx(np.random.RandomState(n), (batch, n, n), dt)
    got = P.expand(torch.from_numpy(u).to(DEV), n_out).cpu().numpy()
    d = (n_out - n) // 2
    exp = np.stack([np.pad(u[i], d, mode="edge") for i in range(batch)])
    np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("n,m,batch", [(64, 2, 3), (256, 4, 2), (2048, 2, 1), (48, 3, 2),
                                       (16, 16, 1), (32, 1, 1)])
def test_coarse_matches_block_mean(dt, n, m, batch):
    from lightprop2d_amd import propagation as P
    u = cplx(np.random.RandomState(n + m), (batch, n, n), dt)
    got = P.coarse(torch.from_numpy(u).to(DEV), m).cpu().numpy()
    exp = u.astype(np.complex128).reshape((batch, n // m, m, n // m, m)).mean((2, 4))
    tol = 1e-6 if dt == np.complex64 else 1e-14
    assert rel_l2(got, exp) < tol


def test_coarse_rejects_a_non_divisor():
    from lightprop2d_amd import propagation as P
    with pytest.raises(ValueError):
        P.coarse(torch.zeros((64, 64), dtype=torch.complex64, device=DEV), 3)


@pytest.mark.parametrize("bits", [64, 128])
def test_beam2d_expand_then_coarse_vs_oracle(bits):
    from lightprop2d_amd import Beam2D, um, nm
    from lightprop2d_amd import profiles as PP
    n, area, wl = 256, 200 * um, 632 * nm
    b = Beam2D(area, n, wl, init_field_gen=PP.gaussian_beam, init_gen_args=(1, 25 * um),
               complex_bits=bits)
    x = np.arange(-n // 2, n // 2, 1) * (area / n)
    o = OracleBeam2D(area, n, wl, init_field=np.exp(-(x[None, :] ** 2 + x[:, None] ** 2)
                                                    / (25 * um) ** 2))
    tol = 1e-5 if bits == 64 else 1e-10
    for op, arg in (("propagate", 300 * um), ("expand", 2 * area), ("propagate", 500 * um),
                    ("coarse", 2), ("propagate", 200 * um)):
        getattr(b, op)(arg)
        getattr(o, op)(arg)
        assert b.npoints == o.npoints and abs(b.area_size - o.area_size) < 1e-15
        assert abs(b.dL - o.dL) < 1e-18
        assert rel_l2(b.field.cpu().numpy(), o.field) < tol, op
        assert rel_l2(b.spectrum.cpu().numpy(), o.spectrum) < tol, op
    assert tuple(b.X.shape) == (n,) and tuple(b.Kz.shape) == (n, n)


@pytest.mark.parametrize("n", [64, 1024])
def test_coarse_vector_and_generic_paths_agree_bitwise(n):
    """The complex64 2 x 2 coarse with 16-byte row loads against the one-element kernel,
    which runs for an input that is not 16-byte aligned: same values, same summation
    order, so bit for bit.  (The 4-element expand is checked bit-exact against numpy
    above; widths that are not a multiple of 4 take the one-element kernel.)"""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(n)
    u = torch.from_numpy(cplx(rs, (2, n, n), np.complex64)).to(DEV)
    flat = torch.empty(2 * n * n + 1, dtype=torch.complex64, device=DEV)
    um = flat[1:].view(2, n, n)  # 8-byte offset: the generic kernel
    um.copy_(u)
    assert torch.equal(P.coarse(u, 2), P.coarse(um, 2))


CROPS = [(256, 3, 0.6, 0), (512, 2, 0.5, 128), (1024, 1, 0.37, 2048), (64, 2, 3 / 64, 16),
         (128, 2, 1.0, 0)]


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("n,batch,frac,npoints", CROPS,
                         ids=[f"{c[0]}-{c[2]:.3g}-{c[3]}" for c in CROPS])
def test_crop_matches_oracle(dt, n, batch, frac, npoints):
    """Zoom (npoints kept), down- and up-sampling, the smallest window (2 samples) and
    a crop to the full area."""
    from lightprop2d_amd import propagation as P
    area, wl = 200e-4, 632e-7
    u = cplx(np.random.RandomState(n + batch), (batch, n, n), dt)
    print(f"crop n={n} batch={batch} frac={frac} npoints={npoints} {np.dtype(dt)}", flush=True)
    got = P.crop(torch.from_numpy(u).to(DEV), area, frac * area, npoints).cpu().numpy()
    tol = 1e-7 if dt == np.complex64 else 1e-13
    for i in range(batch):
        o = OracleBeam2D(area, n, wl, init_field=u[i].copy())
        o.crop(frac * area, npoints)
        assert got.shape[1:] == o.field.shape
        assert rel_l2(got[i], o.field) < tol


LENSES = [(256, 2, 1.0, 2.0), (256, 1, 2.0, 1.0), (128, 2, -1.5, 1.0), (128, 1, 0.5, 0.5),
          (512, 1, 3.0, 4.0)]


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("n,batch,l1,l2", LENSES, ids=[f"{c[0]}-{c[2]}-{c[3]}" for c in LENSES])
def test_lens_image_matches_oracle(dt, n, batch, l1, l2):
    """Magnification 0.5 and 2 (queries beyond the grid clamp), a positive scale (no
    reordering), l1 = l2 (the one-sample shift of the sorted queries)."""
    from lightprop2d_amd import propagation as P
    area, wl = 200e-4, 632e-7
    u = cplx(np.random.RandomState(n + 7), (batch, n, n), dt)
    print(f"lens_image n={n} batch={batch} l1={l1} l2={l2} {np.dtype(dt)}", flush=True)
    got = P.lens_image(torch.from_numpy(u).to(DEV), area, wl, l1, l2).cpu().numpy()
    tol = 1e-7 if dt == np.complex64 else 1e-13
    for i in range(batch):
        o = OracleBeam2D(area, n, wl, init_field=u[i].copy())
        o.lens_image(5.0, l1, l2)
        assert rel_l2(got[i], o.field) < tol


@pytest.mark.parametrize("bits", [64, 128])
def test_beam2d_crop_and_lens_image_chain_vs_oracle(bits):
    from lightprop2d_amd import Beam2D, um, nm
    from lightprop2d_amd import profiles as PP
    n, area, wl = 256, 400 * um, 632 * nm
    b = Beam2D(area, n, wl, init_field_gen=PP.gaussian_beam, init_gen_args=(1, 40 * um),
               complex_bits=bits)
    x = np.arange(-n // 2, n // 2, 1) * (area / n)
    o = OracleBeam2D(area, n, wl, init_field=np.exp(-(x[None, :] ** 2 + x[:, None] ** 2)
                                                    / (40 * um) ** 2))
    tol = 1e-5 if bits == 64 else 1e-10
    steps = (("propagate", (500 * um,)), ("crop", (area / 2,)), ("propagate", (300 * um,)),
             ("crop", (area / 4, 128)), ("lens_image", (1.0, 0.8, 1.2)),
             ("propagate", (200 * um,)))
    for op, args in steps:
        getattr(b, op)(*args)
        getattr(o, op)(*args)
        assert b.npoints == o.npoints and abs(b.dL - o.dL) < 1e-18
        assert b.field.dtype == (torch.complex64 if bits == 64 else torch.complex128)
        assert rel_l2(b.field.cpu().numpy(), o.field) < tol, op
        assert rel_l2(b.spectrum.cpu().numpy(), o.spectrum) < tol, op
    assert torch.equal(b.X.cpu(), torch.from_numpy(o.X))


def test_resampling_to_an_unsupported_grid_leaves_the_beam_unchanged():
    """expand to 1.5x the area (64 -> 96 points) and crop to 100 points have no
    power-of-two FFT on the device, nor does coarse(64) (one point): Lp2dError before
    any state changes; coarse by a non-divisor raises the reference's ValueError, also
    before."""
    from lightprop2d_amd import Beam2D, um, nm
    from lightprop2d_amd._lib import Lp2dError
    b = Beam2D(100 * um, 64, 632 * nm, init_field=np.ones((64, 64)), complex_bits=64)
    f0, s0, x0 = b.field.clone(), b.spectrum.clone(), b.X.clone()
    for op, args, err in (("expand", (150 * um,), Lp2dError), ("coarse", (3,), ValueError),
                          ("coarse", (64,), Lp2dError), ("crop", (50 * um, 100), Lp2dError)):
        with pytest.raises(err):
            getattr(b, op)(*args)
        assert b.npoints == 64 and b.area_size == 100 * um
        assert torch.equal(b.field, f0) and torch.equal(b.spectrum, s0) and torch.equal(b.X, x0)
