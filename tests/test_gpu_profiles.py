"""Device-side generators (lp2d_profile) against the oracle's restatement of
lightprop2d/profiles.py on Beam2D grids.

Masks (bool) must be identical element by element; float outputs agree to the last
bits of exp/sin/cos (relative 1e-15).  The random generators take the same host draw
(numpy RandomState or the legacy global stream) on both sides.
"""
import numpy as np
import pytest
import torch

from oracle import profiles as OP
import lightprop2d_amd.profiles as PP

pytestmark = pytest.mark.gpu

DEV = "cuda"


def grids(n, dL):
    x = np.arange(-n // 2, n // 2, 1) * dL
    return x, x.reshape((-1, 1))


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


CASES = [(256, 200e-4 / 256), (1024, 400e-4 / 1024), (512, 500e-4 / 512)]


@pytest.mark.parametrize("n,dL", CASES)
@pytest.mark.parametrize("name,args", [
    ("plane_wave", ()),
    ("gaussian_beam", (1.0, 25e-4)),
    ("gaussian_beam", (0.7, 50e-4, 3e-4, -2e-4)),
    ("round_hole", (25e-4,)),
    ("round_hole", (40e-4, 1e-4, -3e-4)),
    ("rectangle_hole", (50e-4, 20e-4)),
    ("rectangle_hole", (30e-4, 60e-4, 2e-4, 1e-4)),
    ("square_hole", (30e-4,)),
    ("square_slits", (10e-4, 30e-4)),
    ("square_slits", (8e-4, 20e-4, 1e-4, -1e-4)),
])
def test_device_generators_match_oracle(n, dL, name, args):
    x, y = grids(n, dL)
    exp = np.asarray(getattr(OP, name)(x.copy(), y.copy(), *args))
    xt, yt = dev(x), dev(y)
    got = getattr(PP, name)(xt, yt, *args)
    assert got.device.type == "cuda" and tuple(got.shape) == exp.shape
    g = got.cpu().numpy()
    if exp.dtype == bool:
        assert g.dtype == np.bool_
        np.testing.assert_array_equal(g, exp)
    else:
        assert g.dtype == exp.dtype
        np.testing.assert_allclose(g, exp, rtol=1e-15, atol=1e-300)
    if name == "gaussian_beam" and len(args) > 2:  # the reference shifts the grid in place
        np.testing.assert_array_equal(xt.cpu().numpy(), x - args[2])
        np.testing.assert_array_equal(yt.cpu().numpy(), y - args[3])


def test_round_hole_boundary_through_grid_points():
    # R/sqrt(2) = 37 dL exactly: points on the circle decide the comparison to the ulp
    n, dL = 256, 1e-4
    x, y = grids(n, dL)
    for k in (17, 37, 64):
        R = np.sqrt(2) * k * dL
        exp = OP.round_hole(x, y, R)
        got = PP.round_hole(dev(x), dev(y), R).cpu().numpy()
        np.testing.assert_array_equal(got, exp)


@pytest.mark.parametrize("n,dL", CASES[:2])
@pytest.mark.parametrize("name,args", [
    ("random_wave", ()), ("random_wave", (4,)), ("random_wave_bin", (2,)),
    ("random_round_hole", (25e-4,)), ("random_round_hole_bin", (40e-4, 0, 0, 2)),
    ("random_round_hole_phase", (25e-4,)), ("random_round_hole_phase", (30e-4, 2e-4, 0, 4)),
])
def test_device_random_generators_match_oracle(n, dL, name, args):
    x, y = grids(n, dL)
    exp = np.asarray(getattr(OP, name)(x, y, *args, rng=np.random.RandomState(7)))
    got = getattr(PP, name)(dev(x), dev(y), *args, rng=np.random.RandomState(7))
    g = got.cpu().numpy()
    assert g.shape == exp.shape and g.dtype == exp.dtype
    if name == "random_round_hole_phase":
        np.testing.assert_allclose(g, exp, rtol=1e-15, atol=1e-15)
        np.testing.assert_array_equal(g == 0, exp == 0)
    else:
        np.testing.assert_array_equal(g, exp)
    # the legacy global stream, as the reference draws it
    np.random.seed(11)
    e2 = np.asarray(getattr(OP, name)(x, y, *args))
    np.random.seed(11)
    g2 = getattr(PP, name)(dev(x), dev(y), *args).cpu().numpy()
    np.testing.assert_allclose(g2, e2, rtol=1e-15, atol=1e-15)


def test_beam2d_generator_runs_on_device():
    from lightprop2d_amd import Beam2D, um, nm
    from lightprop2d_amd import profiles as P
    np.random.seed(0)
    b = Beam2D(400 * um, 1024, 632 * nm, init_field_gen=P.random_round_hole_phase,
               init_gen_args=(25 * um,), complex_bits=64)
    np.random.seed(0)
    x, y = grids(1024, 400e-4 / 1024)
    exp = OP.random_round_hole_phase(x, y, 25e-4).astype(np.complex64)
    assert b.field.device.type == "cuda"
    np.testing.assert_allclose(b.field.cpu().numpy(), exp, rtol=1e-6, atol=1e-7)


def test_binning_order_must_divide_the_grid():
    from lightprop2d_amd._lib import Lp2dError
    x, y = grids(256, 1e-4)
    with pytest.raises(Lp2dError):
        PP.random_round_hole(dev(x), dev(y), 25e-4, binning_order=3)
