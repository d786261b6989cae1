"""Pins the CPU oracle (oracle/) before anything is compared against it.

Each check is chosen so that a plausible mistake in the oracle fails it: a wrong FFT
axis or normalisation fails the brute-force DFT fixtures; a wrong sign of the phase
fails the Gouy-phase and H(-z) = conj H(z) checks; a dropped 2*pi fails the analytic
Gaussian width; the evanescent rule is checked mode by mode.
"""
import json
import os

import numpy as np
import pytest
from scipy.special import fresnel

from conftest import rel_l2
import oracle
from oracle import profiles as P
from oracle.angular_spectrum import (OracleBeam2D, construct_grids, transfer_function,
                                     propagate_field, propagate_field_adjoint, fft2, ifft2,
                                     nm, um, mm)


# ----------------------------------------------------------------- grids / H / FFT
def test_units():
    # beam2d.py:37-41
    assert (oracle.nm, oracle.um, oracle.mm, oracle.cm, oracle.m) == (1e-7, 1e-4, 1e-1, 1, 100)


@pytest.mark.parametrize("n", [8, 9, 32])
def test_grids_closed_form(n):
    area, wl = 200 * um, 632 * nm
    g = construct_grids(area, n, wl)
    dL = area / n
    assert g.dL == dL
    assert g.X.shape == (n,) and g.Y.shape == (n, 1)
    assert g.X[0] == (-n // 2) * dL and np.all(np.diff(g.X) > 0)
    ks = [i if i < (n + 1) // 2 else i - n for i in range(n)]
    np.testing.assert_allclose(g.Kx, np.array(ks) / (n * dL), rtol=1e-15)
    assert g.k0 == 1 / wl


def test_kz_evanescent_count():
    area, n, wl = 8 * um, 32, 632 * nm
    g = construct_grids(area, n, wl)
    brute = 0
    for iy in range(n):
        for ix in range(n):
            fx = g.Kx[ix]
            fy = g.Kx[iy]
            if fx * fx + fy * fy > g.k0 ** 2:
                brute += 1
                assert g.Kz[iy, ix] == 0.0
            else:
                assert g.Kz[iy, ix] == pytest.approx(np.sqrt(g.k0 ** 2 - fx * fx - fy * fy), rel=1e-14)
    assert (g.Kz == 0).sum() == brute > 0


def test_transfer_function_properties():
    area, n, wl = 8 * um, 32, 632 * nm
    g = construct_grids(area, n, wl)
    for z in [1 * um, 0.37 * mm, 5 * mm]:
        h = transfer_function(g.Kz, z)
        np.testing.assert_allclose(np.abs(h), 1, atol=1e-14)
        assert np.all(h[g.Kz == 0] == 1)
        np.testing.assert_allclose(transfer_function(g.Kz, -z), np.conj(h), atol=1e-11)
        np.testing.assert_allclose(h, np.exp(-2j * np.pi * g.Kz * z), atol=1e-9)
    h12 = transfer_function(g.Kz, 1 * mm) * transfer_function(g.Kz, 2 * mm)
    np.testing.assert_allclose(h12, transfer_function(g.Kz, 3 * mm), atol=1e-10)


def _dft(n):
    k = np.arange(n)
    return np.exp(-2j * np.pi * (np.outer(k, k) % n) / n)


@pytest.mark.parametrize("n", [4, 8, 16, 32])
def test_fft2_brute_force(n):
    rs = np.random.RandomState(n)
    a = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    W = _dft(n)
    assert rel_l2(fft2(a), W @ a @ W.T) < 1e-13
    assert rel_l2(ifft2(a), np.conj(W) @ a @ np.conj(W).T / n ** 2) < 1e-13
    assert rel_l2(ifft2(fft2(a)), a) < 1e-14


def test_fft2_known_answers():
    n = 16
    d = np.zeros((n, n), complex)
    d[0, 0] = 1
    np.testing.assert_allclose(fft2(d), np.ones((n, n)))           # delta -> constant
    x = np.arange(n)
    tone = np.exp(2j * np.pi * (3 * x[None, :] + 5 * x[:, None]) / n)
    F = fft2(tone)
    assert abs(F[5, 3] - n * n) < 1e-9                               # tone -> delta
    F[5, 3] = 0
    assert np.abs(F).max() < 1e-9


# ----------------------------------------------------------------- golden fixtures
GOLDEN_CASES = ["gauss_propagate_n32", "speckle_aperture_n64", "evanescent_n32",
                "lens_chain_n64", "coord_filter_n16"]


def load_case(golden_dir, name):
    z = np.load(os.path.join(golden_dir, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    arrays = {k[4:]: z[k] for k in z.files if k.startswith("arr_")}
    return meta, z, arrays


def run_oracle_case(meta, field_in, arrays, complex_bits=128):
    b = OracleBeam2D(meta["area"], meta["n"], meta["wl"], init_field=field_in.copy(),
                     complex_bits=complex_bits)
    for op, arg in meta["ops"]:
        val = arrays[arg] if isinstance(arg, str) else arg
        if op == "coordinate_filter":
            b.coordinate_filter(f_init=val)
        elif op == "spectral_filter":
            b.spectral_filter(f_init=val)
        else:
            getattr(b, op)(val)
    return b


@pytest.mark.parametrize("name", GOLDEN_CASES)
def test_oracle_matches_golden(golden_dir, name):
    meta, z, arrays = load_case(golden_dir, name)
    b = run_oracle_case(meta, z["field_in"], arrays)
    assert rel_l2(b.field, z["field_out"]) < 1e-10
    assert rel_l2(b.spectrum, z["spectrum_out"]) < 1e-10


@pytest.mark.parametrize("name", ["adjoint_n32", "adjoint_evanescent_n16"])
def test_oracle_adjoint_matches_golden(golden_dir, name):
    z = np.load(os.path.join(golden_dir, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    fwd = propagate_field(z["u"], meta["area"], meta["wl"], meta["z"], phase_mask=z["phi"])
    assert rel_l2(fwd, z["fwd"]) < 1e-10
    gu, gp = propagate_field_adjoint(z["g"], z["u"], meta["area"], meta["wl"], meta["z"],
                                     phase_mask=z["phi"])
    assert rel_l2(gu, z["grad_u"]) < 1e-10
    assert rel_l2(gp, z["grad_phi"]) < 1e-10


def test_adjoint_dot_product_and_fd():
    n, area, wl, zz = 16, 100 * um, 632 * nm, 0.2 * mm
    rs = np.random.RandomState(7)
    u = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    g = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    Au = propagate_field(u, area, wl, zz, aperture_radius=0.5 / (2 * area / n))
    AHg, _ = propagate_field_adjoint(g, u, area, wl, zz, aperture_radius=0.5 / (2 * area / n))
    assert abs(np.vdot(g, Au) - np.vdot(AHg, u)) < 1e-10 * abs(np.vdot(g, Au))
    # central finite differences of L = sum |A(u e^{i phi})|^2 * w on a few pixels
    phi = rs.uniform(0, 2 * np.pi, (n, n))
    w = rs.uniform(size=(n, n))

    def loss(ph):
        v = propagate_field(u, area, wl, zz, phase_mask=ph)
        return float(np.sum(w * np.abs(v) ** 2))

    v = propagate_field(u, area, wl, zz, phase_mask=phi)
    _, gphi = propagate_field_adjoint(2 * w * v, u, area, wl, zz, phase_mask=phi)
    h = 1e-6
    for (iy, ix) in [(0, 0), (3, 7), (15, 2), (8, 8)]:
        e = np.zeros((n, n))
        e[iy, ix] = h
        fd = (loss(phi + e) - loss(phi - e)) / (2 * h)
        assert fd == pytest.approx(gphi[iy, ix], rel=1e-5, abs=1e-8)


def test_stack_gradient_central_differences():
    """oracle.propagate_stack_grad against central finite differences of a real loss
    through oracle.propagate_stack: every screen of a 4-plane stack (pixels on the
    pupil edge and in the corners included), and the input-field gradient in both its
    real and imaginary directions (PyTorch convention grad = dL/dRe + i dL/dIm).  A
    dropped conjugate, a sign or a screen order swapped fails it."""
    import oracle
    n, area, wl, dz = 16, 120 * um, 800 * nm, 0.15 * mm
    rs = np.random.RandomState(12)
    u0 = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    screens = rs.uniform(0, 2 * np.pi, (4, n, n))
    w = rs.uniform(size=(n, n))

    def loss(u, scr):
        v = oracle.propagate_stack(u, scr, area, wl, dz)
        return float(np.sum(w * np.abs(v) ** 2))

    v = oracle.propagate_stack(u0, screens, area, wl, dz)
    gu, gs = oracle.propagate_stack_grad(u0, screens, area, wl, dz, 2 * w * v)
    h = 1e-6
    for k in range(4):
        for (iy, ix) in [(0, 0), (5, 11), (15, 15), (8, 3)]:
            e = np.zeros_like(screens)
            e[k, iy, ix] = h
            fd = (loss(u0, screens + e) - loss(u0, screens - e)) / (2 * h)
            assert fd == pytest.approx(gs[k][iy, ix], rel=1e-5, abs=1e-8), (k, iy, ix)
    for (iy, ix) in [(2, 9), (14, 0)]:
        e = np.zeros((n, n), dtype=complex)
        e[iy, ix] = h
        fd_re = (loss(u0 + e, screens) - loss(u0 - e, screens)) / (2 * h)
        fd_im = (loss(u0 + 1j * e, screens) - loss(u0 - 1j * e, screens)) / (2 * h)
        assert fd_re == pytest.approx(gu[iy, ix].real, rel=1e-5, abs=1e-8)
        assert fd_im == pytest.approx(gu[iy, ix].imag, rel=1e-5, abs=1e-8)


# ----------------------------------------------------------------- analytic physics
def test_gaussian_beam_analytic():
    """Reference tests/gaussbeam_diffraction.py:51-79 formulas: on-axis intensity
    1/(1+(z/zR)^2), width w/w0 = sqrt(1+(wl z/(pi R^2))^2) from D4sigma/2, and Gouy
    phase arctan(wl z/(pi R^2)) after removing 2*pi*(trunc(k0 z) - k0 z)."""
    npoints, area, wl, R = 256, 2 * mm, 0.532 * um, 0.1 * mm
    b = OracleBeam2D(area, npoints, wl, init_field_gen=P.gaussian_beam, init_gen_args=(1, R))
    I0 = b.centroid_intensity
    zR = np.pi * R ** 2 / wl
    for z in [2 * mm, 4 * mm, 6 * mm]:
        b.propagate(2 * mm)
        zn = z / zR
        assert b.centroid_intensity / I0 == pytest.approx(1 / (1 + zn ** 2), rel=2e-3)
        assert (b.D4sigma[0] / 2) / R == pytest.approx(np.sqrt(1 + zn ** 2), rel=2e-3)
        ph = b.phiprofile[b.centroid[3], b.centroid[2]] - 2 * np.pi * (np.trunc(b.k0 * z) - b.k0 * z)
        ph = (ph + np.pi) % (2 * np.pi) - np.pi
        assert ph == pytest.approx(np.arctan(zn), abs=2e-3)


def test_config1_gaussian_vs_analytic_field():
    """256^2 c128, area 200 um, wl 632 nm, rho0 25 um, propagate 100 um: compare with
    the paraxial Gaussian (R/w)^2 exp(-2 r^2/w^2) intensity (paraxial gap ~1e-4)."""
    n, area, wl, R = 256, 200 * um, 632 * nm, 25 * um
    b = OracleBeam2D(area, n, wl, init_field_gen=P.gaussian_beam, init_gen_args=(1, R))
    z = 100 * um
    b.propagate(z)
    zR = np.pi * R ** 2 / wl
    w = R * np.sqrt(1 + (z / zR) ** 2)
    r2 = b.X ** 2 + b.Y ** 2
    theory = (R / w) ** 2 * np.exp(-2 * r2 / w ** 2)
    assert rel_l2(b.iprofile, theory) < 1e-3
    assert b.iprofile[n // 2, n // 2] == pytest.approx(1 / (1 + (z / zR) ** 2), rel=1e-4)


def test_round_hole_on_axis():
    """Reference tests/hole_diffraction.py:44-47 formula 4 sin^2(pi a^2/(2 wl z)), with
    the true radius a = R/sqrt(2) of round_hole(R) (profiles.py:79-80)."""
    npoints, area, wl, R = 512, 3 * mm, 532e-7, 150 * um
    b = OracleBeam2D(area, npoints, wl, init_field_gen=P.round_hole, init_gen_args=(R,))
    a2 = (R / np.sqrt(2)) ** 2
    # pixelated disc: use its actual area for the effective radius
    a2 = (b.field.real > 0).sum() * b.dL ** 2 / np.pi
    for z in [40 * 0.02, 60 * 0.02, 90 * 0.02]:
        b.propagate(z - b.z)
        expect = 4 * np.sin(np.pi * a2 / (2 * wl * z)) ** 2
        assert b.iprofile[npoints // 2, npoints // 2] == pytest.approx(expect, abs=0.15)


def test_square_hole_on_axis():
    """Reference tests/hole_diffraction.py:76-100 (Fresnel integrals)."""
    npoints, area, wl = 512, 3 * mm, 532e-7
    d = area / 8
    b = OracleBeam2D(area, npoints, wl, init_field_gen=P.square_hole, init_gen_args=(d,))
    deff = np.sqrt((b.field.real > 0).sum()) * b.dL
    for z in [40 * 0.02, 60 * 0.02]:
        b.propagate(z - b.z)
        zn = z * wl * 2 / deff ** 2
        c1, s1 = fresnel(-np.sqrt(1 / zn))
        c2, s2 = fresnel(np.sqrt(1 / zn))
        expect = 0.25 * ((c1 - c2) ** 2 + (s1 - s2) ** 2) ** 2
        assert b.iprofile[npoints // 2, npoints // 2] == pytest.approx(expect, rel=0.05)


def test_propagate_invariants():
    n, area, wl = 64, 200 * um, 632 * nm
    rs = np.random.RandomState(11)
    f = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    a = OracleBeam2D(area, n, wl, init_field=f.copy())
    b = OracleBeam2D(area, n, wl, init_field=f.copy())
    a.propagate(0.3 * mm)
    a.propagate(0.5 * mm)
    b.propagate(0.8 * mm)
    assert rel_l2(a.field, b.field) < 1e-10                       # composition
    a.propagate(-0.8 * mm)
    assert rel_l2(a.field, f) < 1e-10                             # reversibility
    n0 = np.sum(np.abs(f) ** 2)
    assert np.sum(np.abs(b.field) ** 2) == pytest.approx(n0, rel=1e-12)   # Parseval
    z0 = b.z
    b.propagate(0)
    assert b.z == z0                                              # z == 0 early return


def test_evanescent_pass_through():
    n, area, wl = 32, 8 * um, 632 * nm
    rs = np.random.RandomState(5)
    f = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    b = OracleBeam2D(area, n, wl, init_field=f)
    s0 = b.spectrum.copy()
    b.propagate(3 * um)
    ev = b.Kz == 0
    assert ev.sum() > 0
    np.testing.assert_array_equal(b.spectrum[ev], s0[ev])


def test_lens_focuses_with_effective_focal_length():
    """lens(f) applies exp(i r^2 k0/(2 f)) with k0 = 1/wl (no 2*pi), i.e. a thin lens of
    focal length 2*pi*f under the exp(-2*pi*i*Kz*z) convention; beyond-focus sign check."""
    n, area, wl = 256, 2 * mm, 532 * nm
    b = OracleBeam2D(area, n, wl, init_field_gen=P.gaussian_beam, init_gen_args=(1, 0.4 * mm))
    f = 1.0 / (2 * np.pi) * 10  # effective focal length 10 cm
    b.lens(f)
    b.propagate(10.0)
    w_focus = b.D4sigma[0] / 2
    theory = wl * 10.0 / (np.pi * 0.4 * mm)     # focal spot 1/e^2 radius
    assert w_focus == pytest.approx(theory, rel=0.05)


def test_c64_state_close_to_c128():
    n, area, wl = 64, 400 * um, 632 * nm
    x = (np.arange(n) - n // 2) * area / n
    f = P.random_round_hole_phase(x, x.reshape(-1, 1), 100 * um, rng=np.random.RandomState(0))
    a = OracleBeam2D(area, n, wl, init_field=f.astype(np.complex64), complex_bits=64)
    b = OracleBeam2D(area, n, wl, init_field=f.astype(np.complex64), complex_bits=128)
    for _ in range(3):
        a.propagate(50 * um)
        b.propagate(50 * um)
    assert a.field.dtype == np.complex64
    assert rel_l2(a.field, b.field) < 1e-5


def test_expand_is_edge_padding():
    # beam2d.py:385-404 restated: equals numpy.pad(mode="edge"), dL kept, spectrum = fft2
    rs = np.random.RandomState(4)
    n = 64
    b = OracleBeam2D(100e-4, n, 632e-7,
                     init_field=rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n)))
    f0 = b.field.copy()
    b.expand(200e-4)
    assert b.npoints == 128 and b.area_size == 200e-4
    assert abs(b.dL - 100e-4 / 64) < 1e-18
    np.testing.assert_array_equal(b.field, np.pad(f0, 32, mode="edge"))
    np.testing.assert_allclose(b.spectrum, np.fft.fft2(b.field), rtol=0, atol=1e-12)
    # odd growth is rounded down to an even npoints (beam2d.py:387)
    c = OracleBeam2D(100e-4, 64, 632e-7, init_field=f0)
    c.expand(151e-4)
    assert c.npoints == 96 and c.field.shape == (96, 96)
    np.testing.assert_array_equal(c.field, np.pad(f0, 16, mode="edge"))


def test_coarse_is_block_mean():
    # beam2d.py:454-460 restated: equals an explicit loop over m x m blocks
    rs = np.random.RandomState(5)
    n, m = 48, 3
    f0 = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    b = OracleBeam2D(100e-4, n, 632e-7, init_field=f0)
    b.coarse(m)
    assert b.npoints == n // m
    exp = np.empty((n // m, n // m), complex)
    for J in range(n // m):
        for I in range(n // m):
            exp[J, I] = sum(f0[J * m + u, I * m + v] for u in range(m) for v in range(m)) / m ** 2
    np.testing.assert_allclose(b.field, exp, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(b.spectrum, np.fft.fft2(b.field), rtol=0, atol=1e-12)
    # a constant field stays constant through expand + coarse
    c = OracleBeam2D(100e-4, 32, 632e-7, init_field=np.full((32, 32), 2 - 1j))
    c.expand(200e-4)
    c.coarse(2)
    np.testing.assert_allclose(c.field, np.full((32, 32), 2 - 1j), rtol=0, atol=1e-15)


def test_interp2d_linear_matches_fitpack():
    # interp2d(kind='linear') on a rectangular grid ran FITPACK's regrid_smth (kx = ky =
    # 1, s = 0) and bispev; RectBivariateSpline runs the same two routines (its axes are
    # (y, x) for z[y, x]).  Uneven knots, queries outside, on and between the knots,
    # given unsorted (interp2d sorts them).
    from scipy.interpolate import RectBivariateSpline
    from oracle.angular_spectrum import interp2d_linear
    rs = np.random.RandomState(21)
    x = np.cumsum(rs.uniform(0.2, 1.5, 17)) - 6.0
    y = np.cumsum(rs.uniform(0.1, 2.0, 11)) - 4.0
    z = rs.standard_normal((11, 17)) + 1j * rs.standard_normal((11, 17))
    xq = np.concatenate([rs.uniform(x[0] - 3, x[-1] + 3, 40), x[[0, 5, 16]]])
    yq = np.concatenate([rs.uniform(y[0] - 3, y[-1] + 3, 30), y[[0, 4, 10]]])
    rs.shuffle(xq)
    rs.shuffle(yq)
    got = interp2d_linear(x, y, z, xq, yq)
    xs, ys = np.sort(xq), np.sort(yq)
    exp = (RectBivariateSpline(y, x, z.real, kx=1, ky=1, s=0)(ys, xs)
           + 1j * RectBivariateSpline(y, x, z.imag, kx=1, ky=1, s=0)(ys, xs))
    assert got.shape == (len(yq), len(xq))
    np.testing.assert_allclose(got, exp, rtol=1e-13, atol=1e-13)
    # a transposed or unsorted evaluation would not agree
    assert np.abs(interp2d_linear(x, y, z, xq, yq)[:, ::-1] - exp).max() > 0.1


def test_interp2d_linear_closed_forms():
    # exact on a + b x + c y + d x y, the samples themselves at the knots, the edge
    # samples outside the grid
    from oracle.angular_spectrum import interp2d_linear
    x = np.arange(-8, 8) * 0.25
    y = np.arange(-5, 7) * 0.5

    def fxy(X, Y):
        return (1.5 - 2j) + 0.75 * X + (0.25 + 1j) * Y - 0.5j * X * Y

    z = fxy(x[None, :], y[:, None])
    xq = np.linspace(x[0], x[-1], 53)
    yq = np.linspace(y[0], y[-1], 29)
    np.testing.assert_allclose(interp2d_linear(x, y, z, xq, yq), fxy(xq[None, :], yq[:, None]),
                               rtol=0, atol=1e-13)
    np.testing.assert_allclose(interp2d_linear(x, y, z, x, y), z, rtol=0, atol=1e-15)
    out = interp2d_linear(x, y, z, [-100.0, 100.0], [-100.0, 100.0])
    np.testing.assert_allclose(out, z[np.ix_([0, -1], [0, -1])], rtol=0, atol=1e-15)


def test_crop_keeps_the_window_on_the_old_samples():
    # crop to 2h samples of the same spacing: the new grid is the old window, so the
    # field is that window (interp2d at its own knots)
    rs = np.random.RandomState(22)
    n, area = 64, 128e-4
    f0 = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    b = OracleBeam2D(area, n, 632e-7, init_field=f0)
    dL = area / n
    b.crop(33 * dL, 32)        # |X| < 16.5 dL keeps X = -16 .. 16 dL: 33 samples, n = 16
    assert b.npoints == 32 and abs(b.dL - 33 * dL / 32) < 1e-18
    c = OracleBeam2D(area, n, 632e-7, init_field=f0)
    c.crop(32 * dL, 32)        # |X| < 16 dL: 31 samples -> the 30 around the centre
    assert c.field.dtype == np.complex128 and c.field.shape == (32, 32)
    # new X = -16 .. 15 dL on the old spacing: X[1:31] are the window's knots
    # (f0[17:47]); X = -16 and 15 dL clamp to the window's edges
    np.testing.assert_allclose(c.field[1:31, 1:31], f0[17:47, 17:47], rtol=0, atol=1e-14)
    np.testing.assert_allclose(c.field[0, 1:31], f0[17, 17:47], rtol=0, atol=1e-14)
    np.testing.assert_allclose(c.field[31, 31], f0[46, 46], rtol=0, atol=1e-14)
    np.testing.assert_allclose(c.spectrum, np.fft.fft2(c.field), rtol=0, atol=1e-12)
    # npoints = 0 keeps the number of points (a 2x zoom here): |X| < 16 dL keeps the 30
    # samples X = -15 .. 14 dL (f0[17:47]); the new even samples sit at X = -16 .. 15 dL,
    # the two ends clamped to the window's edge
    d = OracleBeam2D(area, n, 632e-7, init_field=f0)
    d.crop(area / 2)
    assert d.npoints == n and d.field.shape == (n, n)
    np.testing.assert_allclose(d.field[2:62:2, 2:62:2], f0[17:47, 17:47], rtol=0, atol=1e-14)
    np.testing.assert_allclose(d.field[0, 2:62:2], f0[17, 17:47], rtol=0, atol=1e-14)
    np.testing.assert_allclose(d.field[62, 62], f0[46, 46], rtol=0, atol=1e-14)
    # the odd samples are the means of their neighbours
    np.testing.assert_allclose(d.field[3, 2:62:2], 0.5 * (f0[17, 17:47] + f0[18, 17:47]),
                               rtol=0, atol=1e-14)


def test_lens_image_closed_forms():
    # l1 = l2: scale -1, interp2d's sorted queries are X shifted by one sample (the last
    # clamped), so the image is not flipped; chirp exp(-i k0 r^2 / l).  l1 = -l2: scale
    # +1, no chirp, the field unchanged.
    rs = np.random.RandomState(23)
    n, area, wl, l = 32, 64e-4, 632e-7, 0.5
    f0 = rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))
    b = OracleBeam2D(area, n, wl, init_field=f0)
    b.lens_image(1.0, l, l)
    k = np.minimum(np.arange(n) + 1, n - 1)
    r2 = b.X[None, :] ** 2 + b.X[:, None] ** 2
    exp = -f0[np.ix_(k, k)] * np.exp(-1j * r2 / wl / l)
    np.testing.assert_allclose(b.field, exp, rtol=0, atol=1e-12)
    assert np.abs(b.X - (np.arange(-n // 2, n // 2) * area / n)).max() == 0
    c = OracleBeam2D(area, n, wl, init_field=f0)
    c.lens_image(1.0, -l, l)
    np.testing.assert_allclose(c.field, f0, rtol=0, atol=1e-14)


def test_threaded_baseline_matches_the_oracle():
    # the multi-threaded CPU baseline computes the same propagation
    from oracle.angular_spectrum import propagate_field_threaded
    rs = np.random.RandomState(8)
    u = rs.standard_normal((128, 128)) + 1j * rs.standard_normal((128, 128))
    a = oracle.propagate_field(u, 200e-4, 632e-7, 3e-3)
    b = propagate_field_threaded(u, 200e-4, 632e-7, 3e-3, 2)
    assert rel_l2(b, a) < 1e-13


def _orthonormal_modes(m, nb, seed, real=False):
    # columns of a QR factor: an orthonormal set of m flattened nb x nb modes
    rs = np.random.RandomState(seed)
    a = rs.standard_normal((nb * nb, m))
    if not real:
        a = a + 1j * rs.standard_normal((nb * nb, m))
    q, _ = np.linalg.qr(a)
    return q.T.copy()  # (m, nb^2), rows orthonormal


def test_modes_round_trip_and_padding():
    # construct then deconstruct recovers the coefficients for a real orthonormal basis
    # (the reference's unconjugated 'scalar' product is exact there), for any basis by
    # least squares, and the padded basis is numpy.pad with zeros
    n, nb, m = 48, 32, 12
    modes = _orthonormal_modes(m, nb, 1, real=True)
    c = np.random.RandomState(2).standard_normal(m) + 1j * np.random.RandomState(3).standard_normal(m)
    b = OracleBeam2D(100e-4, n, 632e-7, init_field=np.zeros((n, n)))
    b.construct_by_modes(list(modes), c)
    e = b._expand_basis(list(modes))
    np.testing.assert_array_equal(e[3], np.pad(modes[3].reshape(nb, nb), 8))
    np.testing.assert_allclose(b.spectrum, np.fft.fft2(b.field), atol=1e-12)
    np.testing.assert_allclose(b.deconstruct_by_modes(list(modes)), c, atol=1e-12)
    cm = _orthonormal_modes(m, nb, 4) * np.exp(1j * 0.3)  # complex, not Hermitian-orthonormal
    cm = cm + 0.2 * np.roll(cm, 1, axis=0)                # and not orthogonal at all
    b.construct_by_modes(list(cm), c)
    np.testing.assert_allclose(b.deconstruct_by_modes(list(cm), mode="lstsq"), c, atol=1e-10)
    gram = cm.dot(cm.T)
    np.testing.assert_allclose(b.fast_deconstruct_by_modes(cm, gram), c, atol=1e-9)
    # a unit coefficient reconstructs its own padded mode
    b.construct_by_modes(list(cm), np.eye(m)[5])
    np.testing.assert_allclose(b.field, np.pad(cm[5].reshape(nb, nb), 8), atol=1e-15)


def test_prepared_grid_propagation_in_both_state_precisions():
    """propagate_field_prepared (grids built once, the CPU baseline's form) agrees with
    propagate_field: to rounding in complex128 (numpy and scipy.fft), to complex64
    rounding in a complex64 state, whose spectrum and field stay complex64."""
    from oracle.angular_spectrum import construct_grids, propagate_field, propagate_field_prepared
    rs = np.random.RandomState(3)
    n, area, wl, z = 128, 300e-4, 632e-7, 0.02
    u = (rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))).astype(np.complex64)
    g = construct_grids(area, n, wl)
    ref = propagate_field(u, area, wl, z)
    a = propagate_field_prepared(u, g, z, np.complex128)
    b = propagate_field_prepared(u, g, z, np.complex128, workers=2)
    c = propagate_field_prepared(u, g, z, np.complex64)
    assert a.dtype == np.complex128 and c.dtype == np.complex64
    assert np.linalg.norm(a - ref) / np.linalg.norm(ref) < 1e-14
    assert np.linalg.norm(b - ref) / np.linalg.norm(ref) < 1e-14
    assert 1e-9 < np.linalg.norm(c - ref) / np.linalg.norm(ref) < 1e-6
