"""GPU parity: the HIP path against the oracle on the same seeded inputs.

Tolerances (relative L2 over the whole array):
  complex64  results vs the complex128 oracle:  1e-5  (the stated target)
  complex128 results vs the complex128 oracle:  1e-10
All calls go through liblp2d (C ABI) via lightprop2d_amd.
"""
import json
import os

import numpy as np
import pytest
import torch

from conftest import rel_l2
import oracle
from oracle import profiles as OP
from oracle.angular_spectrum import OracleBeam2D, construct_grids, circular_aperture

pytestmark = pytest.mark.gpu

TOL64 = 1e-5
TOL128 = 1e-10
DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from lightprop2d_amd import _lib
    _lib.load()


def cplx(rs, shape, dt=np.complex128):
    return (rs.standard_normal(shape) + 1j * rs.standard_normal(shape)).astype(dt)


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def host(t):
    return t.detach().cpu().numpy()


def tol(dt):
    return TOL64 if dt == np.complex64 else TOL128


# ------------------------------------------------------------------ FFT
FFT_SIZES = [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 2048, 4096]


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("n", FFT_SIZES)
def test_fft2_parity(n, dt):
    from lightprop2d_amd import propagation as P
    batch = 3 if n <= 512 else 1
    rs = np.random.RandomState(n)
    a = cplx(rs, (batch, n, n), dt)
    ta = to_dev(a)
    F = host(P.fft2(ta))
    assert rel_l2(F, np.fft.fft2(a.astype(np.complex128))) < tol(dt) / 10
    G = host(P.ifft2(ta))
    assert rel_l2(G, np.fft.ifft2(a.astype(np.complex128))) < tol(dt) / 10
    # in place
    tb = ta.clone()
    P.fft2(tb, out=tb)
    assert rel_l2(host(tb), F) < 1e-12


# ------------------------------------------------------------------ golden fixtures via Beam2D
GOLDEN = ["gauss_propagate_n32", "speckle_aperture_n64", "evanescent_n32",
          "lens_chain_n64", "coord_filter_n16"]


def run_product_case(meta, z, bits):
    from lightprop2d_amd import Beam2D
    arrays = {k[4:]: z[k] for k in z.files if k.startswith("arr_")}
    b = Beam2D(meta["area"], meta["n"], meta["wl"], init_field=z["field_in"], complex_bits=bits)
    for op, arg in meta["ops"]:
        val = arrays[arg] if isinstance(arg, str) else arg
        if op == "coordinate_filter":
            b.coordinate_filter(f_init=val)
        elif op == "spectral_filter":
            b.spectral_filter(f_init=val)
        else:
            getattr(b, op)(val)
    return b


@pytest.mark.parametrize("bits", [64, 128])
@pytest.mark.parametrize("name", GOLDEN)
def test_golden_beam2d(golden_dir, name, bits):
    z = np.load(os.path.join(golden_dir, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    b = run_product_case(meta, z, bits)
    t = TOL64 if bits == 64 else TOL128
    assert rel_l2(host(b.field), z["field_out"]) < t
    assert rel_l2(host(b.spectrum), z["spectrum_out"]) < t


# ------------------------------------------------------------------ fused propagation
CASES = [
    # n, area, wl, z, phase, aperture(frac of Nyquist), lens_f
    (16, 8e-4, 632e-7, 1e-4, False, 0.0, 0.0),
    (64, 200e-4, 632e-7, 100e-4, True, 0.0, 0.0),
    (128, 200e-4, 532e-7, 0.1, True, 0.6, 0.0),
    (256, 200e-4, 632e-7, 0.05, True, 0.0, 0.5),
    (512, 500e-4, 1064e-7, 0.1, True, 0.0, 0.5),
    (1024, 400e-4, 632e-7, 50e-4, True, 0.4, 0.0),
    (256, 6e-4, 632e-7, 2e-4, False, 0.0, 0.0),   # evanescent modes present
]


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
@pytest.mark.parametrize("case", CASES, ids=[f"n{c[0]}-{i}" for i, c in enumerate(CASES)])
def test_propagate_fused(case, dt):
    from lightprop2d_amd import propagation as P
    n, area, wl, z, use_phase, apf, lens = case
    rs = np.random.RandomState(n + 7)
    batch = 2 if n <= 512 else 1
    u = cplx(rs, (batch, n, n), dt)
    phi = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32) if use_phase else None
    ap = apf * 0.5 / (area / n) if apf else 0.0
    tu = to_dev(u)
    out = P.propagate(tu, z, area, wl, phase=to_dev(phi) if use_phase else None,
                      aperture_radius=ap, lens_f=lens)
    exp = np.stack([oracle.propagate_field(u[i], area, wl, z,
                                           phase_mask=phi.astype(np.float64) if use_phase else None,
                                           aperture_radius=ap if ap else None,
                                           lens_f=lens if lens else None)
                    for i in range(batch)])
    assert rel_l2(host(out), exp) < tol(dt)
    # in place gives the same
    P.propagate(tu, z, area, wl, phase=to_dev(phi) if use_phase else None, aperture_radius=ap,
                lens_f=lens, out=tu)
    assert rel_l2(host(tu), host(out)) < 1e-12


def test_propagate_spectrum_out():
    from lightprop2d_amd import propagation as P
    n, area, wl, z = 256, 300e-4, 532e-7, 0.02
    u = cplx(np.random.RandomState(1), (n, n), np.complex64)
    spec = torch.empty((n, n), dtype=torch.complex64, device=DEV)
    P.propagate(to_dev(u), z, area, wl, spectrum_out=spec)
    g = construct_grids(area, n, wl)
    exp = np.fft.fft2(u.astype(np.complex128)) * oracle.transfer_function(g.Kz, z)
    assert rel_l2(host(spec), exp) < TOL64


# ------------------------------------------------------------------ adjoint
@pytest.mark.parametrize("name", ["adjoint_n32", "adjoint_evanescent_n16"])
@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
def test_adjoint_golden(golden_dir, name, dt):
    from lightprop2d_amd import propagation as P
    z = np.load(os.path.join(golden_dir, name + ".npz"))
    meta = json.loads(str(z["meta"]))
    tu = to_dev(z["u"].astype(dt))
    tphi = to_dev(z["phi"].astype(np.float32))
    fwd = P.propagate(tu, meta["z"], meta["area"], meta["wl"], phase=tphi)
    phi32 = z["phi"].astype(np.float32).astype(np.float64)
    fwd_o = oracle.propagate_field(z["u"].astype(dt), meta["area"], meta["wl"], meta["z"],
                                   phase_mask=phi32)
    assert rel_l2(fwd_o, z["fwd"]) < 1e-6      # only the float32 rounding of the phase
    assert rel_l2(host(fwd), fwd_o) < (TOL64 if dt == np.complex64 else 1e-9)
    gu, gp = P.propagate_adjoint(to_dev(z["g"].astype(dt)), tu, meta["z"], meta["area"],
                                 meta["wl"], phase=tphi)
    # the phase was rounded to float32 on the way in: compare with the oracle at that phase
    gu_o, gp_o = oracle.propagate_field_adjoint(z["g"].astype(dt), z["u"].astype(dt), meta["area"],
                                                meta["wl"], meta["z"], phase_mask=phi32)
    t = TOL64 if dt == np.complex64 else 1e-9
    assert rel_l2(host(gu), gu_o) < t
    # grad_phase is float32 by contract (the phase screen is float32)
    assert rel_l2(host(gp), gp_o) < (TOL64 if dt == np.complex64 else 1e-7)


@pytest.mark.parametrize("n", [64, 256, 1024])
def test_adjoint_dot_product(n):
    """<A u, g> == <u, A^H g> for the fused operator with aperture and lens."""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(n)
    area, wl, z = 300e-4, 632e-7, 0.03
    u = cplx(rs, (n, n))
    g = cplx(rs, (n, n))
    ap = 0.3 / (area / n)
    Au = host(P.propagate(to_dev(u), z, area, wl, aperture_radius=ap, lens_f=0.7))
    AHg, _ = P.propagate_adjoint(to_dev(g), to_dev(u), z, area, wl, aperture_radius=ap, lens_f=0.7)
    lhs = np.vdot(g, Au)
    rhs = np.vdot(host(AHg), u)
    assert abs(lhs - rhs) < 1e-10 * abs(lhs)


@pytest.mark.parametrize("n,batch", [(128, 3), (512, 5), (1024, 12)])
def test_adjoint_shared_phase_batch_reduction(n, batch):
    """Shared screen: the phase gradient is summed over the batch inside the kernels
    (fused into the last row pass for n >= 256); grad_in can be skipped."""
    from lightprop2d_amd import propagation as P
    area, wl, z = 200e-4, 532e-7, 0.01
    rs = np.random.RandomState(9)
    u = cplx(rs, (batch, n, n), np.complex64)
    g = cplx(rs, (batch, n, n), np.complex64)
    phi = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    gu, gp = P.propagate_adjoint(to_dev(g), to_dev(u), z, area, wl, phase=to_dev(phi))
    exp_u, exp_p = [], 0
    for i in range(batch):
        a, b = oracle.propagate_field_adjoint(g[i], u[i], area, wl, z, phase_mask=phi.astype(np.float64))
        exp_u.append(a)
        exp_p = exp_p + b
    assert gp.shape == (n, n)
    assert rel_l2(host(gp), exp_p) < TOL64
    assert rel_l2(host(gu), np.stack(exp_u)) < TOL64
    gu2, gp2 = P.propagate_adjoint(to_dev(g), to_dev(u), z, area, wl, phase=to_dev(phi),
                                   want_field_grad=False)
    assert gu2 is None
    assert rel_l2(host(gp2), host(gp)) < 1e-6


# ------------------------------------------------------------------ grids / intensity
@pytest.mark.parametrize("n,area,wl", [(32, 8e-4, 632e-7), (256, 200e-4, 632e-7), (1024, 6e-3, 800e-7)])
def test_kz_grid_bit_exact(n, area, wl):
    from lightprop2d_amd import propagation as P
    kz = host(P.kz_grid(n, area, wl))
    np.testing.assert_array_equal(kz, construct_grids(area, n, wl).Kz)


def test_intensity():
    from lightprop2d_amd import propagation as P
    u = cplx(np.random.RandomState(2), (2, 64, 64), np.complex64)
    np.testing.assert_allclose(host(P.intensity(to_dev(u))), np.abs(u) ** 2, rtol=1e-6)
    u = u.astype(np.complex128)
    np.testing.assert_allclose(host(P.intensity(to_dev(u))), np.abs(u) ** 2, rtol=1e-14)


# ------------------------------------------------------------------ Beam2D semantics
def test_beam2d_state_semantics():
    from lightprop2d_amd import Beam2D, um, nm, gaussian_beam
    b = Beam2D(200 * um, 64, 632 * nm, init_field_gen=gaussian_beam, init_gen_args=(1, 25 * um))
    assert b.field.dtype == torch.complex128 and b.field.is_cuda
    s0 = b.spectrum
    b.propagate(0)
    assert b.z == 0 and b.spectrum is s0
    b.propagate(10 * um)
    b.propagate(5 * um)
    assert b.z == pytest.approx(15 * um)
    assert b.spectrum is s0                    # multiplied in place, like the reference
    assert isinstance(b.iprofile, np.ndarray)
    b.numpy_output = False
    assert isinstance(b.iprofile, torch.Tensor)
    with pytest.raises(ValueError):
        Beam2D(200 * um, 64, 632 * nm)
    with pytest.raises(AttributeError):
        Beam2D(200 * um, 64, 632 * nm, init_field=np.ones((64, 64)), complex_bits=32)
    assert "Beam 64x64 points" in repr(b)


def test_beam2d_device_init_field_is_aliased():
    """A device init_field of the beam's dtype becomes the beam's field itself, as the
    reference keeps the array it is given (beam2d.py:122-129, _asxp :154-187), so
    coordinate_filter's in-place multiply (`field *=`, :338-341) reaches the caller's
    tensor.  A host array is copied to the device (the reference's cupy path copies
    host arrays too) and another dtype is converted, so neither is touched."""
    from lightprop2d_amd import Beam2D, um, nm
    n = 64
    t = torch.ones((n, n), dtype=torch.complex128, device="cuda")
    b = Beam2D(200 * um, n, 632 * nm, init_field=t)
    assert b.field is t
    mask = torch.full((n, n), 0.5, dtype=torch.float64, device="cuda")
    b.coordinate_filter(f_init=mask)
    assert b.field is t and torch.equal(t, torch.full_like(t, 0.5))
    h = np.ones((n, n), dtype=np.complex128)
    b2 = Beam2D(200 * um, n, 632 * nm, init_field=h)
    b2.coordinate_filter(f_init=mask)
    assert np.all(h == 1)
    t64 = torch.ones((n, n), dtype=torch.complex64, device="cuda")
    b3 = Beam2D(200 * um, n, 632 * nm, init_field=t64)
    assert b3.field is not t64 and b3.field.dtype == torch.complex128
    b3.coordinate_filter(f_init=mask)
    assert torch.equal(t64, torch.ones_like(t64))


def test_beam2d_filters_with_generators():
    """coordinate_filter / spectral_filter with f_gen on the device grids vs the oracle."""
    from lightprop2d_amd import Beam2D, round_hole, um, nm
    n, area, wl = 256, 400 * um, 632 * nm
    x = (np.arange(n) - n // 2) * area / n
    f = OP.random_round_hole_phase(x, x.reshape(-1, 1), 60 * um, rng=np.random.RandomState(4))
    a = Beam2D(area, n, wl, init_field=f)
    o = OracleBeam2D(area, n, wl, init_field=f.copy())
    for bm in (a, o):
        gen = round_hole if bm is a else OP.round_hole
        bm.coordinate_filter(f_gen=gen, fargs=(50 * um, 3 * um, -2 * um))
        bm.propagate(80 * um)
        bm.spectral_filter(f_gen=gen, fargs=(0.25 / wl,))
        bm.lens(0.5)
        bm.propagate(0.3)
    assert rel_l2(host(a.field), o.field) < TOL128
    assert rel_l2(host(a.spectrum), o.spectrum) < TOL128


def test_beam2d_lazy_state_observes_reference_values():
    """propagate / spectral_filter leave the field stale (one elementwise sweep each);
    every read in a mixed chain equals the reference state at that point, including a
    spectrum changed in place by the caller (the field of that state is fixed before
    the spectrum is handed out) and assignments of either attribute."""
    from lightprop2d_amd import Beam2D, round_hole, um, nm
    n, area, wl = 128, 300 * um, 632 * nm
    x = (np.arange(n) - n // 2) * area / n
    f = OP.random_round_hole_phase(x, x.reshape(-1, 1), 80 * um, rng=np.random.RandomState(7))
    a = Beam2D(area, n, wl, init_field=f)
    o = OracleBeam2D(area, n, wl, init_field=f.copy())
    ap = 0.3 / wl

    def both(fn_a, fn_o):
        fn_a(a)
        fn_o(o)

    def check():
        assert rel_l2(host(a.field), o.field) < TOL128
        assert rel_l2(host(a.spectrum), o.spectrum) < TOL128

    both(lambda b: b.propagate(40 * um), lambda b: b.propagate(40 * um))
    both(lambda b: b.propagate(60 * um), lambda b: b.propagate(60 * um))   # two pending
    both(lambda b: b.spectral_filter(f_gen=round_hole, fargs=(ap,)),
         lambda b: b.spectral_filter(f_gen=OP.round_hole, fargs=(ap,)))
    assert a._field is None                      # nothing inverse-transformed yet
    check()
    both(lambda b: b.propagate(0.1), lambda b: b.propagate(0.1))
    s_a, s_o = a.spectrum, o.spectrum            # handed out ...
    s_a *= 0.5                                   # ... and changed in place
    s_o *= 0.5
    check()
    both(lambda b: b.lens(0.7), lambda b: b.lens(0.7))
    both(lambda b: b.propagate(20 * um), lambda b: b.propagate(20 * um))
    assert rel_l2(a.iprofile, o.iprofile) < TOL128
    g = np.exp(1j * np.linspace(0, 1, n * n)).reshape(n, n)
    a.field = g.copy()                           # spectrum stays (with H applied)
    o.field = g.copy()
    check()
    a.spectrum = np.fft.fft2(g)
    o.spectrum = np.fft.fft2(g)
    both(lambda b: b.propagate(-30 * um), lambda b: b.propagate(-30 * um))
    check()
    assert a.z == pytest.approx(o.z)


def test_beam2d_numpy_written_generators():
    """f_gen callables written for numpy arrays (np.exp, np.abs) raise on device
    tensors; they are run on host copies of the grids (beam2d.py:335-338, 366-369
    call them on the numpy grids), including one that shifts its arguments in place
    the way gaussian_beam does (profiles.py:69-70)."""
    from lightprop2d_amd import Beam2D, um, nm
    n, area, wl = 128, 300 * um, 632 * nm
    x = (np.arange(n) - n // 2) * area / n
    f = OP.random_round_hole_phase(x, x.reshape(-1, 1), 90 * um, rng=np.random.RandomState(8))

    def soft_disc(xx, yy, r):
        return np.exp(-(xx ** 2 + yy ** 2) / r ** 2)

    def box(kx, ky, c):
        return (np.abs(kx) < c) & (np.abs(ky) < c)

    def shifted(xx, yy, x0):
        xx -= x0          # numpy in-place shift, as gaussian_beam does
        return np.exp(-np.abs(xx) / (50 * um)) + 0 * yy

    a = Beam2D(area, n, wl, init_field=f)
    o = OracleBeam2D(area, n, wl, init_field=f.copy())
    for b in (a, o):
        b.coordinate_filter(f_gen=soft_disc, fargs=(70 * um,))
        b.propagate(0.05)
        b.spectral_filter(f_gen=box, fargs=(0.2 / wl,))
        b.coordinate_filter(f_gen=shifted, fargs=(3 * um,))
    assert rel_l2(host(a.field), o.field) < TOL128
    np.testing.assert_array_equal(host(a.X), o.X)      # the shift is kept, Y with it
    np.testing.assert_array_equal(host(a.Y), o.Y)
    b = Beam2D(area, n, wl, init_field_gen=lambda xx, yy: soft_disc(xx, yy, 40 * um) + 0j)
    np.testing.assert_allclose(host(b.field), soft_disc(x, x.reshape(-1, 1), 40 * um),
                               rtol=1e-15)


def test_beam2d_assigned_kz():
    """An assigned Kz (here the paraxial one) is the one propagate uses, as in the
    reference (beam2d.py:485)."""
    from lightprop2d_amd import Beam2D, gaussian_beam, um, nm
    n, area, wl = 128, 300 * um, 632 * nm
    a = Beam2D(area, n, wl, init_field_gen=gaussian_beam, init_gen_args=(1, 30 * um))
    o = OracleBeam2D(area, n, wl, init_field_gen=OP.gaussian_beam, init_gen_args=(1, 30 * um))
    kz = o.k0 - (o.Kx ** 2 + o.Ky ** 2) / (2 * o.k0)
    o.Kz = kz
    a.Kz = kz.copy()
    for b in (a, o):
        b.propagate(0.2)
        b.spectral_filter(f_init=np.ones((n, n)))
    assert rel_l2(host(a.field), o.field) < TOL128
    full = oracle.propagate_field(host(Beam2D(area, n, wl, init_field_gen=gaussian_beam,
                                              init_gen_args=(1, 30 * um)).field),
                                  area, wl, 0.2)
    assert rel_l2(host(a.field), full) > 1e-6           # not the exact transfer function


# ------------------------------------------------------------------ workloads
def test_config1_gaussian_c128():
    """256^2 c128, area 200 um, wl 632 nm, Gaussian rho0 25 um, propagate 100 um."""
    from lightprop2d_amd import Beam2D, gaussian_beam, um, nm
    n, area, wl, R, z = 256, 200 * um, 632 * nm, 25 * um, 100 * um
    b = Beam2D(area, n, wl, init_field_gen=gaussian_beam, init_gen_args=(1, R))
    o = OracleBeam2D(area, n, wl, init_field_gen=OP.gaussian_beam, init_gen_args=(1, R))
    b.propagate(z)
    o.propagate(z)
    assert rel_l2(host(b.field), o.field) < TOL128
    zR = np.pi * R ** 2 / wl
    w = R * np.sqrt(1 + (z / zR) ** 2)
    r2 = o.X ** 2 + o.Y ** 2
    theory = (R / w) ** 2 * np.exp(-2 * r2 / w ** 2)
    assert rel_l2(b.iprofile, theory) < 1e-3                      # paraxial gap ~1e-4


def test_config2_speckle_c64():
    """1024^2 c64, area 400 um, wl 632 nm, random-phase disc of radius 25 um (reference
    round_hole(R) has radius R/sqrt(2), so R = 25*sqrt(2) um), seed 0; 10 x
    [propagate(50 um), spectral aperture |f| <= NA/wl, NA 0.2]."""
    from lightprop2d_amd import Beam2D, um, nm
    from lightprop2d_amd import propagation as P
    n, area, wl, NA = 1024, 400 * um, 632 * nm, 0.2
    x = (np.arange(n) - n // 2) * area / n
    np.random.seed(0)
    f = OP.random_round_hole_phase(x, x.reshape(-1, 1), 25 * um * np.sqrt(2))
    g = construct_grids(area, n, wl)
    mask = circular_aperture(g.Kx, g.Ky, NA / wl)
    o = OracleBeam2D(area, n, wl, init_field=f.astype(np.complex64), complex_bits=128)
    b = Beam2D(area, n, wl, init_field=f, complex_bits=64)
    tmask = to_dev(mask)
    for _ in range(10):
        o.propagate(50 * um)
        o.spectral_filter(f_init=mask)
        b.propagate(50 * um)
        b.spectral_filter(f_init=tmask)
    assert rel_l2(host(b.field), o.field) < TOL64
    # the fused field-to-field path with the on-chip aperture gives the same field
    u = to_dev(f.astype(np.complex64))
    for _ in range(10):
        P.propagate(u, 50 * um, area, wl, aperture_radius=NA / wl, out=u)
    assert rel_l2(host(u), o.field) < TOL64


def test_uneven_batch_vs_oracle():
    """A scaled-down config-3 mix: 64^2, 128^2, 256^2, 512^2 fields, random z each."""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(0)
    area, wl = 200e-4, 532e-7
    sizes = [512] * 2 + [256] * 3 + [128] * 4 + [64] * 2
    srcs, zs, host_in = [], [], []
    for n in sizes:
        x = (np.arange(n) - n // 2) * area / n
        f = OP.random_round_hole_phase(x, x.reshape(-1, 1), 25e-4, rng=rs).astype(np.complex64)
        host_in.append(f)
        srcs.append(to_dev(f))
        zs.append(rs.uniform(10e-4, 0.1))
    dsts = [torch.empty_like(s) for s in srcs]
    ub = P.UnevenBatch(srcs, dsts, zs, area, wl)
    ub.run()
    torch.cuda.synchronize()
    for f, d, z in zip(host_in, dsts, zs):
        assert rel_l2(host(d), oracle.propagate_field(f, area, wl, z)) < TOL64
    # graph capture + replays give the same answer; in-place batch too
    ref = [host(d) for d in dsts]
    for d in dsts:
        d.zero_()
    gph = ub.capture_graph()
    for _ in range(3):
        gph.replay()
        torch.cuda.synchronize()
        for d, r in zip(dsts, ref):
            assert rel_l2(host(d), r) < 1e-12
    ip = [s.clone() for s in srcs]
    P.UnevenBatch(ip, ip, zs, area, wl).run()
    for d, r in zip(ip, ref):
        assert rel_l2(host(d), r) < 1e-12


def test_full_size_round_trip_4096():
    """4096^2 c64 (config-5 plane size): P(z) then P(-z) is the identity on every mode
    (H(-z) = conj H(z), evanescent H = 1), and Parseval holds without evanescent modes."""
    from lightprop2d_amd import propagation as P
    n, area, wl, z = 4096, 2e-1, 800e-7, 0.2
    rs = np.random.RandomState(5)
    u = cplx(rs, (n, n), np.complex64)
    tu = to_dev(u)
    v = P.propagate(tu, z, area, wl)
    e0 = float((tu.abs() ** 2).sum())
    e1 = float((v.abs() ** 2).sum())
    assert e1 == pytest.approx(e0, rel=1e-5)
    w = P.propagate(v, -z, area, wl)
    assert rel_l2(host(w), u) < 2e-6


def test_4096_batch_of_6_in_column_groups_of_4():
    """6 fields of 4096^2 c64: the column pass takes groups of at most 4 fields at this
    length (a full group of 4 and a partial one of 2) over the 2 x 4 block-tiled plane;
    a field of each group against the oracle, every field through the round trip."""
    from lightprop2d_amd import propagation as P
    n, area, wl, z, na = 4096, 2e-1, 800e-7, 0.2, 0.05
    B = 6
    rs = np.random.RandomState(6)
    u = cplx(rs, (B, n, n), np.complex64)
    tu = to_dev(u)
    v = P.propagate(tu, z, area, wl, aperture_radius=na / wl)
    for b in (1, 5):
        print(f"n=4096 batch=6 field={b}", flush=True)
        exp = oracle.propagate_field(u[b], area, wl, z, aperture_radius=na / wl)
        assert rel_l2(host(v[b]), exp) < 1e-5
    w = P.propagate(P.propagate(tu, z, area, wl), -z, area, wl)
    assert rel_l2(host(w), u) < 2e-6


def test_batch_composition_and_concurrent_streams():
    """32 fields of 1024^2 c64 in one call, in two calls of 16 on one stream, and in the
    same two calls on two streams at once: the concurrent calls equal the sequential
    ones bit for bit (no state shared between calls in flight), and the split equals
    the whole call to 1e-6 (relative L2 per field).  The pipelined passes peel each
    workgroup's first row tile or field off their loop, and the compiler schedules the
    peeled copy a little differently, so a field whose place in that order changes
    with the batch may differ in the last bits."""
    from lightprop2d_amd import propagation as P
    n, B, area, wl, z = 1024, 32, 400e-4, 632e-7, 50e-4
    rs = np.random.RandomState(32)
    u = to_dev(cplx(rs, (B, n, n), np.complex64))
    whole = P.propagate(u, z, area, wl)
    seq = torch.empty_like(u)
    P.propagate(u[:16], z, area, wl, out=seq[:16])
    P.propagate(u[16:], z, area, wl, out=seq[16:])
    con = torch.empty_like(u)
    main = torch.cuda.current_stream()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    s1.wait_stream(main)
    s2.wait_stream(main)
    with torch.cuda.stream(s1):
        P.propagate(u[:16], z, area, wl, out=con[:16])
    with torch.cuda.stream(s2):
        P.propagate(u[16:], z, area, wl, out=con[16:])
    main.wait_stream(s1)
    main.wait_stream(s2)
    torch.cuda.synchronize()
    assert torch.equal(con, seq)
    for b in range(B):
        r = float((seq[b] - whole[b]).norm() / whole[b].norm())
        assert r < 1e-6, (b, r)


# ------------------------------------------------------------------ diagnostics (beam2d.py:669-752)
@pytest.mark.parametrize("bits", [64, 128])
def test_diagnostics_vs_oracle(bits):
    from lightprop2d_amd import Beam2D, gaussian_beam, um, nm
    n, area, wl = 256, 400 * um, 632 * nm
    b = Beam2D(area, n, wl, init_field_gen=gaussian_beam, init_gen_args=(1, 30 * um, 20 * um, -15 * um),
               complex_bits=bits)
    o = OracleBeam2D(area, n, wl, init_field_gen=OP.gaussian_beam,
                     init_gen_args=(1, 30 * um, 20 * um, -15 * um), complex_bits=bits)
    b.propagate(0.05)
    o.propagate(0.05)
    cb, co = b.centroid, o.centroid
    assert cb[2:] == co[2:]
    assert cb[0] == pytest.approx(co[0], rel=1e-9 if bits == 128 else 1e-5)
    assert cb[1] == pytest.approx(co[1], rel=1e-9 if bits == 128 else 1e-5)
    assert b.D4sigma == pytest.approx(o.D4sigma, rel=1e-9 if bits == 128 else 1e-5)
    assert b.centroid_intensity == pytest.approx(o.centroid_intensity, rel=1e-9 if bits == 128 else 1e-5)
    # phase is compared where the beam has power (elsewhere it is the phase of noise)
    live = o.iprofile > 1e-6 * o.iprofile.max()
    d = np.angle(np.exp(1j * (b.phiprofile[live] - o.phiprofile[live])))
    assert np.abs(d).max() < (1e-9 if bits == 128 else 1e-3)


def test_reference_gaussian_diffraction_script_on_gpu():
    """The reference's analytic check (tests/gaussbeam_diffraction.py:18-79) run through
    the GPU Beam2D: 512^2, area 2 mm, 532 nm, R = 0.1 mm, 2 mm steps; on-axis
    intensity 1/(1+(z/zR)^2), width w/w0 from D4sigma/2, Gouy phase arctan(z/zR)."""
    from lightprop2d_amd import Beam2D, gaussian_beam, mm, um
    npoints, area_size, wl0, R = 512, 2 * mm, 0.532 * um, 0.1 * mm
    beam = Beam2D(area_size, npoints, wl0, init_field_gen=gaussian_beam, init_gen_args=(1, R),
                  use_gpu=True)
    dz = 2 * mm
    z_grid = np.arange(0, 40, dz)[:8]
    intensities, widths, phases = [], [], []
    beam.propagate(z_grid[0])
    for z in z_grid:
        if z > z_grid[0]:
            beam.propagate(dz)
        intensities.append(beam.centroid_intensity)
        widths.append(beam.D4sigma[0] / 2)
        phases.append(beam.phiprofile[beam.centroid[3], beam.centroid[2]])
    zn = 2 * z_grid / (2 * np.pi * beam.k0 * R ** 2)
    np.testing.assert_allclose(np.array(intensities) / intensities[0], 1 / (1 + zn ** 2), rtol=2e-3)
    np.testing.assert_allclose(np.array(widths) / R, np.sqrt(1 + (wl0 * z_grid / np.pi / R ** 2) ** 2),
                               rtol=2e-3)
    ph = np.array(phases) - 2 * np.pi * (np.trunc(beam.k0 * z_grid) - beam.k0 * z_grid)
    ph = (ph + np.pi) % (2 * np.pi) - np.pi
    np.testing.assert_allclose(ph, np.arctan(wl0 * z_grid / np.pi / R ** 2), atol=2e-3)


@pytest.mark.parametrize("dt", [np.complex64, np.complex128], ids=["c64", "c128"])
def test_uneven_batch_per_field_optics(dt):
    """Per-field z, area, wavelength and aperture (some fields without one), sizes
    from 16 to 1024 in one batch, against the oracle field by field."""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(3)
    sizes = [1024, 256, 16, 512, 64, 128, 256, 32]
    srcs, dsts, zs, areas, wls, aps, host_in = [], [], [], [], [], [], []
    for i, n in enumerate(sizes):
        f = cplx(rs, (n, n), dt)
        host_in.append(f)
        srcs.append(to_dev(f))
        dsts.append(torch.empty_like(srcs[-1]))
        zs.append(rs.uniform(1e-4, 0.05))
        areas.append(rs.uniform(100e-4, 400e-4))
        wls.append(rs.uniform(400e-7, 1100e-7))
        aps.append(0.0 if i % 3 == 0 else 0.3 / (areas[-1] / n))
    P.UnevenBatch(srcs, dsts, zs, areas, wls, aps).run()
    torch.cuda.synchronize()
    for f, d, z, a, w, ap in zip(host_in, dsts, zs, areas, wls, aps):
        exp = oracle.propagate_field(f, a, w, z, aperture_radius=ap if ap else None)
        assert rel_l2(host(d), exp) < tol(dt)


def test_uneven_batch_of_4096_fields():
    """Five 4096^2 complex64 fields and one 2048^2 in an uneven batch: the 4096 bucket
    runs the column pass over per-field pointer and optics tables on the 2 x 4
    block-tiled plane, in groups of at most 4 fields (4 + 1).  Per-field z and
    apertures; two 4096 fields (one per group) and the 2048 field against the oracle."""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(9)
    area, wl = 2e-1, 800e-7
    sizes = [4096, 4096, 2048, 4096, 4096, 4096]
    zs = [0.2, 0.1, 0.15, 0.05, 0.3, 0.25]
    aps = [0.05 / wl, 0.0, 0.04 / wl, 0.03 / wl, 0.0, 0.06 / wl]
    host_in = [cplx(rs, (n, n), np.complex64) for n in sizes]
    srcs = [to_dev(f) for f in host_in]
    dsts = [torch.empty_like(t) for t in srcs]
    P.UnevenBatch(srcs, dsts, zs, area, wl, aperture_radius=aps).run()
    torch.cuda.synchronize()
    for i in (0, 2, 5):
        print(f"uneven 4096 batch field={i} n={sizes[i]}", flush=True)
        exp = oracle.propagate_field(host_in[i], area, wl, zs[i],
                                     aperture_radius=aps[i] or None)
        assert rel_l2(host(dsts[i]), exp) < 1e-5


def _config3_inputs(seed=3):
    """Configuration 3 exactly: 2048 x 128^2, 1536 x 256^2, 512 x 512^2 seeded
    random-phase round holes (R = 25 um, area 200 um), then z_i ~ U[10 um, 1 mm], one
    RandomState stream consumed in field order (profiles.py:104-113 per field).
    Returns {n: complex64 array (count, n, n)}, the sizes in field order and zs."""
    area, R = 200e-4, 25e-4
    rs = np.random.RandomState(seed)
    counts = [(128, 2048), (256, 1536), (512, 512)]
    pools = {}
    for n, cnt in counts:
        x = (np.arange(n) - n // 2) * area / n
        hole = OP.round_hole(x, x.reshape(-1, 1), R)
        out = np.empty((cnt, n, n), dtype=np.complex64)
        for c0 in range(0, cnt, 64):  # draws in field order, (n, n) per field
            u = rs.uniform(size=(min(64, cnt - c0), n, n))
            ang = u * hole * 2 * np.pi
            out[c0:c0 + len(u)] = hole * (np.cos(ang) + 1j * np.sin(ang))
        pools[n] = out
    sizes = [n for n, cnt in counts for _ in range(cnt)]
    zs = rs.uniform(10e-4, 0.1, size=len(sizes))
    return pools, sizes, zs


def test_config3_batch_of_4096_fields():
    """The full configuration-3 launch: 4096 fields in one uneven batch (per-field
    tables of 2048, 1536 and 512 entries), out of place and in place.  Twelve sampled
    fields (first, middle and last of each size bucket, and three more) against the
    oracle at 1e-5; Parseval on every field (no evanescent modes at 532 nm: the grid's
    corner frequency is below k0 for all three sizes), in-place results equal to the
    out-of-place ones."""
    from lightprop2d_amd import propagation as P
    area, wl, R = 200e-4, 532e-7, 25e-4
    print("config 3: generating 4096 inputs", flush=True)
    pools, sizes, zs = _config3_inputs()
    # the vectorised draw is the reference generator, field by field
    x = (np.arange(128) - 64) * area / 128
    rs0 = np.random.RandomState(3)
    for k in range(2):
        g = OP.random_round_hole_phase(x, x.reshape(-1, 1), R, rng=rs0).astype(np.complex64)
        assert np.array_equal(g, pools[128][k])
    dev_in = {n: to_dev(p) for n, p in pools.items()}
    dev_out = {n: torch.empty_like(t) for n, t in dev_in.items()}
    srcs, dsts, where = [], [], []
    idx = {128: 0, 256: 0, 512: 0}
    for n in sizes:
        srcs.append(dev_in[n][idx[n]])
        dsts.append(dev_out[n][idx[n]])
        where.append((n, idx[n]))
        idx[n] += 1
    print("config 3: out of place", flush=True)
    ub = P.UnevenBatch(srcs, dsts, zs.tolist(), area, wl)
    ub.run()
    torch.cuda.synchronize()
    first = {128: 0, 256: 2048, 512: 3584}
    counts = {128: 2048, 256: 1536, 512: 512}
    picks = []
    for n in (128, 256, 512):
        picks += [first[n], first[n] + counts[n] // 2, first[n] + counts[n] - 1]
    picks += [77, 2048 + 1001, 3584 + 300]
    for i in picks:
        n, k = where[i]
        exp = oracle.propagate_field(pools[n][k], area, wl, float(zs[i]))
        err = rel_l2(host(dev_out[n][k]), exp)
        assert err < TOL64, f"field {i} (n={n}, z={zs[i]:.4g}): {err:.3e}"
    for n in (128, 256, 512):
        e_in = (dev_in[n].abs() ** 2).sum((1, 2), dtype=torch.float64)
        e_out = (dev_out[n].abs() ** 2).sum((1, 2), dtype=torch.float64)
        rel = ((e_out - e_in).abs() / e_in).max().item()
        assert rel < 1e-5, f"n={n}: Parseval off by {rel:.3e}"
    print("config 3: in place", flush=True)
    ip = {n: t.clone() for n, t in dev_in.items()}
    idx = {128: 0, 256: 0, 512: 0}
    views = []
    for n in sizes:
        views.append(ip[n][idx[n]])
        idx[n] += 1
    P.UnevenBatch(views, views, zs.tolist(), area, wl).run()
    torch.cuda.synchronize()
    for n in (128, 256, 512):
        assert torch.equal(ip[n], dev_out[n]), f"n={n}: in-place result differs"


@pytest.mark.parametrize("dt,tol", [(np.complex64, 1e-5), (np.complex128, 1e-10)],
                         ids=["c64", "c128"])
def test_maximum_size_8192(dt, tol):
    """The largest supported grid (lp2d_max_n = 8192): one field with a pupil against the
    complex128 oracle, then the P(z) P(-z) round trip (pupil-free) and Parseval, at the
    schedules tuned or defaulted for n = 8192."""
    import oracle
    from lightprop2d_amd import propagation as P
    n, area, wl, z = 8192, 4e-1, 800e-7, 0.1
    ap = 0.05 / wl
    rs = np.random.RandomState(8)
    u = cplx(rs, (n, n), dt)
    print(f"n={n} {dt.__name__} with pupil", flush=True)
    tu = to_dev(u)
    got = host(P.propagate(tu, z, area, wl, aperture_radius=ap))
    exp = oracle.propagate_field(u, area, wl, z, aperture_radius=ap)
    assert rel_l2(got, exp) < tol
    del got, exp
    print(f"n={n} {dt.__name__} round trip", flush=True)
    v = P.propagate(tu, z, area, wl)
    assert float((v.abs() ** 2).sum()) == pytest.approx(float((tu.abs() ** 2).sum()), rel=1e-5)
    w = P.propagate(v, -z, area, wl)
    assert rel_l2(host(w), u) < (4e-6 if dt == np.complex64 else 1e-12)
