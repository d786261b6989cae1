"""Mode-basis projection / reconstruction (beam2d.py:538-667) on the device against
the oracle.  The projection multiplies inputs in the field's precision and
accumulates in FP64, so it agrees with numpy to the rounding of the inputs (relative
1e-6 in complex64, 1e-13 in complex128); reconstruction likewise; lstsq and the fast
variant to 1e-9 (complex128)."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
from oracle.angular_spectrum import OracleBeam2D

pytestmark = pytest.mark.gpu

DEV = "cuda"


def modes_basis(m, nb, seed, real=False):
    rs = np.random.RandomState(seed)
    a = rs.standard_normal((nb * nb, m))
    if not real:
        a = a + 1j * rs.standard_normal((nb * nb, m))
    q, _ = np.linalg.qr(a)
    return q.T.copy()


def cplx(rs, shape):
    return rs.standard_normal(shape) + 1j * rs.standard_normal(shape)


@pytest.mark.parametrize("dt,tol", [(np.complex64, 1e-6), (np.complex128, 1e-13)],
                         ids=["c64", "c128"])
@pytest.mark.parametrize("n,nb,m,batch", [(64, 64, 7, 1), (128, 64, 20, 5), (256, 128, 64, 3),
                                          (1024, 256, 16, 2), (32, 30, 1, 9)])
def test_project_and_construct_match_numpy(dt, tol, n, nb, m, batch):
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(n + m)
    modes = cplx(rs, (m, nb * nb)).astype(dt)
    f = cplx(rs, (batch, n, n)).astype(dt)
    d = (n - nb) // 2
    win = f[:, d:n - d, d:n - d].reshape(batch, -1).astype(np.complex128)
    exp = win @ modes.astype(np.complex128).T
    got = P.modes_project(torch.from_numpy(f).to(DEV), torch.from_numpy(modes).to(DEV))
    assert got.dtype == torch.complex128 and tuple(got.shape) == (batch, m)
    assert rel_l2(got.cpu().numpy(), exp) < tol
    c = cplx(rs, (batch, m))
    fe = np.einsum('ijk,bi->bjk', np.pad(modes.astype(np.complex128).reshape(m, nb, nb),
                                          ((0, 0), (d, d), (d, d))), c)
    fg = P.modes_construct(torch.from_numpy(modes).to(DEV), torch.from_numpy(c), n,
                           dtype=torch.from_numpy(f).dtype, device=torch.device(DEV))
    assert rel_l2(fg.cpu().numpy(), fe) < tol


@pytest.mark.parametrize("bits", [64, 128])
def test_beam2d_mode_methods_vs_oracle(bits):
    from lightprop2d_amd import Beam2D, um, nm
    n, nb, m = 256, 128, 24
    area, wl = 200 * um, 632 * nm
    real = modes_basis(m, nb, 3, real=True)
    cm = modes_basis(m, nb, 4) + 0.2 * np.roll(modes_basis(m, nb, 5), 1, axis=0)
    c = cplx(np.random.RandomState(6), m)
    tol = 1e-5 if bits == 64 else 1e-10
    b = Beam2D(area, n, wl, init_field=np.zeros((n, n), complex), complex_bits=bits)
    o = OracleBeam2D(area, n, wl, init_field=np.zeros((n, n), complex))
    for basis in (real, cm):
        b.construct_by_modes(list(basis), c)
        o.construct_by_modes(list(basis), c)
        assert rel_l2(b.field.cpu().numpy(), o.field) < tol
        assert rel_l2(b.spectrum.cpu().numpy(), o.spectrum) < tol
        assert rel_l2(b.deconstruct_by_modes(list(basis)).cpu().numpy(),
                      o.deconstruct_by_modes(list(basis))) < tol
    # least squares recovers the coefficients of a field in the span of any basis
    b.construct_by_modes(list(cm), c)
    assert rel_l2(b.deconstruct_by_modes(list(cm), mode="lstsq").cpu().numpy(), c) < tol
    assert rel_l2(b.fast_deconstruct_by_modes(cm, cm.dot(cm.T)).cpu().numpy(), c) < tol * 10
    # the orthonormal real basis: the scalar product is exact
    b.construct_by_modes(list(real), c)
    assert rel_l2(b.deconstruct_by_modes(list(real)).cpu().numpy(), c) < tol
    e = b._expand_basis(list(real)).cpu().numpy()
    np.testing.assert_allclose(e[2], np.pad(real[2].reshape(nb, nb), 64), atol=1e-6)


def test_modes_reject_bad_windows():
    from lightprop2d_amd import propagation as P
    from lightprop2d_amd._lib import Lp2dError
    f = torch.zeros((64, 64), dtype=torch.complex64, device=DEV)
    with pytest.raises(Lp2dError):
        P.modes_project(f, torch.zeros((3, 63 * 63), dtype=torch.complex64, device=DEV))
    with pytest.raises(Lp2dError):
        P.modes_project(f, torch.zeros((3, 10), dtype=torch.complex64, device=DEV))


@pytest.mark.parametrize("dt,tol", [(np.complex64, 1e-6), (np.complex128, 1e-13)],
                         ids=["c64", "c128"])
@pytest.mark.parametrize("n,nb,m,batch", [(512, 256, 64, 32), (256, 198, 70, 33),
                                          (128, 100, 17, 8), (64, 64, 130, 40), (64, 2, 3, 1)])
def test_project_on_matrix_cores(dt, tol, n, nb, m, batch):
    """Batches through the wide kernels.  The FP64-MFMA projection (64-mode x 32-field tiles, 32-pixel chunks, per-range
    partial sums) against numpy and against the vector kernel: ragged mode, field and
    pixel counts (70 modes, 33 fields, 198^2 pixels) leave partial tiles and chunks;
    both accumulate in FP64, so they agree to summation-order rounding."""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(n + m + batch)
    modes = cplx(rs, (m, nb * nb)).astype(dt)
    f = cplx(rs, (batch, n, n)).astype(dt)
    d = (n - nb) // 2
    win = f[:, d:n - d, d:n - d].reshape(batch, -1).astype(np.complex128)
    exp = win @ modes.astype(np.complex128).T
    ft, mt = torch.from_numpy(f).to(DEV), torch.from_numpy(modes).to(DEV)
    print(f"n={n} nb={nb} modes={m} batch={batch} {dt.__name__}", flush=True)
    got = P.modes_project(ft, mt, matrix_cores=True).reshape(batch, m)
    vec = P.modes_project(ft, mt, matrix_cores=False).reshape(batch, m)
    assert rel_l2(got.cpu().numpy(), exp) < tol
    assert rel_l2(got.cpu().numpy(), vec.cpu().numpy()) < 1e-13
    # reconstruction of the same batch (32 fields per pass above 8, coefficients of
    # 64-mode blocks staged in LDS: 70 and 130 modes leave partial blocks)
    c = cplx(rs, (batch, m))
    fe = np.einsum('ijk,bi->bjk', np.pad(modes.astype(np.complex128).reshape(m, nb, nb),
                                          ((0, 0), (d, d), (d, d))), c)
    fg = P.modes_construct(mt, torch.from_numpy(c), n, dtype=ft.dtype, device=torch.device(DEV))
    assert rel_l2(fg.cpu().numpy().reshape(batch, n, n), fe) < tol
