"""Every registered kernel variant against the oracle.

The autotuner may select any row-pass or column-pass variant for a size, so each one
is run here on the same seeded batch and compared with the complex128 oracle
(relative L2 1e-5 for complex64, 1e-10 for complex128).  Batches of 5 fields leave a
partial last group in the column classes that loop over groups of fields; the
second configuration is dominated by evanescent modes and also checks the spectrum
output; the uneven batch runs the per-field pointer and optics tables.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from conftest import rel_l2
import oracle
from oracle.angular_spectrum import construct_grids

pytestmark = pytest.mark.gpu

DEV = "cuda"
CASES = [(256, "c64"), (512, "c64"), (1024, "c64"), (2048, "c64"), (4096, "c64"),
         (256, "c128"), (1024, "c128"), (4096, "c128")]
# (area [cm], wl [cm], z [cm], NA of the spectral pupil or 0)
OPTICS = {"pupil": (400e-4, 632e-7, 50e-4, 0.2), "evanescent": (6e-4, 632e-7, 3e-4, 0.0)}


def _lib():
    from lightprop2d_amd import _lib as L
    return L.lib()


def _names(lib, n, dt, kind):
    buf = C.create_string_buffer(128)
    out = []
    for v in range(lib.lp2d_variant_count(n, dt, kind)):
        lib.lp2d_variant_name(n, dt, kind, v, buf, 128)
        out.append(buf.value.decode())
    return out


@pytest.fixture
def restore_selection():
    lib = _lib()
    saved = {(n, dt, k): lib.lp2d_get_variant(n, dt, k)
             for n, _ in CASES for dt in (0, 1) for k in (0, 1)}
    yield lib
    for (n, dt, k), v in saved.items():
        lib.lp2d_set_variant(n, dt, k, v)


@pytest.mark.parametrize("which", sorted(OPTICS))
@pytest.mark.parametrize("n,dtype", CASES, ids=[f"{n}-{d}" for n, d in CASES])
def test_every_variant_matches_oracle(restore_selection, n, dtype, which):
    from lightprop2d_amd import propagation as P
    lib = restore_selection
    dt = 0 if dtype == "c64" else 1
    npdt = np.complex64 if dtype == "c64" else np.complex128
    tol = 1e-5 if dtype == "c64" else 1e-10
    area, wl, z, na = OPTICS[which]
    ap = na / wl
    batch = 5 if n <= 1024 else (3 if n <= 2048 else (2 if dtype == "c64" else 1))
    rs = np.random.RandomState(n + (1 if which == "pupil" else 2))
    u = (rs.standard_normal((batch, n, n)) + 1j * rs.standard_normal((batch, n, n))).astype(npdt)
    exp = np.stack([oracle.propagate_field(u[i], area, wl, z, aperture_radius=ap or None)
                    for i in range(batch)])
    spec_exp = None
    if which == "evanescent":
        g = construct_grids(area, n, wl)
        spec_exp = np.fft.fft2(u.astype(np.complex128)) * oracle.transfer_function(g.Kz, z)
    tu = torch.from_numpy(u).to(DEV)
    for kind in (0, 1):
        default = lib.lp2d_get_variant(n, dt, kind)
        for v, name in enumerate(_names(lib, n, dt, kind)):
            print(f"n={n} {dtype} {which} kind={kind} variant={v} [{name}]", flush=True)
            assert lib.lp2d_set_variant(n, dt, kind, v) == 0
            spec = torch.empty_like(tu) if spec_exp is not None else None
            out = P.propagate(tu, z, area, wl, aperture_radius=ap, spectrum_out=spec)
            got = out.cpu().numpy()
            assert rel_l2(got, exp) < tol, name
            if spec is not None:
                assert rel_l2(spec.cpu().numpy(), spec_exp) < tol, name
        lib.lp2d_set_variant(n, dt, kind, default)


@pytest.mark.parametrize("kind", [0, 1], ids=["rows", "cols"])
@pytest.mark.parametrize("dtype", ["c64", "c128"])
def test_every_variant_on_uneven_tables(restore_selection, dtype, kind):
    """Every row and every column variant of 256 and 512 points over the per-field
    pointer, distance and aperture tables of an uneven batch, with separate
    destinations (a pipelined row variant runs its plain class of the same shape
    there)."""
    from lightprop2d_amd import propagation as P
    lib = restore_selection
    dt = 0 if dtype == "c64" else 1
    npdt = np.complex64 if dtype == "c64" else np.complex128
    tol = 1e-5 if dtype == "c64" else 1e-10
    rs = np.random.RandomState(11)
    sizes = [256, 512, 256, 512, 256]
    zs = [20e-4, 35e-4, 50e-4, 5e-4, 80e-4]
    area, wl = 200e-4, 532e-7
    aps = [0.0, 0.15 / wl, 0.3 / wl, 0.0, 0.1 / wl]
    us = [(rs.standard_normal((s, s)) + 1j * rs.standard_normal((s, s))).astype(npdt)
          for s in sizes]
    exps = [oracle.propagate_field(u, area, wl, z, aperture_radius=a or None)
            for u, z, a in zip(us, zs, aps)]
    srcs = [torch.from_numpy(u).to(DEV) for u in us]
    dsts = [torch.empty_like(s) for s in srcs]
    ub = P.UnevenBatch(srcs, dsts, zs, area, wl, aperture_radius=aps)
    for n in (256, 512):
        default = lib.lp2d_get_variant(n, dt, kind)
        for v, name in enumerate(_names(lib, n, dt, kind)):
            print(f"uneven {dtype} n={n} kind={kind} variant={v} [{name}]", flush=True)
            assert lib.lp2d_set_variant(n, dt, kind, v) == 0
            for d in dsts:
                d.zero_()
            ub.run()
            for s, d, e in zip(sizes, dsts, exps):
                if s == n:
                    assert rel_l2(d.cpu().numpy(), e) < tol, name
        lib.lp2d_set_variant(n, dt, kind, default)


@pytest.mark.parametrize("n,batch", [(128, 8300), (512, 300), (1024, 72), (256, 200), (256, 1)])
def test_pipelined_groups_at_large_and_single_batches(n, batch):
    """Batches that make the pipelined kernels cap their groups at 32 fields per
    workgroup (8300 x 128^2, 300 x 512^2), a partial last group (72 x 1024^2) and one
    field; a sample of fields is checked against the oracle."""
    from lightprop2d_amd import propagation as P
    area, wl, z, na = 200e-4, 532e-7, 80e-4, 0.25
    g = torch.Generator(device="cpu").manual_seed(n + batch)
    u = torch.randn((batch, n, n), dtype=torch.complex64, generator=g).to(DEV)
    out = P.propagate(u, z, area, wl, aperture_radius=na / wl)
    pick = sorted({min(b, batch - 1) for b in (0, batch // 3, batch // 2 + 1, batch - 1)})
    for b in pick:
        print(f"n={n} batch={batch} field={b}", flush=True)
        exp = oracle.propagate_field(u[b].cpu().numpy(), area, wl, z, aperture_radius=na / wl)
        assert rel_l2(out[b].cpu().numpy(), exp) < 1e-5
