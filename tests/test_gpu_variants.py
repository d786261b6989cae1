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
    # through propagation, which installs the tuned selection on first use, so that
    # restore_selection saves (and puts back) the selection every other test runs with
    from lightprop2d_amd import propagation as P
    return P._lib()


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


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
def test_every_row_variant_with_a_phase_screen(restore_selection, n):
    """The forward row pass with a phase screen (the prologue of lp2d_propagate with
    `phase`): every row variant of complex64 n, a screen shared by the batch and one
    per field, against the oracle.  The pipelined class carries the screen itself
    (prefetched with the tile) except at 512 points, where the plain class runs."""
    from lightprop2d_amd import propagation as P
    lib = restore_selection
    area, wl, z, na = OPTICS["pupil"]
    batch = 3 if n <= 1024 else 2   # per-field screens reach field 1 at every size
    rs = np.random.RandomState(3 * n)
    u = (rs.standard_normal((batch, n, n)) + 1j * rs.standard_normal((batch, n, n))).astype(
        np.complex64)
    shared = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    per = rs.uniform(0, 2 * np.pi, (batch, n, n)).astype(np.float32)
    tu = torch.from_numpy(u).to(DEV)
    default = lib.lp2d_get_variant(n, 0, 0)
    for ph in (shared, per):
        exp = np.stack([oracle.propagate_field(u[i], area, wl, z, aperture_radius=na / wl,
                                               phase_mask=(ph if ph.ndim == 2 else ph[i])
                                               .astype(np.float64))
                        for i in range(batch)])
        tph = torch.from_numpy(ph).to(DEV)
        for v, name in enumerate(_names(lib, n, 0, 0)):
            print(f"n={n} screen {ph.shape} row variant={v} [{name}]", flush=True)
            assert lib.lp2d_set_variant(n, 0, 0, v) == 0
            out = P.propagate(tu, z, area, wl, phase=tph, aperture_radius=na / wl)
            assert rel_l2(out.cpu().numpy(), exp) < 1e-5, name
    lib.lp2d_set_variant(n, 0, 0, default)


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


@pytest.mark.parametrize("in_place", [False, True], ids=["out", "in_place"])
@pytest.mark.parametrize("n,batch", [(128, 8300), (512, 300), (1024, 72), (256, 200), (256, 1)])
def test_pipelined_groups_at_large_and_single_batches(n, batch, in_place):
    """Batches that make the pipelined kernels cap their groups at 32 fields per
    workgroup (8300 x 128^2, 300 x 512^2), a partial last group (72 x 1024^2) and one
    field; a sample of fields is checked against the oracle.  Out of place and in
    place: the persistent row passes loop over many tiles per workgroup, and a row pass
    converting between the natural and the block-tiled layout in place must not let
    one wave store over data another wave of its workgroup has yet to load (the inverse
    rows always convert in place; in place, the forward rows do too)."""
    from lightprop2d_amd import propagation as P
    area, wl, z, na = 200e-4, 532e-7, 80e-4, 0.25
    g = torch.Generator(device="cpu").manual_seed(n + batch)
    u = torch.randn((batch, n, n), dtype=torch.complex64, generator=g).to(DEV)
    pick = sorted({min(b, batch - 1) for b in (0, batch // 3, batch // 2 + 1, batch - 1)})
    src = {b: u[b].cpu().numpy() for b in pick}
    out = P.propagate(u, z, area, wl, aperture_radius=na / wl, out=u if in_place else None)
    for b in pick:
        print(f"n={n} batch={batch} field={b} in_place={in_place}", flush=True)
        exp = oracle.propagate_field(src[b], area, wl, z, aperture_radius=na / wl)
        assert rel_l2(out[b].cpu().numpy(), exp) < 1e-5


@pytest.mark.parametrize("n", [256, 512, 1024, 4096])
def test_every_chain_variant_matches_oracle(n):
    """Every chained-plane row pass (kind 3) of lp2d_propagate_stack: a 3-screen stack
    with a pupil against the oracle for one field, the stored middle planes included,
    and equal to the plane-by-plane propagation to rounding."""
    from lightprop2d_amd import propagation as P
    lib = _lib()
    area, wl, dz, K = 300e-4, 800e-7, 0.02, 3
    ap = 0.3 / (area / n)
    B = 1 if n >= 4096 else 3
    rs = np.random.RandomState(n + 7)
    u = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(np.complex64)
    scr = rs.uniform(0, 2 * np.pi, (K, n, n)).astype(np.float32)
    exp = [u[B - 1].astype(np.complex128)]
    for k in range(K):
        exp.append(oracle.propagate_field(exp[-1], area, wl, dz,
                                          phase_mask=scr[k].astype(np.float64),
                                          aperture_radius=ap))
    tu = torch.from_numpy(u).to(DEV)
    ts = torch.from_numpy(scr).to(DEV)
    seq = tu
    for k in range(K):
        seq = P.propagate(seq, dz, area, wl, phase=ts[k], aperture_radius=ap)
    names = _names(lib, n, 0, 3)
    assert len(names) >= 1
    default = lib.lp2d_get_variant(n, 0, 3)
    try:
        for v, name in enumerate(names):
            print(f"chain n={n} variant={v} [{name}]", flush=True)
            assert lib.lp2d_set_variant(n, 0, 3, v) == 0
            planes = [None] + [torch.empty_like(tu) for _ in range(K - 1)]
            out = P.propagate_stack(tu, ts, dz, area, wl, aperture_radius=ap, planes=planes)
            for k in range(1, K + 1):
                got = out if k == K else planes[k]
                assert rel_l2(got[B - 1].cpu().numpy(), exp[k]) < 1e-5, (name, k)
            assert rel_l2(out.cpu().numpy(), seq.cpu().numpy()) < 1e-6, name
    finally:
        lib.lp2d_set_variant(n, 0, 3, default)


def test_autotuner_runs_and_installs_a_measured_choice():
    """autotune.tune at 256^2 complex64: every row, column and chained-pass variant is
    timed, the fastest is installed, and the propagation still matches the oracle."""
    from lightprop2d_amd import autotune
    from lightprop2d_amd import propagation as P
    lib = _lib()
    n = 256
    saved = {k: lib.lp2d_get_variant(n, 0, k) for k in (0, 1, 3)}
    try:
        res = autotune.tune(n, "c64", batch=4, reps=1, save=False)
        names = autotune.variant_names(n, "c64")
        assert len(res["rows"]) == len(names["rows"]) and len(res["cols"]) == len(names["cols"])
        assert len(res["chain"]) == len(names["chain"]) >= 1
        for kind, key in ((0, "rows"), (1, "cols"), (3, "chain")):
            best = min(range(len(res[key])), key=lambda i: res[key][i]["ms"])
            assert lib.lp2d_get_variant(n, 0, kind) == res["best"][key] == best
        rs = np.random.RandomState(21)
        u = (rs.standard_normal((2, n, n)) + 1j * rs.standard_normal((2, n, n))).astype(np.complex64)
        area, wl, z = 200e-4, 632e-7, 30e-4
        out = P.propagate(torch.from_numpy(u).to(DEV), z, area, wl)
        assert rel_l2(out[1].cpu().numpy(), oracle.propagate_field(u[1], area, wl, z)) < 1e-5
    finally:
        for k, v in saved.items():
            lib.lp2d_set_variant(n, 0, k, v)
