"""The allocation-free adjoint (lp2d_propagate_adjoint_noalloc) against the adjoint it
stands in for (lp2d_propagate_adjoint), both called through the C ABI on the same
inputs: field gradients and the batch-summed phase gradient of a shared screen must be
identical (same kernels, same summation order; only the home of the partial planes
differs).  Two calls in a row, the second smaller: a three-sweep size whose gradient
pass sums per-group planes in the scratch array, then a whole-field size whose
per-field planes go there."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu
DEV = "cuda"


def _call(lib, L, fn, g, u, phi, psb, o, gin, gph, *extra):
    from lightprop2d_amd import propagation as P
    n, batch, dt = P._field_geom(g, "grad_out")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    args = [n, batch, dt, C.c_void_p(g.data_ptr()), C.c_void_p(u.data_ptr()),
            C.c_void_p(phi.data_ptr()), psb, C.byref(o), C.c_void_p(gin.data_ptr()),
            C.c_void_p(gph.data_ptr()), 0]
    rc = getattr(lib, fn)(*args, *extra, stream)
    assert rc == 0, L.lib().lp2d_last_error().decode()


@pytest.mark.parametrize("dtype", [torch.complex64, torch.complex128], ids=["c64", "c128"])
def test_noalloc_adjoint_matches_the_allocating_one(dtype):
    from lightprop2d_amd import _lib as L
    from lightprop2d_amd import propagation as P
    lib = P._lib()                      # loads the library and runs lp2d_init
    area, wl = 300e-4, 632e-7
    for n, batch, z in ((512, 6, 0.05), (128, 3, 0.02)):
        print(f"noalloc adjoint n={n} batch={batch} {dtype}", flush=True)
        g = torch.Generator(device="cpu").manual_seed(n + batch)
        gout = torch.randn((batch, n, n), dtype=dtype, generator=g).to(DEV)
        u = torch.randn((batch, n, n), dtype=dtype, generator=g).to(DEV)
        phi = (torch.rand((n, n), generator=g) * 6.28).to(torch.float32).to(DEV)
        o = P.optics(area, wl, z=z, aperture_radius=0.3 / (area / n), lens_f=0.8)
        # the allocating adjoint: per-group / per-field planes in grad_phase, sum in plane 0
        gin_ref = torch.empty_like(gout)
        gph_ref = torch.empty((batch, n, n), dtype=torch.float32, device=DEV)
        _call(lib, L, "lp2d_propagate_adjoint", gout, u, phi, 0, o, gin_ref, gph_ref)
        # the allocation-free one: one n x n plane for the sum
        gin = torch.full_like(gout, float("nan"))
        gph = torch.full((n, n), float("nan"), dtype=torch.float32, device=DEV)
        _call(lib, L, "lp2d_propagate_adjoint_noalloc", gout, u, phi, 0, o, gin, gph)
        torch.cuda.synchronize()
        assert torch.equal(gin, gin_ref)
        assert torch.equal(gph, gph_ref[0])
        # per-field screens take no scratch: the same results again
        phib = (torch.rand((batch, n, n), generator=g) * 6.28).to(torch.float32).to(DEV)
        gph_ref = torch.empty((batch, n, n), dtype=torch.float32, device=DEV)
        _call(lib, L, "lp2d_propagate_adjoint", gout, u, phib, n * n, o, gin_ref, gph_ref)
        gph = torch.empty_like(gph_ref)
        _call(lib, L, "lp2d_propagate_adjoint_noalloc", gout, u, phib, n * n, o, gin, gph)
        torch.cuda.synchronize()
        assert torch.equal(gin, gin_ref)
        assert torch.equal(gph, gph_ref)


def test_noalloc_python_wrapper_under_graph_capture():
    """The Python wrapper allocates nothing either: captured into a HIP graph and
    replayed, each replay equals the eager result (program work between replays)."""
    from lightprop2d_amd import propagation as P
    P._lib()
    n, batch, area, wl, z = 256, 4, 300e-4, 632e-7, 0.03
    g = torch.Generator(device="cpu").manual_seed(5)
    gout = torch.randn((batch, n, n), dtype=torch.complex64, generator=g).to(DEV)
    u = torch.randn((batch, n, n), dtype=torch.complex64, generator=g).to(DEV)
    phi = (torch.rand((n, n), generator=g) * 6.28).to(torch.float32).to(DEV)
    gin = torch.empty_like(gout)
    gph = torch.empty((n, n), dtype=torch.float32, device=DEV)
    ref_in, ref_ph = P.propagate_adjoint(gout, u, z, area, wl, phase=phi, use_workspace=False)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            P.propagate_adjoint_noalloc(gout, u, z, area, wl, gin, phase=phi, grad_phase=gph)
    torch.cuda.current_stream().wait_stream(s)
    for k in range(3):
        gin.zero_()
        gph.zero_()
        _ = P.propagate(u, z, area, wl)            # other work between the replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(gin, ref_in), k
        assert torch.equal(gph, ref_ph), k
