"""Plans (lp2d_plan_*): a fixed geometry with the kernel schedule copied at creation.

A plan call runs the same kernels as the plan-less entry point under the selection
the plan was made with, so its results equal the plan-less call bit for bit when the
selection has not changed since; a later lp2d_set_variant changes the plan-less call
(seen in the tracer's kernel names) but not the plan.  The adjoint takes the caller's
workspace and, with a shared screen, returns the batch sum in one plane.  Parity with
the oracle at the plan level is checked once more on a configuration-2 field.
"""
import ctypes as C
import re

import numpy as np
import pytest
import torch

from conftest import rel_l2
import oracle

pytestmark = pytest.mark.gpu

DEV = "cuda"


def _modules():
    from lightprop2d_amd import propagation as P, _lib as L
    return P, L, P._lib()   # with the tuned selection installed


def _rand(shape, dt, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    return torch.randn(shape, dtype=dt, generator=g).to(DEV)


@pytest.mark.parametrize("n,B,dt", [(128, 3, torch.complex64), (1024, 4, torch.complex64),
                                    (256, 2, torch.complex128), (4096, 1, torch.complex64)])
def test_plan_calls_equal_plan_less_calls(n, B, dt):
    P, L, lib = _modules()
    print(f"plan n={n} B={B} {dt}", flush=True)
    area, wl, z = 400e-4, 632e-7, 50e-4
    ap = 0.3 / (area / n)
    u = _rand((B, n, n), dt, n + B)
    ph = (torch.rand((n, n), generator=torch.Generator().manual_seed(3)) * 6).to(DEV)
    plan = P.Plan(n, B, dt)
    # FFT both ways
    assert torch.equal(plan.fft2(u), P.fft2(u))
    assert torch.equal(plan.fft2(u, inverse=True), P.fft2(u, inverse=True))
    # field to field with a shared screen and a pupil, spectrum side output
    s1, s2 = torch.empty_like(u), torch.empty_like(u)
    v1 = plan.propagate(u, z, area, wl, phase=ph, aperture_radius=ap, spectrum_out=s1)
    v2 = P.propagate(u, z, area, wl, phase=ph, aperture_radius=ap, spectrum_out=s2)
    assert torch.equal(v1, v2) and torch.equal(s1, s2)
    # adjoint: shared screen (one summed plane) and per-field screens
    g = _rand((B, n, n), dt, 7 * n)
    a1, gp1 = plan.propagate_adjoint(g, u, z, area, wl, phase=ph, aperture_radius=ap)
    a2, gp2 = P.propagate_adjoint(g, u, z, area, wl, phase=ph, aperture_radius=ap)
    assert gp1.shape == (n, n)
    assert torch.equal(a1, a2) and torch.equal(gp1, gp2)
    phb = (torch.rand((B, n, n), generator=torch.Generator().manual_seed(4)) * 6).to(DEV)
    a1, gp1 = plan.propagate_adjoint(g, u, z, area, wl, phase=phb)
    a2, gp2 = P.propagate_adjoint(g, u, z, area, wl, phase=phb)
    assert gp1.shape == (B, n, n)
    assert torch.equal(a1, a2) and torch.equal(gp1, gp2)
    # a caller workspace one byte short is refused before any launch
    need = plan.workspace_size(L.LP2D_OP_ADJOINT)
    if need:
        short = torch.empty(need - 1, dtype=torch.uint8, device=DEV)
        with pytest.raises(P.Lp2dError):
            plan.propagate_adjoint(g, u, z, area, wl, phase=ph, work=short)
    # a field of another geometry is refused
    with pytest.raises(P.Lp2dError):
        plan.propagate(u[:1].contiguous() if B > 1 else torch.cat([u, u]), z, area, wl)
    plan.close()
    with pytest.raises(P.Lp2dError):
        plan.fft2(u)


def test_plan_keeps_its_schedule():
    """After lp2d_set_variant picks another column schedule for 1024^2, the plan-less
    call runs the new one and the plan still runs the one it was made with; both
    agree with each other to rounding."""
    P, L, lib = _modules()
    n, B, dt = 1024, 8, L.LP2D_C64
    area, wl, z = 400e-4, 632e-7, 50e-4
    u = _rand((B, n, n), torch.complex64, 11)
    before = lib.lp2d_get_variant(n, dt, 1)
    names = []
    buf = C.create_string_buffer(128)
    for v in range(lib.lp2d_variant_count(n, dt, 1)):
        lib.lp2d_variant_name(n, dt, 1, v, buf, 128)
        names.append(buf.value.decode())
    # another schedule with a different shape (E, W: the tracer's name) than the current
    shape = [re.search(r"E=(\d+) W=(\d+)", nm).groups() for nm in names]
    other = next(i for i in range(len(names)) if shape[i] != shape[before])
    plan = P.Plan(n, B, torch.complex64)

    def traced(fn):
        L.check(lib.lp2d_trace_enable(1))
        out = fn()
        torch.cuda.synchronize()
        L.check(lib.lp2d_trace_enable(0))
        arr = (L.lp2d_trace_stat * 64)()
        cnt = C.c_int()
        L.check(lib.lp2d_trace_collect(arr, 64, C.byref(cnt)))
        return out, sorted(arr[i].name.decode() for i in range(cnt.value)
                           if arr[i].name.decode().startswith("cols"))

    try:
        L.check(lib.lp2d_set_variant(n, dt, 1, other))
        v_plan, k_plan = traced(lambda: plan.propagate(u, z, area, wl))
        v_free, k_free = traced(lambda: P.propagate(u, z, area, wl))
        print("plan:", k_plan, "plan-less:", k_free, flush=True)
        assert k_plan != k_free
        assert rel_l2(v_plan.cpu().numpy(), v_free.cpu().numpy()) < 2e-6
        L.check(lib.lp2d_set_variant(n, dt, 1, before))
        v_back, k_back = traced(lambda: P.propagate(u, z, area, wl))
        assert k_back == k_plan and torch.equal(v_back, v_plan)
    finally:
        lib.lp2d_set_variant(n, dt, 1, before)
        plan.close()


def test_plan_stack_and_stack_adjoint_through_the_c_abi():
    """lp2d_plan_propagate_stack / lp2d_plan_propagate_stack_adjoint equal the
    plan-less calls bit for bit (512^2, 3 fields, 3 screens)."""
    P, L, lib = _modules()
    n, B, K, area, wl, dz = 512, 3, 3, 300e-4, 800e-7, 0.02
    u = _rand((B, n, n), torch.complex64, 21)
    scr = (torch.rand((K, n, n), generator=torch.Generator().manual_seed(5)) * 6).to(DEV)
    plan = P.Plan(n, B, torch.complex64)
    o = P.optics(area, wl, z=dz)
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = torch.empty_like(u)
    planes = [u] + [torch.empty_like(u) for _ in range(K - 1)]
    parr = (C.c_void_p * K)(None, *[p.data_ptr() for p in planes[1:]])
    L.check(lib.lp2d_plan_propagate_stack(plan._handle(), C.c_void_p(u.data_ptr()),
                                          C.c_void_p(out.data_ptr()),
                                          C.c_void_p(scr.data_ptr()), K, C.byref(o), parr, s))
    ref = P.propagate_stack(u, scr, dz, area, wl)
    assert torch.equal(out, ref)
    g = _rand((B, n, n), torch.complex64, 22)
    gin = torch.empty_like(u)
    gs = torch.empty((K, n, n), dtype=torch.float32, device=DEV)
    need = plan.workspace_size(L.LP2D_OP_STACK_ADJOINT)
    assert need == lib.lp2d_propagate_stack_adjoint_ws_size(n, B, L.LP2D_C64)
    work = plan.workspace(L.LP2D_OP_STACK_ADJOINT)
    pin = (C.c_void_p * K)(*[p.data_ptr() for p in planes])
    L.check(lib.lp2d_plan_propagate_stack_adjoint(
        plan._handle(), C.c_void_p(g.data_ptr()), pin, C.c_void_p(scr.data_ptr()), K,
        C.byref(o), C.c_void_p(gin.data_ptr()), C.c_void_p(gs.data_ptr()), 0,
        C.c_void_p(work.data_ptr()) if work is not None else None, need, s))
    gin2, gs2 = P.propagate_stack_adjoint(g, planes, scr, dz, area, wl)
    assert torch.equal(gin, gin2) and torch.equal(gs, gs2)
    plan.close()


def test_plan_config2_speckle_vs_oracle():
    """Configuration 2's first step through a plan: the speckle field (seed 0) with the
    NA 0.2 pupil, against the oracle (relative L2 1e-5)."""
    P, L, lib = _modules()
    from lightprop2d_amd import profiles as PP
    n, area, wl, z, NA = 1024, 400e-4, 632e-7, 50e-4, 0.2
    x = (np.arange(n) - n // 2) * area / n
    f = PP.random_round_hole_phase(x, x.reshape(-1, 1), 25e-4 * np.sqrt(2),
                                   rng=np.random.RandomState(0)).astype(np.complex64)
    plan = P.Plan(n, 1, torch.complex64)
    u = torch.from_numpy(f).to(DEV)
    v = plan.propagate(u, z, area, wl, aperture_radius=NA / wl)
    ref = oracle.propagate_field(f, area, wl, z, aperture_radius=NA / wl)
    assert rel_l2(v.cpu().numpy(), ref) < 1e-5
    plan.close()


def test_plan_calls_under_graph_capture():
    """A plan's propagate and adjoint (caller workspace) captured into one HIP graph:
    every replay, with other work and clears between replays, equals the eager
    results bit for bit."""
    P, L, lib = _modules()
    n, B, area, wl, z = 1024, 4, 400e-4, 632e-7, 50e-4
    ap = 0.3 / (area / n)
    u = _rand((B, n, n), torch.complex64, 31)
    g = _rand((B, n, n), torch.complex64, 32)
    ph = (torch.rand((n, n), generator=torch.Generator().manual_seed(6)) * 6).to(DEV)
    plan = P.Plan(n, B, torch.complex64)
    work = plan.workspace(L.LP2D_OP_ADJOINT)
    ref_v = plan.propagate(u, z, area, wl, phase=ph, aperture_radius=ap)
    ref_g, ref_p = plan.propagate_adjoint(g, u, z, area, wl, phase=ph, aperture_radius=ap,
                                          work=work)
    out, gin = torch.empty_like(u), torch.empty_like(u)
    graph = torch.cuda.CUDAGraph()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        with torch.cuda.graph(graph, stream=s):
            plan.propagate(u, z, area, wl, phase=ph, aperture_radius=ap, out=out)
            _, gp = plan.propagate_adjoint(g, u, z, area, wl, phase=ph, aperture_radius=ap,
                                           out=gin, work=work)
    torch.cuda.current_stream().wait_stream(s)
    for k in range(3):
        out.zero_()
        gin.zero_()
        gp.zero_()
        _ = P.propagate(u, 2 * z, area, wl)          # other work between the replays
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref_v), k
        assert torch.equal(gin, ref_g), k
        assert torch.equal(gp, ref_p), k
    plan.close()
