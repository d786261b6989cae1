"""Benchmark: batched field-propagations/s at 1024^2 complex64 on one MI355X.

A step is one pass of the hot path over one batch of synthetic fields resident in
HBM: B independent 1024x1024 complex64 fields, each propagated field -> field,
u_out = F^-1[H(z) F[u_in]] (forward 2-D FFT, transfer function generated on chip,
inverse FFT).  The forward+adjoint leg times the differentiable step used for
phase-mask optimisation: forward with a phase screen, then the adjoint with the
phase-gradient epilogue.

Prints ONE JSON line (rank 0, N = 1).  Per-kernel times come from HIP events recorded
on the launch stream by liblp2d's tracer; the CPU baseline runs the oracle (a numpy
restatement of the reference's numpy path) on this host, bounded in time.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0
METRIC = ("field-propagations/s at 1024² c64 (fwd & fwd+adjoint), "
          "% HBM roofline vs CPU")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=200)
    p.add_argument("--warmup", type=int, default=20)
    p.add_argument("--batch", type=int, default=32)
    p.add_argument("--n", type=int, default=1024)
    p.add_argument("--cpu-seconds", type=float, default=12.0,
                   help="time budget of the CPU baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--autotune", action="store_true", help="pick kernel variants first")
    p.add_argument("--no-extra", action="store_true",
                   help="skip the extra workloads (uneven batch, speckle chain, large "
                        "fields, generators, resampling, mode bases)")
    p.add_argument("--leg", choices=sorted(LEGS),
                   help="run only this extra workload and print its JSON (for profiling)")
    return p.parse_args()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:  # pragma: no cover
        pass
    return "unknown"


def _host_threads():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


def _time_cpu(fn, budget_s, max_iter):
    """Warm-up, then at least 5 timed calls until budget_s: (mean rate, best-of-5 rate,
    calls, seconds)."""
    fn()
    times = []
    t0 = time.perf_counter()
    while len(times) < 5 or (time.perf_counter() - t0 < budget_s and len(times) < max_iter):
        t1 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t1)
    el = sum(times)
    return len(times) / el, 1.0 / min(times[:5]), len(times), el


def cpu_baseline(n, budget_s, workers=None):
    """The oracle's restatement of Beam2D.propagate (numpy.fft on one thread: the
    reference's own fallback, beam2d.py:31-34, 217, 229; or scipy.fft on `workers`
    threads standing in for its pyfftw threads=-1 branch, beam2d.py:27-29) on a field
    of a geometry built once, H rebuilt per call as the reference does.  The complex64
    state (the GPU's arithmetic type) is `value`; complex128 (the reference default)
    is reported beside it."""
    import oracle.angular_spectrum as OA
    rs = np.random.RandomState(0)
    area, wl, z = 400e-4, 632e-7, 50e-4
    u = (rs.standard_normal((n, n)) + 1j * rs.standard_normal((n, n))).astype(np.complex64)
    g = OA.construct_grids(area, n, wl)
    res = {}
    for name, dt in (("c64", np.complex64), ("c128", np.complex128)):
        res[name] = _time_cpu(lambda: OA.propagate_field_prepared(u, g, z, dt, workers),
                              budget_s / 2, 400)
    mean64, best64, k64, el64 = res["c64"]
    mean128, best128, k128, el128 = res["c128"]
    lib = "numpy.fft pocketfft, one thread" if workers is None else f"scipy.fft, {workers} workers"
    return {"value": mean64, "unit": "field-propagations/s",
            "cores": 1 if workers is None else workers, "kind": "port",
            "sample": f"{k64} x {n}x{n} complex64 field-propagations ({lib}; grids built "
                      f"once, H rebuilt per call) in {el64:.1f} s; complex128: {k128} in "
                      f"{el128:.1f} s",
            "best_of_5": best64, "value_c128": mean128, "best_of_5_c128": best128,
            "cpu_model": _cpu_model(), "host_threads": _host_threads(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS", "")}


TRAFFIC_FILE = "hbm_traffic_propagate_1024_c64.md"


def profiled_traffic(n, batch):
    """HBM bytes per launch of each sweep of the forward step, keyed by the tracer's
    kernel names, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary
    (profiles/hbm_traffic_propagate_1024_c64.md, lines
    ``sweep_traffic_bytes[<name>]: <bytes>``), or {} for another workload."""
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", TRAFFIC_FILE)
    out = {}
    try:
        text = open(p).read()
    except OSError:
        return out
    kv = {}
    for line in text.splitlines():
        m = re.match(r"^sweep_traffic_bytes\[(.+)\]: (\d+)$", line.strip())
        if m:
            out[m.group(1)] = int(m.group(2))
        elif ": " in line:
            k, v = line.split(": ", 1)
            kv[k.strip()] = v.strip()
    if kv.get("workload_n") != str(n) or kv.get("workload_batch") != str(batch):
        return {}
    return out


def trace_stats(lib, L):
    import ctypes as C
    arr = (L.lp2d_trace_stat * 64)()
    cnt = C.c_int()
    L.check(lib.lp2d_trace_collect(arr, 64, C.byref(cnt)), "trace_collect")
    return {arr[i].name.decode(): (arr[i].launches, arr[i].total_ms) for i in range(cnt.value)}


def extra_legs(torch, P, L, lib, dev):
    """The other workloads, timed the same way (inputs resident, HIP-event kernel
    times): the uneven batch of 4096 fields (one launch per size bucket and sweep over
    per-field descriptor tables, captured in a HIP graph), the 10-step speckle chain at
    1024^2, large fields, the profile generators, resampling and mode bases.  alg_bytes
    counts one read and one write of every field (16 B per complex64 element), the
    least any propagation can move."""
    out = {"uneven_batch": uneven_leg(torch, P, L, lib, dev),
           "speckle_chain": speckle_leg(torch, P, dev),
           "beam2d_chain": beam2d_leg(torch, P, dev),
           "published": published_leg(torch, dev),
           "phase_mask_design": phase_mask_leg(torch, dev),
           "gaussian_256_c128": gaussian_leg(torch, dev),
           "stack": stack_leg(torch, P, L, lib, dev)}
    out["large_fields"] = large_field_leg(torch, P, L, lib, dev)
    out["generators"] = generator_leg(torch, dev)
    out["resample"] = resample_leg(torch, P, dev)
    out["modes"] = modes_leg(torch, P, dev)
    return out


def uneven_leg(torch, P, L, lib, dev):
    """Configuration 3's uneven batch: 4096 fields of three sizes, size-bucketed launches
    over per-field descriptor tables, replayed from a HIP graph."""
    # uneven batch: 2048 x 128^2 + 1536 x 256^2 + 512 x 512^2, complex64, z ~ U[10 um, 1 mm]
    g = torch.Generator(device="cpu").manual_seed(1)
    sizes = [128] * 2048 + [256] * 1536 + [512] * 512
    zs = (torch.rand(len(sizes), generator=g, dtype=torch.float64) * (0.1 - 10e-4) + 10e-4).tolist()
    pools = {}
    for nn in (128, 256, 512):
        cnt = sizes.count(nn)
        pools[nn] = torch.randn((cnt, nn, nn), dtype=torch.complex64, generator=g).to(dev)
    idx = {128: 0, 256: 0, 512: 0}
    srcs = []
    for nn in sizes:
        srcs.append(pools[nn][idx[nn]])
        idx[nn] += 1
    ub = P.UnevenBatch(srcs, srcs, zs, 200e-4, 532e-7)   # in place
    graph = ub.capture_graph()
    for _ in range(2):
        graph.replay()
    torch.cuda.synchronize()
    K = 10
    L.check(lib.lp2d_trace_enable(1))
    ub.run()
    torch.cuda.synchronize()
    L.check(lib.lp2d_trace_enable(0))
    sweeps = trace_stats(lib, L)
    t0 = time.perf_counter()
    for _ in range(K):
        graph.replay()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / K
    alg = 16.0 * sum(nn * nn for nn in sizes)
    res = {
        "workload": "4096 fields (2048x128^2 + 1536x256^2 + 512x512^2) complex64, per-field z, "
                    "area 200 um, 532 nm; size-bucketed launches over descriptor tables, "
                    "HIP graph replay",
        "value": round(len(sizes) / el, 1), "unit": "field-propagations/s",
        "ms_per_call": round(el * 1e3, 4), "alg_bytes": int(alg),
        "end_to_end_GBs": round(alg / el / 1e9, 1),
        "end_to_end_frac": round(alg / el / 1e9 / HBM_PEAK_GBS, 4),
        "sweeps_us": {k: round(ms * 1e3 / c, 1) for k, (c, ms) in sweeps.items()}}
    del ub, graph, srcs, pools
    return res


def speckle_leg(torch, P, dev):
    """The 10-step speckle chain at 1024^2 (8 fields, NA 0.2 aperture, in place)."""
    # speckle chain: 1024^2, 10 x [propagate 50 um, NA 0.2 aperture], field to field
    n, area, wl = 1024, 400e-4, 632e-7
    g = torch.Generator(device="cpu").manual_seed(2)
    u = torch.randn((8, n, n), dtype=torch.complex64, generator=g).to(dev)
    for _ in range(2):
        P.propagate(u, 50e-4, area, wl, aperture_radius=0.2 / wl, out=u)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        for _ in range(10):
            P.propagate(u, 50e-4, area, wl, aperture_radius=0.2 / wl, out=u)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 5
    res = {
        "workload": "8 speckle fields 1024^2 complex64, 10 x [propagate 50 um, NA 0.2 aperture], "
                    "in place",
        "value": round(8 * 10 / el, 1), "unit": "field-propagations/s",
        "ms_per_chain": round(el * 1e3, 4)}
    del u
    return res


def large_field_leg(torch, P, L, lib, dev):
    """Field-to-field propagation of 8 fields of 4096^2 complex64 (the field size of the
    multi-plane stack: area 2 mm, 800 nm, 2 mm steps), 1 GiB per sweep: whole-call rate
    and the per-sweep kernel times from the tracer."""
    n, B, area, wl, z = 4096, 8, 2000e-4, 800e-7, 2000e-4
    g = torch.Generator(device="cpu").manual_seed(5)
    u = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    out = torch.empty_like(u)
    for _ in range(2):
        P.propagate(u, z, area, wl, out=out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        P.propagate(u, z, area, wl, out=out)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    L.check(lib.lp2d_trace_enable(1))
    P.propagate(u, z, area, wl, out=out)
    torch.cuda.synchronize()
    L.check(lib.lp2d_trace_enable(0))
    sweeps = trace_stats(lib, L)
    alg = 16.0 * B * n * n
    return {"workload": "8 fields 4096^2 complex64, area 2 mm, 800 nm, z 2 mm, field to field",
            "value": round(B / el, 1), "unit": "field-propagations/s",
            "ms_per_call": round(el * 1e3, 4),
            "sweeps_us": {k: round(ms * 1e3 / c, 1) for k, (c, ms) in sweeps.items()},
            "end_to_end_GBs": round(alg / el / 1e9, 1),
            "end_to_end_frac": round(alg / el / 1e9 / HBM_PEAK_GBS, 4)}


def stack_leg(torch, P, L, lib, dev):
    """Configuration 5's forward: 8 fields of 4096^2 complex64 through 5 shared phase
    screens 2 mm apart (800 nm, area 2 mm), as one fused stack (lp2d_propagate_stack:
    the inverse rows of a plane, the next screen and the next forward rows in one
    sweep) against five separate propagations; its adjoint as one fused stack adjoint
    (lp2d_propagate_stack_adjoint) against five separate adjoints; one checkpointed
    gradient step through torch.autograd; per-sweep kernel times of the fused forms
    from the tracer."""
    n, B, K, area, wl, dz = 4096, 8, 5, 2000e-4, 800e-7, 2000e-4
    g = torch.Generator(device="cpu").manual_seed(6)
    u = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    scr = (torch.randn((K, n, n), generator=g)).to(torch.float32).to(dev)
    out = torch.empty_like(u)
    tmp = torch.empty_like(u)

    def fused():
        P.propagate_stack(u, scr, dz, area, wl, out=out)

    def per_plane():
        P.propagate(u, dz, area, wl, phase=scr[0], out=tmp)
        for k in range(1, K):
            P.propagate(tmp, dz, area, wl, phase=scr[k], out=tmp)

    res = {}
    for name, fn in (("fused", fused), ("per_plane", per_plane)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / 5
        res[name] = {"ms_per_stack": round(el * 1e3, 3), "ms_per_plane": round(el * 1e3 / K, 3),
                     "plane_GBs": round(16.0 * B * n * n * K / el / 1e9, 1)}
    L.check(lib.lp2d_trace_enable(1))
    fused()
    torch.cuda.synchronize()
    L.check(lib.lp2d_trace_enable(0))
    sw = trace_stats(lib, L)
    res["fused_sweeps_us"] = {k: round(ms * 1e3 / c, 1) for k, (c, ms) in sw.items()}
    res["speedup"] = round(res["per_plane"]["ms_per_stack"] / res["fused"]["ms_per_stack"], 3)
    # the adjoint, given the plane inputs: one fused stack adjoint (2K + 1 sweeps, screen
    # gradients summed over the batch) against K separate adjoints
    planes = [u] + [torch.empty_like(u) for _ in range(K - 1)]
    P.propagate_stack(u, scr, dz, area, wl, out=out, planes=[None] + planes[1:])
    gout = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    gin = torch.empty_like(u)
    gs = torch.empty((K, n, n), dtype=torch.float32, device=dev)

    def fused_adj():
        P.propagate_stack_adjoint(gout, planes, scr, dz, area, wl, out=gin, grad_screens=gs)

    def per_plane_adj():
        gg = gout
        for k in range(K - 1, -1, -1):
            gg, gk = P.propagate_adjoint(gg, planes[k], dz, area, wl, phase=scr[k])
            gs[k].copy_(gk)

    adj = {}
    for name, fn in (("fused", fused_adj), ("per_plane", per_plane_adj)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / 5
        adj[name] = {"ms_per_stack": round(el * 1e3, 3), "ms_per_plane": round(el * 1e3 / K, 3)}
    L.check(lib.lp2d_trace_enable(1))
    fused_adj()
    torch.cuda.synchronize()
    L.check(lib.lp2d_trace_enable(0))
    sw = trace_stats(lib, L)
    adj["fused_sweeps_us"] = {k: round(ms * 1e3 / c, 1) for k, (c, ms) in sw.items()}
    adj["speedup"] = round(adj["per_plane"]["ms_per_stack"] / adj["fused"]["ms_per_stack"], 3)
    res["adjoint"] = adj
    del planes, gout, gin, gs
    # one gradient step of the checkpointed stack (every plane input kept): fused
    # forward storing the planes, fused adjoint with the screen gradients
    from lightprop2d_amd import autograd as AG
    ur = u.detach()
    sr = scr.clone().requires_grad_(True)
    w = torch.rand((n, n), generator=g).to(dev)

    def step():
        sr.grad = None
        v = AG.propagate_stack(ur, sr, dz, area, wl)
        # |v|^2 through the real view: its backward is a view as well (no copies of the
        # 1 GiB field into real and imaginary parts)
        (w * torch.view_as_real(v).square().sum(-1)).sum().backward()

    step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        step()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / 3
    res["grad_step"] = {"ms": round(el * 1e3, 3), "ms_per_plane": round(el * 1e3 / K, 3)}
    res["workload"] = ("8 x 4096^2 complex64, 5 shared phase screens, 2 mm spacing, 800 nm, "
                       "area 2 mm: fused stack vs 5 propagations, fused stack adjoint vs 5 "
                       "adjoints, and one checkpointed gradient step (loss sum(w |u_5|^2))")
    del u, scr, out, tmp, sr, w
    return res


def gaussian_leg(torch, dev):
    """Configuration 1 through the drop-in Beam2D on the device: 256^2 complex128 (the
    reference's default precision), area 200 um, 632 nm, Gaussian beam of radius
    25 um, one propagate(100 um), the intensity read on the device (the reference's
    tests/gaussbeam_diffraction.py sequence).  Timed per whole sequence, construction
    included, synchronised."""
    from lightprop2d_amd import Beam2D, gaussian_beam, um, nm

    def run():
        b = Beam2D(200 * um, 256, 632 * nm, init_field_gen=gaussian_beam,
                   init_gen_args=(1, 25 * um), numpy_output=False)
        b.propagate(100 * um)
        return b.iprofile

    for _ in range(3):
        run()
    torch.cuda.synchronize()
    reps = 100
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
        torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    return {"us_per_sequence": round(el * 1e6, 1), "sequences_per_s": round(1 / el, 1),
            "workload": "configuration 1: Beam2D 256^2 complex128, Gaussian 25 um, "
                        "propagate 100 um, iprofile on the device; construction included"}


def phase_mask_leg(torch, dev):
    """Configuration 4: a 512^2 float32 phase mask (seeded U[0, 2pi)) on a Gaussian
    input (rho0 50 um, area 500 um, 1064 nm): propagate 1 mm through the mask, lens
    f = 5 mm (reference convention) and propagate 5 mm, MSE of the intensity against a
    ring (radius 30 um, width 5 um); 200 Adam steps (lr 0.1) through the hand-written
    adjoint (lightprop2d_amd.autograd).  One field per step, so each step is bound by
    launch latency; reported as ms per step (forward, loss, backward, optimiser)."""
    from lightprop2d_amd import autograd as AG
    from lightprop2d_amd import profiles as PP
    n, area, wl, R = 512, 500e-4, 1064e-7, 50e-4
    x = (np.arange(n) - n // 2) * area / n
    u0 = PP.gaussian_beam(x.copy(), x.reshape(-1, 1).copy(), 1.0, R)
    r = np.sqrt(x ** 2 + x.reshape(-1, 1) ** 2)
    target = torch.from_numpy(np.exp(-(r - 30e-4) ** 2 / (2 * (5e-4) ** 2)).astype(np.float32)).to(dev)
    u0_t = torch.from_numpy(np.asarray(u0).astype(np.complex64)).to(dev)
    phi = torch.from_numpy(np.random.RandomState(0).uniform(0, 2 * np.pi, (n, n))
                           .astype(np.float32)).to(dev).requires_grad_(True)
    opt = torch.optim.Adam([phi], lr=0.1)

    def step():
        opt.zero_grad()
        u1 = AG.propagate(u0_t, 1e-1, area, wl, phase=phi)
        u2 = AG.propagate(u1, 5e-1, area, wl, lens_f=5e-1)
        loss = ((u2.real ** 2 + u2.imag ** 2 - target) ** 2).mean()
        loss.backward()
        opt.step()
        return loss.detach()

    first = step()
    for _ in range(4):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = [step() for _ in range(200)]
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    # the same step captured once into a HIP graph and replayed (static tensors, Adam
    # with its step count on the device; the library allocates nothing after its
    # initialisation, the autograd buffers come from the graph's pool)
    phg = torch.from_numpy(np.random.RandomState(0).uniform(0, 2 * np.pi, (n, n))
                           .astype(np.float32)).to(dev).requires_grad_(True)
    optg = torch.optim.Adam([phg], lr=0.1, capturable=True)

    def body():
        u1 = AG.propagate(u0_t, 1e-1, area, wl, phase=phg)
        u2 = AG.propagate(u1, 5e-1, area, wl, lens_f=5e-1)
        loss = ((u2.real ** 2 + u2.imag ** 2 - target) ** 2).mean()
        loss.backward()
        optg.step()
        return loss

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            optg.zero_grad(set_to_none=True)
            body()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    optg.zero_grad(set_to_none=True)
    with torch.cuda.graph(graph):
        static_loss = body()
    for _ in range(2):
        graph.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(200):
        graph.replay()
    torch.cuda.synchronize()
    elg = time.perf_counter() - t0
    return {"ms_per_step": round(el / 200 * 1e3, 4), "steps_per_s": round(200 / el, 1),
            "ms_per_step_graph": round(elg / 200 * 1e3, 4),
            "steps_per_s_graph": round(200 / elg, 1),
            "loss_first": float(first), "loss_last": float(losses[-1]),
            "loss_last_graph": float(static_loss),
            "workload": "configuration 4: 512^2 phase mask, Gaussian 50 um, area 500 um, "
                        "1064 nm, P(1 mm) -> lens 5 mm -> P(5 mm), MSE to a ring, Adam "
                        "(lr 0.1), 200 timed steps after 5 warm-up steps; eager, and the "
                        "step captured in one HIP graph"}


def published_leg(torch, dev):
    """The reference's own published timings (the %timeit transcripts of
    Beam2D._update_obj's docstring, beam2d.py:236-269, 1024^2 complex128, the default
    precision; old micrometre API Beam2D(200, 1024, 0.632, ...) = area 200 um, wl
    632 nm, Gaussian rho0 50 um), repeated through the drop-in Beam2D: construction
    from gaussian_beam, from an array, from an array and a spectrum, and _update_obj
    with and without a spectrum.  Each call is timed from the host to the end of its
    GPU work (synchronised, `us`) and, as %timeit on CuPy measures, without waiting
    (`async_us`); both are means over repeated calls.  The published figures are from
    an unstated NVIDIA GPU (CuPy 11.2) and an unstated CPU."""
    from lightprop2d_amd import Beam2D, gaussian_beam, um, nm
    area, n, wl = 200 * um, 1024, 632 * nm
    b = Beam2D(area, n, wl, init_field_gen=gaussian_beam, init_gen_args=(1, 50 * um))
    a = Beam2D(area, n, wl, init_field=b.field)
    ops = {
        "construct_from_gaussian_beam": (
            lambda: Beam2D(area, n, wl, init_field_gen=gaussian_beam, init_gen_args=(1, 50 * um)),
            2750.0, 81000.0),
        "construct_from_array": (lambda: Beam2D(area, n, wl, init_field=b.field), 2160.0, 55800.0),
        "update_obj_field": (lambda: a._update_obj(b.field), 66.6, 36400.0),
        "construct_with_spectrum": (
            lambda: Beam2D(area, n, wl, init_field=b.field, init_spectrum=b.spectrum),
            1100.0, 17200.0),
        "update_obj_field_spectrum": (lambda: a._update_obj(b.field, spectrum=b.spectrum),
                                      1.39, 1.12),
    }
    out = {}
    for name, (fn, pub_gpu, pub_cpu) in ops.items():
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        reps = 50
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
            torch.cuda.synchronize()
        sync_us = (time.perf_counter() - t0) / reps * 1e6
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        async_us = (time.perf_counter() - t0) / reps * 1e6
        torch.cuda.synchronize()
        out[name] = {"us": round(sync_us, 2), "async_us": round(async_us, 2),
                     "published_gpu_us": pub_gpu, "published_cpu_us": pub_cpu,
                     "speedup_vs_published_gpu": round(pub_gpu / sync_us, 2),
                     "speedup_vs_published_gpu_async": round(pub_gpu / async_us, 2)}
    out["workload"] = ("the reference's published Beam2D timings at 1024^2 complex128 "
                       "(beam2d.py:236-269), through the drop-in Beam2D")
    return out


def beam2d_leg(torch, P, dev):
    """Configuration 2 through the drop-in Beam2D: 1024^2 complex64 speckle field,
    10 x [propagate(50 um), spectral_filter(NA 0.2 pupil array)], then iprofile on the
    device.  Beam2D keeps the field lazily: each propagate folds its transfer function
    into the next spectral multiply, each spectral_filter is one elementwise sweep over
    the spectrum (16 n^2 B with the pupil's 1 n^2 B), and the one inverse FFT runs at
    the iprofile read.  Reported per Beam2D call (21 calls per chain)."""
    from lightprop2d_amd import Beam2D
    from lightprop2d_amd import profiles as PP
    n, area, wl, NA = 1024, 400e-4, 632e-7, 0.2
    x = (np.arange(n) - n // 2) * area / n
    f = PP.random_round_hole_phase(x, x.reshape(-1, 1), 25e-4 * np.sqrt(2),
                                   rng=np.random.RandomState(0))
    kx = np.fft.fftfreq(n, area / n)
    pupil = torch.from_numpy((kx ** 2 + kx.reshape(-1, 1) ** 2) <= (NA / wl) ** 2).to(dev)
    b = Beam2D(area, n, wl, init_field=f, complex_bits=64, numpy_output=False)

    def chain():
        for _ in range(10):
            b.propagate(50e-4)
            b.spectral_filter(f_init=pupil)
        return b.iprofile

    chain()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    reps = 10
    for _ in range(reps):
        chain()
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / reps
    return {"workload": "Beam2D 1024^2 complex64: 10 x [propagate 50 um, spectral_filter(NA 0.2 "
                        "pupil)] + iprofile on the device",
            "ms_per_chain": round(el * 1e3, 4), "value": round(21 / el, 1),
            "unit": "Beam2D calls/s",
            "bytes_per_filter_call": 17 * n * n,
            "sweeps_per_chain": "10 elementwise + 2 (one inverse FFT) + 1 (intensity)"}


MFMA_FP64_PEAK_TFS = 78.6  # MI355X dense FP64 matrix peak


def modes_leg(torch, P, dev):
    """Mode-basis projection and reconstruction: 64 modes of 256^2 points against 32
    complex64 fields of 512^2 (central 256^2 window).  Algorithmic bytes: the mode
    matrix once plus the field windows (projection) or the fields (reconstruction)."""
    M, nb, n, B = 64, 256, 512, 32
    g = torch.Generator(device="cpu").manual_seed(4)
    modes = torch.randn((M, nb * nb), dtype=torch.complex64, generator=g).to(dev)
    f = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    c = torch.randn((B, M), dtype=torch.complex128, generator=g).to(dev)
    res = {}
    f1 = f[0].contiguous()
    for name, call, alg in (("project", lambda: P.modes_project(f, modes),
                             8 * M * nb * nb + 8 * B * nb * nb + 16 * B * M),
                            ("project_one_field", lambda: P.modes_project(f1, modes),
                             8 * M * nb * nb + 8 * nb * nb + 16 * M),
                            ("construct", lambda: P.modes_construct(modes, c, n, torch.complex64,
                                                                   dev),
                             8 * M * nb * nb + 8 * B * n * n + 16 * B * M)):
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            call()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / 20
        res[name] = {"us_per_call": round(el * 1e6, 1), "GBs": round(alg / el / 1e9, 1),
                     "hbm_frac": round(alg / el / 1e9 / HBM_PEAK_GBS, 4)}
    # the batch projection runs on the FP64 matrix cores: 4 real products of
    # (M x P) by (P x B) per call, 2 flops per multiply-add
    fl = 8.0 * M * B * nb * nb
    tfs = fl / (res["project"]["us_per_call"] * 1e-6) / 1e12
    res["project"].update({"path": "FP64 MFMA (v_mfma_f64_16x16x4f64)",
                           "mfma_TFLOPs": round(tfs, 2),
                           "mfma_frac": round(tfs / MFMA_FP64_PEAK_TFS, 4)})
    mh = modes[:, :].cpu().numpy().astype(np.complex128)
    fh = f[0, 128:384, 128:384].reshape(-1).cpu().numpy().astype(np.complex128)
    t0 = time.perf_counter()
    mh @ fh
    res["numpy_one_field_project_us"] = round((time.perf_counter() - t0) * 1e6, 1)
    res["workload"] = ("64 modes x 256^2, 32 complex64 fields of 512^2 (central window); "
                       "coefficients complex128")
    return res


def resample_leg(torch, P, dev):
    """Beam2D.expand / coarse / crop / lens_image kernels on 8 fields of complex64:
    1024^2 -> 2048^2 (edge replication), 2048^2 -> 1024^2 (2x2 block mean), crop to
    half the area at 1024 points (bilinear 2x zoom of the central 512^2), lens_image at
    magnification 1/2 (l1 = 1, l2 = 2: the central 512^2 read, chirped).  Algorithmic
    bytes = one read of the samples used plus one write of the output; the numpy
    equivalent of one field's expand + coarse is timed beside it."""
    B, n = 8, 1024
    area, wl = 400e-4, 632e-7
    g = torch.Generator(device="cpu").manual_seed(3)
    u = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    big = P.expand(u, 2 * n)
    res = {}
    quarter = (n // 2) ** 2
    for name, call, alg in (("expand_1024_to_2048", lambda: P.expand(u, 2 * n),
                             8 * B * (n * n + 4 * n * n)),
                            ("coarse_2048_to_1024", lambda: P.coarse(big, 2),
                             8 * B * (4 * n * n + n * n)),
                            ("crop_half_area_1024", lambda: P.crop(u, area, area / 2),
                             8 * B * (quarter + n * n)),
                            ("lens_image_m0.5_1024", lambda: P.lens_image(u, area, wl, 1.0, 2.0),
                             8 * B * (quarter + n * n))):
        call()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(20):
            call()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / 20
        res[name] = {"us_per_call": round(el * 1e6, 1), "GBs": round(alg / el / 1e9, 1),
                     "hbm_frac": round(alg / el / 1e9 / HBM_PEAK_GBS, 4)}
    f = np.random.RandomState(0).standard_normal((n, n)).astype(np.complex64)
    t0 = time.perf_counter()
    fb = np.pad(f, n // 2, mode="edge")
    fb.reshape((n, 2, n, 2)).mean((1, 3))
    res["numpy_one_field_us"] = round((time.perf_counter() - t0) * 1e6, 1)
    res["workload"] = ("8 fields 1024^2 complex64: expand x2, coarse(2), crop(area/2), "
                       "lens_image(l1=1, l2=2); device-resident")
    return res


def generator_leg(torch, dev):
    """Profile generators at 1024^2 on the Beam2D grid (area 400 um): the device path
    (lp2d_profile; the random kind uploads only its host draw) against the same
    generator evaluated with numpy on the host and copied to the device."""
    from lightprop2d_amd import profiles as PP
    n, dL = 1024, 400e-4 / 1024
    xh = np.arange(-n // 2, n // 2, 1) * dL
    yh = xh.reshape((-1, 1))
    xd = torch.from_numpy(xh).to(dev)
    yd = xd.reshape((-1, 1))
    res = {}
    for name, args in (("gaussian_beam", (1.0, 25e-4)), ("round_hole", (25e-4,)),
                       ("random_round_hole_phase", (25e-4,))):
        fn = getattr(PP, name)

        def device_call():
            return fn(xd.clone(), yd.clone(), *args)

        def host_call():
            return torch.from_numpy(np.asarray(fn(xh.copy(), yh.copy(), *args))).to(dev)

        row = {}
        for label, call in (("device_us", device_call), ("host_numpy_plus_copy_us", host_call)):
            call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(10):
                call()
            torch.cuda.synchronize()
            row[label] = round((time.perf_counter() - t0) / 10 * 1e6, 1)
        res[name] = row
    res["workload"] = ("one 1024x1024 generator call on the Beam2D grid (area 400 um), "
                       "result resident on the device; random draws from numpy's RNG on "
                       "the host in both paths")
    return res


# extra workloads that run alone with --leg (uniform signature)
LEGS = {
    "uneven_batch": lambda torch, P, L, lib, dev: uneven_leg(torch, P, L, lib, dev),
    "speckle_chain": lambda torch, P, L, lib, dev: speckle_leg(torch, P, dev),
    "beam2d_chain": lambda torch, P, L, lib, dev: beam2d_leg(torch, P, dev),
    "published": lambda torch, P, L, lib, dev: published_leg(torch, dev),
    "phase_mask_design": lambda torch, P, L, lib, dev: phase_mask_leg(torch, dev),
    "gaussian_256_c128": lambda torch, P, L, lib, dev: gaussian_leg(torch, dev),
    "stack": lambda torch, P, L, lib, dev: stack_leg(torch, P, L, lib, dev),
    "large_fields": lambda torch, P, L, lib, dev: large_field_leg(torch, P, L, lib, dev),
    "resample": lambda torch, P, L, lib, dev: resample_leg(torch, P, dev),
    "modes": lambda torch, P, L, lib, dev: modes_leg(torch, P, dev),
    "generators": lambda torch, P, L, lib, dev: generator_leg(torch, dev),
}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1 "
                                   "(batches of independent fields shard by field with no "
                                   "communication)"}))
        sys.exit(2)
    import torch
    from lightprop2d_amd import _lib as L
    from lightprop2d_amd import propagation as P

    lib = L.load()
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.leg:
        # a plain 256 MiB device copy first: the calibration dispatch for rocprof's
        # FETCH_SIZE / WRITE_SIZE counters in a profiled run of one leg
        a = torch.empty(1 << 25, dtype=torch.complex64, device=dev)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
        print(json.dumps({args.leg: LEGS[args.leg](torch, P, L, lib, dev)}), flush=True)
        return
    n, B = args.n, args.batch
    area, wl, z = 400e-4, 632e-7, 50e-4
    if args.autotune:
        from lightprop2d_amd import autotune
        autotune.tune(n, "c64", batch=B)

    g = torch.Generator(device="cpu").manual_seed(0)
    u = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    out = torch.empty_like(u)
    phi = (torch.rand((n, n), generator=g) * 2 * np.pi).to(torch.float32).to(dev)
    gout = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(dev)
    gin = torch.empty_like(u)

    def fwd_step():
        P.propagate(u, z, area, wl, out=out)

    def fwd_adj_step():
        P.propagate(u, z, area, wl, phase=phi, out=out)
        P.propagate_adjoint(gout, u, z, area, wl, phase=phi, out=gin)

    def timed(step, K, W):
        """K steps after W warm-up ones: (wall seconds, GPU seconds between two HIP
        events recorded on the launch stream around the K steps, per-kernel times).
        The per-kernel times come from a separate traced pass of min(K, 10) steps:
        events around every launch perturb back-to-back kernels, so they stay out of
        the timed region."""
        for _ in range(W):
            step()
        torch.cuda.synchronize()
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()  # torch's current stream: the stream every lp2d call launches on
        for _ in range(K):
            step()
        e1.record()
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        gpu = e0.elapsed_time(e1) / 1e3
        L.check(lib.lp2d_trace_enable(1))
        for _ in range(min(K, 10)):
            step()
        torch.cuda.synchronize()
        L.check(lib.lp2d_trace_enable(0))
        return el, gpu, trace_stats(lib, L)

    # plain device copy of the same bytes: achievable-bandwidth reference, and the
    # calibration dispatch for rocprof's FETCH_SIZE / WRITE_SIZE counters
    for _ in range(3):
        out.copy_(u)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(10):
        out.copy_(u)
    ev1.record()
    torch.cuda.synchronize()
    copy_s = ev0.elapsed_time(ev1) / 1e3 / 10
    copy_gbs = 2 * u.numel() * 8 / copy_s / 1e9

    el, gpu_s, kt = timed(fwd_step, args.steps, args.warmup)
    props = B * args.steps
    value = props / el
    K2 = max(1, args.steps // 2)
    el2, gpu2_s, kt2 = timed(fwd_adj_step, K2, 2)
    fa_rate = B * K2 / el2

    # The unit of work is one field-propagation, spatial field in HBM -> spatial field
    # in HBM: B_alg = 16 n^2 bytes (the field read once and written once, complex64).
    # One step launches the three sweeps (forward rows, columns with H, inverse rows)
    # over the B fields, so the step's algorithmic bytes are 16 n^2 B, and the
    # roofline fraction is that over the step's GPU time (HIP events around the timed
    # steps).  Each sweep also reads and writes every field once, so a sweep alone is
    # credited with the same bytes in `sweeps` (its own rate, not the step's).
    alg_step = 16.0 * n * n * B
    achieved = alg_step * args.steps / gpu_s / 1e9
    wall_gbs = 16.0 * n * n * value / 1e9
    traffic = profiled_traffic(n, B)
    sweeps = {}
    for name, (c, ms) in kt.items():
        avg = ms / 1e3 / c
        sw = {"avg_us": round(avg * 1e6, 2), "alg_bytes": int(alg_step),
              "GBs": round(alg_step / avg / 1e9, 1),
              "frac": round(alg_step / avg / 1e9 / HBM_PEAK_GBS, 4)}
        if name in traffic:
            sw["traffic"] = traffic[name]
            sw["traffic_ratio"] = round(traffic[name] / alg_step, 3)
        sweeps[name] = sw
    step_traffic = (sum(traffic[k] for k in kt) if traffic and all(k in traffic for k in kt)
                    else None)
    kernel_sum_s = sum(ms / 1e3 / c for c, ms in kt.values())
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": step_traffic,
                "traffic_ratio": round(step_traffic / alg_step, 3) if step_traffic else None,
                "traffic_per_propagation": round(step_traffic / B) if step_traffic else None,
                "traffic_source": ("profiles/" + TRAFFIC_FILE) if step_traffic else None,
                "launch": (f"one step: the {len(kt)} sweeps of a field-to-field "
                           f"propagation over {B} fields of {n}^2 complex64"),
                "alg_bytes_per_launch": int(alg_step),
                "alg_bytes_per_propagation": int(16 * n * n),
                "step_gpu_us": round(gpu_s / args.steps * 1e6, 2),
                "kernel_sum_us": round(kernel_sum_s * 1e6, 2),
                "kernel_sum_frac": round(alg_step / kernel_sum_s / 1e9 / HBM_PEAK_GBS, 4),
                "wall_GBs": round(wall_gbs, 1),
                "wall_frac": round(wall_gbs / HBM_PEAK_GBS, 4),
                "sweeps": sweeps,
                "copy_GBs": round(copy_gbs, 1)}
    kernels = {k: {"launches": c, "avg_us": round(ms * 1e3 / c, 2)} for k, (c, ms) in kt.items()}

    extra = {} if args.no_extra else extra_legs(torch, P, L, lib, dev)

    cpu = cpu_mt = None
    if not args.no_cpu:
        cpu = cpu_baseline(n, args.cpu_seconds)
        cpu_mt = cpu_baseline(n, min(args.cpu_seconds, 6.0),
                              workers=max(1, min(_host_threads(), 16)))

    line = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "field-propagations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {"workload": f"batched {n}x{n} complex64 field propagation "
                               f"F^-1[H F u], {B} fields per step, area 400 um, "
                               f"wl 632 nm, z 50 um, inputs resident in HBM",
                   "n": n, "batch": B, "parallelism": "single GPU"},
        "fwd_adjoint": {"value": round(fa_rate, 1), "unit": "field-propagations/s",
                        "ms_per_step": round(el2 / K2 * 1e3, 4),
                        "gpu_ms_per_step": round(gpu2_s / K2 * 1e3, 4),
                        "ratio_to_fwd": round(value / fa_rate, 3),
                        "workload": "forward with phase screen + adjoint with phase gradient"},
        "roofline": roofline,
        "kernels": kernels,
        "kernels_fwd_adjoint": {k: {"launches": c, "avg_us": round(ms * 1e3 / c, 2)}
                                for k, (c, ms) in kt2.items()},
        "cpu_baseline": cpu,
        "cpu_baseline_threaded": cpu_mt,
        "workloads": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
