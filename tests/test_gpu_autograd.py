"""Hand-written adjoint behind torch.autograd.Function, checked three ways:
against torch autograd through a complex128 torch.fft restatement of the same
forward, against central finite differences on the complex128 oracle (config 4:
64 seeded pixels), and, for the checkpointed multi-plane stack (config 5), against
the oracle's reverse-mode gradient and between checkpoint policies."""
import numpy as np
import pytest
import torch

from conftest import rel_l2
import oracle
from oracle import profiles as OP
from oracle.angular_spectrum import construct_grids, transfer_function, lens_phase

pytestmark = pytest.mark.gpu
DEV = "cuda"


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    assert torch.cuda.is_available()


def torch_reference(u, phase, z, area, wl, ap=0.0, lens=0.0):
    """Same forward, written with torch.fft on the CPU in complex128 (torch autograd)."""
    n = u.shape[-1]
    g = construct_grids(area, n, wl)
    H = torch.from_numpy(transfer_function(g.Kz, z))
    v = u * torch.exp(1j * phase.to(torch.float64)) if phase is not None else u
    if lens:
        ph = torch.from_numpy(lens_phase(g.X, g.Y, g.k0, lens))
        v = v * torch.exp(1j * ph)
    S = torch.fft.fft2(v) * H
    if ap:
        S = S * torch.from_numpy(((g.Kx ** 2 + g.Ky ** 2) <= ap ** 2).astype(np.float64))
    return torch.fft.ifft2(S)


@pytest.mark.parametrize("n,dt", [(64, torch.complex128), (256, torch.complex128),
                                  (256, torch.complex64), (1024, torch.complex64)])
def test_adjoint_vs_torch_autograd_c128(n, dt):
    from lightprop2d_amd import autograd as AG
    rs = np.random.RandomState(n)
    area, wl, z, lens = 300e-4, 632e-7, 0.04, 0.9
    ap = 0.4 / (area / n)
    u_np = rs.standard_normal((2, n, n)) + 1j * rs.standard_normal((2, n, n))
    phi_np = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    w_np = rs.uniform(size=(2, n, n))
    # reference (CPU, complex128, torch autograd)
    ur = torch.from_numpy(u_np).requires_grad_(True)
    pr = torch.from_numpy(phi_np.astype(np.float64)).requires_grad_(True)
    vr = torch_reference(ur, pr, z, area, wl, ap, lens)
    Lr = (torch.from_numpy(w_np) * vr.abs() ** 2).sum()
    Lr.backward()
    # ours
    ud = torch.from_numpy(u_np).to(dt).to(DEV).requires_grad_(True)
    pd = torch.from_numpy(phi_np).to(DEV).requires_grad_(True)
    vd = AG.propagate(ud, z, area, wl, phase=pd, aperture_radius=ap, lens_f=lens)
    Ld = (torch.from_numpy(w_np).to(DEV) * vd.abs() ** 2).sum()
    Ld.backward()
    tol = 1e-9 if dt == torch.complex128 else 2e-5
    assert rel_l2(vd.detach().cpu().numpy(), vr.detach().numpy()) < tol
    assert rel_l2(ud.grad.cpu().numpy(), ur.grad.numpy()) < tol
    assert rel_l2(pd.grad.cpu().numpy(), pr.grad.numpy()) < max(tol, 1e-7)


def _config4(n=512):
    area, wl, R = 500e-4, 1064e-7, 50e-4
    x = (np.arange(n) - n // 2) * area / n
    u0 = OP.gaussian_beam(x.copy(), x.reshape(-1, 1).copy(), 1.0, R)
    r = np.sqrt(x ** 2 + x.reshape(-1, 1) ** 2)
    target = np.exp(-(r - 30e-4) ** 2 / (2 * (5e-4) ** 2))
    return area, wl, u0, target


def _loss_oracle(phi, u0, target, area, wl):
    u1 = oracle.propagate_field(u0, area, wl, 1e-1 * 1, phase_mask=phi)         # 1 mm
    u2 = oracle.propagate_field(u1, area, wl, 5e-1, lens_f=5e-1)                # lens 5 mm, 5 mm
    return float(np.mean((np.abs(u2) ** 2 - target) ** 2))


def _loss_gpu(phi_t, u0_t, target_t, area, wl):
    from lightprop2d_amd import autograd as AG
    u1 = AG.propagate(u0_t, 1e-1, area, wl, phase=phi_t)
    u2 = AG.propagate(u1, 5e-1, area, wl, lens_f=5e-1)
    return ((u2.abs() ** 2 - target_t) ** 2).mean()


def test_config4_gradient_vs_central_fd_on_oracle():
    """512^2 FP32 phase mask (seeded U[0,2pi)), Gaussian 50 um, area 500 um, 1064 nm:
    P(1 mm) -> lens(5 mm, reference convention) -> P(5 mm) -> MSE to a ring
    (radius 30 um, Gaussian width sigma 5 um).  Central FD, h = 1e-4 rad, on 64 seeded
    pixels of the complex128 oracle vs the complex64 hand-written adjoint."""
    n = 512
    area, wl, u0, target = _config4(n)
    rs = np.random.RandomState(4)
    phi = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    phi_t = torch.from_numpy(phi).to(DEV).requires_grad_(True)
    u0_t = torch.from_numpy(u0.astype(np.complex64)).to(DEV)
    t_t = torch.from_numpy(target.astype(np.float32)).to(DEV)
    L = _loss_gpu(phi_t, u0_t, t_t, area, wl)
    L.backward()
    g = phi_t.grad.cpu().numpy().astype(np.float64)
    # pick 64 pixels where the beam has power (elsewhere the gradient is ~0)
    inten = np.abs(u0) ** 2
    cand = np.argwhere(inten > 1e-3 * inten.max())
    pix = cand[rs.choice(len(cand), 64, replace=False)]
    h = 1e-4
    phi64 = phi.astype(np.float64)
    fd = []
    for iy, ix in pix:
        p = phi64.copy(); p[iy, ix] += h
        m = phi64.copy(); m[iy, ix] -= h
        fd.append((_loss_oracle(p, u0, target, area, wl) - _loss_oracle(m, u0, target, area, wl)) / (2 * h))
    fd = np.array(fd)
    ours = np.array([g[iy, ix] for iy, ix in pix])
    assert np.linalg.norm(ours - fd) / np.linalg.norm(fd) < 1e-3


def test_config4_adam_200_steps():
    from lightprop2d_amd import autograd as AG  # noqa: F401
    n = 512
    area, wl, u0, target = _config4(n)
    rs = np.random.RandomState(0)
    phi_t = torch.from_numpy(rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)).to(DEV)
    phi_t.requires_grad_(True)
    u0_t = torch.from_numpy(u0.astype(np.complex64)).to(DEV)
    t_t = torch.from_numpy(target.astype(np.float32)).to(DEV)
    opt = torch.optim.Adam([phi_t], lr=0.1)
    losses = []
    for _ in range(200):
        opt.zero_grad()
        L = _loss_gpu(phi_t, u0_t, t_t, area, wl)
        L.backward()
        opt.step()
        losses.append(float(L))
    assert np.isfinite(losses).all()
    assert min(losses[-10:]) < 0.6 * losses[0]


def _screens(K, n, area, rs):
    """White N(0,1) -> Gaussian spectral filter (20 um correlation length) -> 1 rad rms."""
    g = construct_grids(area, n, 800e-7)
    lc = 20e-4
    filt = np.exp(-(np.pi * lc) ** 2 * (g.Kx ** 2 + g.Ky ** 2))
    out = []
    for _ in range(K):
        w = rs.standard_normal((n, n))
        s = np.real(np.fft.ifft2(np.fft.fft2(w) * filt))
        out.append((s / s.std()).astype(np.float32))
    return np.stack(out)


def _screens_gpu(K, n, area, rs):
    """_screens at full size, filtered with the library's own (validated) c128 FFT."""
    from lightprop2d_amd import propagation as P
    g = construct_grids(area, n, 800e-7)
    filt = torch.from_numpy(np.exp(-(np.pi * 20e-4) ** 2 * (g.Kx ** 2 + g.Ky ** 2))).to(DEV)
    out = []
    for _ in range(K):
        w = torch.from_numpy(rs.standard_normal((n, n))).to(torch.complex128).to(DEV)
        s = P.ifft2(P.fft2(w) * filt).real
        out.append((s / s.std()).to(torch.float32))
    return torch.stack(out).contiguous()


@pytest.mark.parametrize("stride,mode", [(1, "segment"), (2, "segment"), (5, "segment"),
                                         (3, "plane")])
def test_stack_grad_vs_oracle(stride, mode):
    from lightprop2d_amd import autograd as AG
    n, area, wl, dz, K = 128, 2e-1 / 16, 800e-7, 0.2, 5
    rs = np.random.RandomState(1)
    scr = _screens(K, n, area, rs)
    u0 = (rs.standard_normal((2, n, n)) + 1j * rs.standard_normal((2, n, n)))
    w = rs.uniform(size=(2, n, n))
    u_t = torch.from_numpy(u0).to(DEV).requires_grad_(True)
    s_t = torch.from_numpy(scr).to(DEV).requires_grad_(True)
    out = AG.propagate_stack(u_t, s_t, dz, area, wl, stride=stride, recompute=mode)
    L = (torch.from_numpy(w).to(DEV) * out.abs() ** 2).sum()
    L.backward()
    # oracle: per batch element, grad_out = 2 w u_out
    gs = np.zeros_like(scr, dtype=np.float64)
    for b in range(2):
        uo = oracle.propagate_stack(u0[b], scr.astype(np.float64), area, wl, dz)
        gu, gl = oracle.propagate_stack_grad(u0[b], scr.astype(np.float64), area, wl, dz, 2 * w[b] * uo)
        gs += np.stack(gl)
        assert rel_l2(u_t.grad[b].cpu().numpy(), gu) < 1e-9
    assert rel_l2(s_t.grad.cpu().numpy(), gs) < 1e-6   # float32 screens


def test_checkpoint_stride_policy():
    from lightprop2d_amd.autograd import checkpoint_stride
    def held(K, s, mode):
        return -(-K // s) + (s if mode == "segment" else 1)

    s, mode = checkpoint_stride(5, 100, 500)        # room for 5 planes
    assert held(5, s, mode) <= 5
    s, mode = checkpoint_stride(16, 100, 800)       # 16 planes in 8
    assert held(16, s, mode) <= 8
    s, mode = checkpoint_stride(5, 100, 300)        # 3 planes: recompute plane by plane
    assert mode == "plane" and -(-5 // s) + 1 <= 3
    assert checkpoint_stride(5, 100, None) == (1, "segment")
    with pytest.raises(Exception):
        checkpoint_stride(5, 100, 100)


def test_config5_full_size_checkpointed():
    """8 x 4096^2 complex64, 5 screens (corr. 20 um, 1 rad rms) 2 mm apart, 800 nm,
    area 2 mm; gradient w.r.t. all screens with a memory budget of 3 planes for the
    stored checkpoints vs. storing every plane: the same loss, gradients equal to
    rounding (a plane input stored by the fused stack's chained row pass and the same
    plane recomputed by a shorter stack end in different passes, whose last bits can
    differ), finite, and a directional derivative that matches a c128 FD."""
    from lightprop2d_amd import autograd as AG
    n, area, wl, dz, K, B = 4096, 2e-1, 800e-7, 0.2, 5, 8
    rs = np.random.RandomState(2)
    scr = _screens_gpu(K, n, area, rs)
    u0 = torch.ones((B, n, n), dtype=torch.complex64, device=DEV)
    tgt = torch.rand((n, n), device=DEV)

    def run(stride=None, budget=None, dtype=torch.complex64, screens=scr):
        s = screens.clone().requires_grad_(True)
        out = AG.propagate_stack(u0.to(dtype), s, dz, area, wl, memory_budget=budget, stride=stride)
        L = ((out.abs() ** 2 - tgt) ** 2).mean()
        L.backward()
        return float(L), s.grad

    plane = B * n * n * 8
    L1, g1 = run(stride=1)
    L3, g3 = run(budget=3 * plane)
    assert L1 == L3
    assert float((g1 - g3).norm() / g1.norm()) < 1e-6
    assert torch.isfinite(g1).all()
    # directional derivative on the complex128 path
    d = torch.from_numpy(rs.standard_normal((K, n, n)).astype(np.float32)).to(DEV)

    def loss128(screens):
        out = AG.propagate_stack(u0[:2].to(torch.complex128), screens, dz, area, wl)
        return float(((out.abs() ** 2 - tgt) ** 2).mean())

    def cfd(eps):
        return (loss128(scr + eps * d) - loss128(scr - eps * d)) / (2 * eps)

    # Richardson extrapolation of two central differences cancels the O(eps^2) term
    fd = (4 * cfd(2e-3) - cfd(4e-3)) / 3
    _, g2 = run(stride=1, dtype=torch.complex64)
    # batch elements are identical (plane wave input): the mean loss of 2 equals that of 8
    assert float((g2.double() * d.double()).sum()) == pytest.approx(fd, rel=1e-3)


@pytest.mark.parametrize("n", [256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("screen", ["shared", "per_field", "none"])
def test_adjoint_workspace_matches_oracle(n, screen):
    """The adjoint with a workspace (column pass out of place, block-tiled intermediate
    read by the gradient schedule: 8 x 8 tiles up to 1024, row-major at 2048, 2 x 4
    tiles at 4096) and without one (every pass in grad_in) against the complex128 oracle adjoint,
    for a shared screen (batch-summed gradient), per-field screens and no screen; batches
    of 5 (3 from 2048 on) leave partial groups in the pipelined passes."""
    from lightprop2d_amd import propagation as P
    rs = np.random.RandomState(n + len(screen))
    area, wl, z = 300e-4, 632e-7, 0.03
    ap = 0.3 / wl
    B = 5 if n <= 1024 else 3
    u = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(np.complex64)
    go = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(np.complex64)
    if screen == "shared":
        phi = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
        phis = [phi] * B
    elif screen == "per_field":
        phi = rs.uniform(0, 2 * np.pi, (B, n, n)).astype(np.float32)
        phis = list(phi)
    else:
        phi, phis = None, [None] * B
    exp_u, exp_p = [], []
    for b in range(B):
        gu, gp = oracle.propagate_field_adjoint(go[b], u[b], area, wl, z, phase_mask=phis[b],
                                                aperture_radius=ap)
        exp_u.append(gu)
        exp_p.append(gp)
    exp_u = np.stack(exp_u)
    ut, gt = torch.from_numpy(u).to(DEV), torch.from_numpy(go).to(DEV)
    pt = torch.from_numpy(phi).to(DEV) if phi is not None else None
    outs = {}
    for ws in (True, False):
        print(f"n={n} screen={screen} workspace={ws}", flush=True)
        gu, gp = P.propagate_adjoint(gt, ut, z, area, wl, phase=pt, aperture_radius=ap,
                                     use_workspace=ws)
        assert rel_l2(gu.cpu().numpy(), exp_u) < 1e-5
        if screen == "shared":
            assert rel_l2(gp.cpu().numpy(), np.sum(exp_p, axis=0)) < 1e-5
        elif screen == "per_field":
            assert rel_l2(gp.cpu().numpy(), np.stack(exp_p)) < 1e-5
        outs[ws] = (gu, gp)
    # same kernels, different intermediate layout: agreement to FP32 rounding
    assert rel_l2(outs[True][0].cpu().numpy(), outs[False][0].cpu().numpy()) < 1e-6
    # a caller-supplied workspace gives the same result as an allocated one
    work = torch.empty_like(gt)
    gu2, _ = P.propagate_adjoint(gt, ut, z, area, wl, phase=pt, aperture_radius=ap, work=work)
    assert torch.equal(gu2, outs[True][0])


@pytest.mark.parametrize("n,dt,B", [(512, torch.complex64, 3), (128, torch.complex64, 2),
                                    (256, torch.complex128, 2), (1024, torch.complex64, 2),
                                    (4096, torch.complex64, 1)])
def test_fused_stack_and_its_planes_vs_oracle(n, dt, B):
    """lp2d_propagate_stack: K = 4 shared screens with a pupil.  At complex64 three-sweep
    sizes (512, 1024) the inverse rows, the next screen and the forward rows of
    consecutive planes run as one chained pass over the block-tiled plane; 128^2 runs
    the whole-field kernel per plane and complex128 the per-plane sweeps.  The output
    and the stored plane inputs u_1 .. u_3 against the oracle, field by field; the
    output equals the plane-by-plane propagation to rounding."""
    from lightprop2d_amd import propagation as P
    area, wl, dz, K = 300e-4, 800e-7, 0.02, (3 if n >= 4096 else 4)
    ap = 0.35 / (area / n)
    rs = np.random.RandomState(n + B)
    u0 = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n)))
    scr = rs.uniform(0, 2 * np.pi, (K, n, n)).astype(np.float32)
    tu = torch.from_numpy(u0).to(dt).to(DEV)
    ts = torch.from_numpy(scr).to(DEV)
    planes = [None] + [torch.empty_like(tu) for _ in range(K - 1)]
    print(f"stack n={n} {dt} B={B}", flush=True)
    out = P.propagate_stack(tu, ts, dz, area, wl, aperture_radius=ap, planes=planes)
    tol = 1e-5 if dt == torch.complex64 else 1e-10
    for b in range(B):
        u = u0[b].astype(np.complex64 if dt == torch.complex64 else np.complex128)
        for k in range(K):
            u = oracle.propagate_field(u, area, wl, dz, phase_mask=scr[k].astype(np.float64),
                                       aperture_radius=ap)
            got = out[b] if k == K - 1 else planes[k + 1][b]
            assert rel_l2(got.cpu().numpy(), u) < tol, (b, k)
    seq = tu
    for k in range(K):
        seq = P.propagate(seq, dz, area, wl, phase=ts[k], aperture_radius=ap)
        got = out if k == K - 1 else planes[k + 1]
        print(f"  plane {k + 1}: max |fused - per-plane| = "
              f"{float((got - seq).abs().max()):.3e}", flush=True)
    assert rel_l2(out.cpu().numpy(), seq.cpu().numpy()) < tol / 10
    # in place, without stored planes
    ip = tu.clone()
    P.propagate_stack(ip, ts, dz, area, wl, aperture_radius=ap, out=ip)
    assert torch.equal(ip, out)


@pytest.mark.parametrize("stride,mode", [(1, "segment"), (2, "segment"), (3, "plane")])
def test_stack_grad_c64_fused_vs_oracle(stride, mode):
    """The checkpointed stack in complex64 at 512^2 (forward and recomputation through
    the fused stack) against the oracle's reverse-mode gradient."""
    from lightprop2d_amd import autograd as AG
    n, area, wl, dz, K = 512, 2e-1 / 8, 800e-7, 0.2, 5
    rs = np.random.RandomState(4)
    scr = _screens(K, n, area, rs)
    u0 = (rs.standard_normal((2, n, n)) + 1j * rs.standard_normal((2, n, n)))
    w = rs.uniform(size=(2, n, n))
    u_t = torch.from_numpy(u0).to(torch.complex64).to(DEV).requires_grad_(True)
    s_t = torch.from_numpy(scr).to(DEV).requires_grad_(True)
    out = AG.propagate_stack(u_t, s_t, dz, area, wl, stride=stride, recompute=mode)
    L = (torch.from_numpy(w).to(DEV) * out.abs() ** 2).sum()
    L.backward()
    gs = np.zeros_like(scr, dtype=np.float64)
    for b in range(2):
        u0b = u0[b].astype(np.complex64).astype(np.complex128)
        uo = oracle.propagate_stack(u0b, scr.astype(np.float64), area, wl, dz)
        gu, gl = oracle.propagate_stack_grad(u0b, scr.astype(np.float64), area, wl, dz,
                                             2 * w[b] * uo)
        gs += np.stack(gl)
        assert rel_l2(u_t.grad[b].cpu().numpy(), gu) < 2e-5
    assert rel_l2(s_t.grad.cpu().numpy(), gs) < 2e-5


@pytest.mark.parametrize("n,dt,B,K", [(256, torch.complex64, 3, 3), (512, torch.complex64, 3, 4),
                                      (1024, torch.complex64, 9, 3),
                                      (4096, torch.complex64, 1, 2),
                                      (2048, torch.complex64, 2, 2), (128, torch.complex64, 3, 3),
                                      (256, torch.complex128, 2, 3)])
def test_fused_stack_adjoint_vs_oracle(n, dt, B, K):
    """lp2d_propagate_stack_adjoint with a pupil, given the plane inputs u_0 .. u_{K-1}
    and grad_out: the gradient of the stack's input (field by field) and the batch-summed
    screen gradients against the oracle's reverse sweep.  256, 512, 1024 and 4096
    complex64 run the chained adjoint (1024 with 9 fields: uneven field groups whose
    gradient planes are summed after each plane); 2048, 128^2 and complex128 the
    plane-by-plane adjoint.  Both forms agree with a plane-by-plane composition of
    lp2d_propagate_adjoint; in place (out = grad_out) the results are unchanged, and
    without screen gradients the field gradient is unchanged (the chained form) or
    equal to rounding (plane by plane: with no shared-screen sum the intermediate may
    take another layout)."""
    from lightprop2d_amd import propagation as P
    area, wl, dz = 300e-4, 800e-7, 0.02
    ap = 0.35 / (area / n)
    rs = np.random.RandomState(7 * n + B)
    c = np.complex64 if dt == torch.complex64 else np.complex128
    u0 = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(c)
    scr = rs.uniform(0, 2 * np.pi, (K, n, n)).astype(np.float32)
    gout = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(c)
    planes = [u0]
    for k in range(K - 1):
        planes.append(oracle.propagate_field(planes[-1], area, wl, dz,
                                             phase_mask=scr[k].astype(np.float64),
                                             aperture_radius=ap).astype(c))
    g = gout.astype(np.complex128)
    gs_ref = [None] * K
    for k in range(K - 1, -1, -1):
        g, gk = oracle.propagate_field_adjoint(g, planes[k], area, wl, dz,
                                               phase_mask=scr[k].astype(np.float64),
                                               aperture_radius=ap)
        gs_ref[k] = gk.sum(axis=0)
    tp = [torch.from_numpy(p).to(DEV) for p in planes]
    ts = torch.from_numpy(scr).to(DEV)
    tg = torch.from_numpy(gout).to(DEV)
    print(f"stack adjoint n={n} {dt} B={B} K={K}", flush=True)
    gu, gs = P.propagate_stack_adjoint(tg, tp, ts, dz, area, wl, aperture_radius=ap)
    tol = 1e-5 if dt == torch.complex64 else 1e-10
    for b in range(B):
        assert rel_l2(gu[b].cpu().numpy(), g[b]) < tol, b
    for k in range(K):
        # float32 gradient planes summed over the batch
        assert rel_l2(gs[k].cpu().numpy(), gs_ref[k]) < (2e-5 if dt == torch.complex64
                                                          else 1e-6), k
    # plane by plane with lp2d_propagate_adjoint
    gseq = tg
    for k in range(K - 1, -1, -1):
        gseq, gk = P.propagate_adjoint(gseq, tp[k], dz, area, wl, phase=ts[k],
                                       aperture_radius=ap)
        assert rel_l2(gs[k].cpu().numpy(), gk.cpu().numpy()) < 1e-5, k
    assert rel_l2(gu.cpu().numpy(), gseq.cpu().numpy()) < tol / 10
    # no screen gradients; in place
    gu2, none = P.propagate_stack_adjoint(tg, tp, ts, dz, area, wl, aperture_radius=ap,
                                          want_screen_grad=False)
    assert none is None
    if n in (256, 512, 1024, 4096) and dt == torch.complex64:
        assert torch.equal(gu2, gu)
    else:
        assert rel_l2(gu2.cpu().numpy(), gu.cpu().numpy()) < tol / 10
    ip = tg.clone()
    gu3, gs3 = P.propagate_stack_adjoint(ip, tp, ts, dz, area, wl, aperture_radius=ap, out=ip)
    assert gu3.data_ptr() == ip.data_ptr() and torch.equal(gu3, gu) and torch.equal(gs3, gs)


def test_stack_adjoint_screen_grad_only_and_errors():
    """LP2D_ADJ_NO_GRAD_IN (screen gradients only) gives the same screen gradients;
    grad_in aliasing a plane input, too few planes and a short workspace are refused
    before any launch."""
    from lightprop2d_amd import propagation as P, _lib as L
    import ctypes as C
    n, B, K, area, wl, dz = 512, 4, 3, 300e-4, 800e-7, 0.02
    rs = np.random.RandomState(11)
    tp = [torch.from_numpy((rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n)))
                           .astype(np.complex64)).to(DEV) for _ in range(K)]
    ts = torch.from_numpy(rs.uniform(0, 6, (K, n, n)).astype(np.float32)).to(DEV)
    tg = torch.from_numpy((rs.standard_normal((B, n, n)) + 0j).astype(np.complex64)).to(DEV)
    gu, gs = P.propagate_stack_adjoint(tg, tp, ts, dz, area, wl)
    none, gs2 = P.propagate_stack_adjoint(tg, tp, ts, dz, area, wl, want_field_grad=False)
    assert none is None and torch.equal(gs2, gs)
    with pytest.raises(P.Lp2dError):
        P.propagate_stack_adjoint(tg, tp, ts, dz, area, wl, out=tp[1])
    with pytest.raises(P.Lp2dError):
        P.propagate_stack_adjoint(tg, tp[:2], ts, dz, area, wl)
    lib = L.lib()
    need = lib.lp2d_propagate_stack_adjoint_ws_size(n, B, L.LP2D_C64)
    assert need > 0   # several field groups or a shared-screen sum over 4 fields
    parr = (C.c_void_p * K)(*[p.data_ptr() for p in tp])
    o = P.optics(area, wl, z=dz)
    work = torch.empty(need // 4, dtype=torch.float32, device=DEV)
    out = torch.empty_like(tg)
    rc = lib.lp2d_propagate_stack_adjoint(n, B, L.LP2D_C64, C.c_void_p(tg.data_ptr()), parr,
                                          C.c_void_p(ts.data_ptr()), K, C.byref(o),
                                          C.c_void_p(out.data_ptr()),
                                          C.c_void_p(gs.data_ptr()), 0,
                                          C.c_void_p(work.data_ptr()), need - 4, None)
    assert rc == L.LP2D_E_SIZE


@pytest.mark.parametrize("dt", [torch.complex64, torch.complex128])
def test_phase_grad_pass_vs_numpy_and_the_fused_adjoint(dt):
    """lp2d_phase_grad against numpy (per field and batch-summed, shared and per-field
    screens, no screen), and an adjoint composed from propagate(-z) and phase_grad
    against lp2d_propagate_adjoint's fused gradient epilogue."""
    from lightprop2d_amd import propagation as P
    n, B, area, wl, z = 256, 3, 300e-4, 800e-7, 0.02
    rs = np.random.RandomState(9)
    c = np.complex64 if dt == torch.complex64 else np.complex128
    u = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(c)
    g = (rs.standard_normal((B, n, n)) + 1j * rs.standard_normal((B, n, n))).astype(c)
    ph1 = rs.uniform(0, 2 * np.pi, (n, n)).astype(np.float32)
    phb = rs.uniform(0, 2 * np.pi, (B, n, n)).astype(np.float32)
    tu, tg = torch.from_numpy(u).to(DEV), torch.from_numpy(g).to(DEV)
    tol = 1e-5 if dt == torch.complex64 else 1e-6     # float32 output
    for ph in (None, ph1, phb):
        e = np.exp(1j * ph.astype(np.float64)) if ph is not None else 1.0
        ref = np.imag(np.conj(u.astype(np.complex128) * e) * g)
        tph = torch.from_numpy(ph).to(DEV) if ph is not None else None
        per = P.phase_grad(tu, tg, tph, sum_batch=False)
        tot = P.phase_grad(tu, tg, tph, sum_batch=True)
        assert rel_l2(per.cpu().numpy(), ref) < tol
        assert rel_l2(tot.cpu().numpy(), ref.sum(axis=0)) < tol
    tph = torch.from_numpy(ph1).to(DEV)
    g1 = P.propagate(tg, -z, area, wl)                 # conj(H(z)) = H(-z)
    composed = P.phase_grad(tu, g1, tph, sum_batch=True)
    _, fused = P.propagate_adjoint(tg, tu, z, area, wl, phase=tph)
    assert rel_l2(composed.cpu().numpy(), fused.cpu().numpy()) < tol


def test_config4_step_replayed_from_a_hip_graph_matches_eager():
    """Configuration 4's Adam step (forward through the screen, lens and both
    propagations, the loss, the hand-written backward, Adam with its step count on the
    device) captured once into a HIP graph: each replay, with other device work and
    synchronisations between replays, gives the eager run's loss for that step bit for
    bit, and the mask after 3 eager + 5 replayed steps equals the eager mask after 8."""
    from lightprop2d_amd import autograd as AG
    from lightprop2d_amd import propagation as P
    n = 512
    area, wl, u0, target = _config4(n)
    phi0 = torch.from_numpy(np.random.RandomState(2).uniform(0, 2 * np.pi, (n, n))
                            .astype(np.float32)).to(DEV)
    u0_t = torch.from_numpy(u0.astype(np.complex64)).to(DEV)
    t_t = torch.from_numpy(target.astype(np.float32)).to(DEV)

    def make():
        phi = phi0.clone().requires_grad_(True)
        return phi, torch.optim.Adam([phi], lr=0.1, capturable=True)

    def body(phi, opt):
        u1 = AG.propagate(u0_t, 1e-1, area, wl, phase=phi)
        u2 = AG.propagate(u1, 5e-1, area, wl, lens_f=5e-1)
        loss = ((u2.real ** 2 + u2.imag ** 2 - t_t) ** 2).mean()
        loss.backward()
        opt.step()
        return loss

    pe, oe = make()
    eager = []
    for _ in range(8):
        oe.zero_grad(set_to_none=True)
        eager.append(float(body(pe, oe).detach()))
    pg, og = make()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(3):
            og.zero_grad(set_to_none=True)
            body(pg, og)
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    og.zero_grad(set_to_none=True)
    with torch.cuda.graph(graph):
        sl = body(pg, og)
    for k in range(3, 8):
        _ = P.propagate(u0_t, 2e-1, area, wl)        # other work between the replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert float(sl.detach()) == eager[k], k
    assert torch.equal(pg.detach(), pe.detach())


def test_config5_fused_stack_adjoint_identity_at_full_size():
    """Configuration 5's size (8 x 4096^2 complex64, 5 screens): the fused stack
    adjoint is the adjoint of the fused stack, Re<g, A u> = Re<A^H g, u> for random u
    and g (sums in FP64), and its screen gradients equal the plane-by-plane adjoints'
    (relative L2 1e-5)."""
    from lightprop2d_amd import propagation as P
    n, area, wl, dz, K, B = 4096, 2e-1, 800e-7, 0.2, 5, 8
    rs = np.random.RandomState(12)
    scr = _screens_gpu(K, n, area, rs)
    g = torch.Generator(device="cpu").manual_seed(12)
    u = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(DEV)
    gout = torch.randn((B, n, n), dtype=torch.complex64, generator=g).to(DEV)
    planes = [u] + [torch.empty_like(u) for _ in range(K - 1)]
    out = P.propagate_stack(u, scr, dz, area, wl, planes=[None] + planes[1:])
    lhs = float((gout.conj() * out).real.double().sum())
    del out
    gin, gs = P.propagate_stack_adjoint(gout, planes, scr, dz, area, wl)
    rhs = float((gin.conj() * u).real.double().sum())
    print(f"<g, A u> = {lhs:.6e}, <A^H g, u> = {rhs:.6e}", flush=True)
    assert rhs == pytest.approx(lhs, rel=1e-5)
    gg = gout
    for k in range(K - 1, -1, -1):
        gg, gk = P.propagate_adjoint(gg, planes[k], dz, area, wl, phase=scr[k])
        assert float((gs[k] - gk).norm() / gk.norm()) < 1e-5, k
    assert float((gin - gg).norm() / gg.norm()) < 1e-5
