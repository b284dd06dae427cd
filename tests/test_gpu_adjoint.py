"""The adjoint of the time step (slevolve_amd.adjoint).

* transposes: <A g, U> = <g, A^T U> and <S U, g> = <U, S^T g> (the real inner products of the
  stored arrays) for random fields at several grids, incl. lmax = nlon/2 (Nyquist order); the
  reverse modal recursion likewise against the advance;
* the frozen-mask model reproduces the device solver's trajectory (masks taken from that run);
* the gradient of the adjoint configuration (256 x 512, lmax 256, 300 x 1 yr, 200 ocean points and
  the J2 rate) against central finite differences of the model on 50 seeded random directions,
  through the torch.autograd.Function;
* the FP64 reference gradient: the same frozen model built on an independent CPU implementation
  of the transforms (dense Legendre tables of the oracle, torch FFTs) and differentiated by torch
  autograd through the unrolled fixed-point iterations, against the hand-written reverse sweep.
"""
import numpy as np
import pytest
import torch

from slevolve_amd import synthetic as syn

pytestmark = pytest.mark.gpu


def rdot(a, b):
    """Real inner product of two arrays as stored (complex: real and imaginary parts)."""
    return float((a.conj() * b).real.sum()) if a.is_complex() else float((a * b).sum())


@pytest.mark.parametrize("nlat,L", [(64, 31), (128, 63), (256, 256), (512, 511)])
def test_transposes_dot_products(nlat, L):
    from slevolve_amd import SHT
    from slevolve_amd.adjoint import DeviceOps

    print(f"[adjoint] transposes nlat {nlat} lmax {L}", flush=True)
    ops = DeviceOps(SHT(nlat, 2 * nlat, L))
    gen = torch.Generator(device="cuda").manual_seed(nlat + L)
    g = torch.randn((2, nlat, 2 * nlat), dtype=torch.float64, device="cuda", generator=gen)
    U = torch.randn((2, ops.nc), dtype=torch.complex128, device="cuda", generator=gen)
    for b in range(2):
        lhs, rhs = rdot(ops.A(g[b:b + 1]), U[b:b + 1]), rdot(g[b:b + 1], ops.AT(U[b:b + 1]))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1e-300) + 1e-13, (lhs, rhs)
        lhs, rhs = rdot(ops.S(U[b:b + 1]), g[b:b + 1]), rdot(U[b:b + 1], ops.ST(g[b:b + 1]))
        assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), 1e-300) + 1e-13, (lhs, rhs)
    # batched calls equal per-field calls
    assert torch.equal(ops.AT(U)[1], ops.AT(U[1:2])[0]) and torch.equal(ops.ST(g)[1], ops.ST(g[1:2])[0])


def test_modal_reverse_dot_product(prem_text):
    from slevolve_amd import SHT, LoveNumbers
    from slevolve_amd.adjoint import DeviceOps, FrozenStepModel
    from slevolve_amd.solver import Planet

    L = 63
    ops = DeviceOps(SHT(64, 128, L))
    z = torch.zeros((1, 64, 128), dtype=torch.float64, device="cuda")
    mdl = FrozenStepModel(ops, LoveNumbers(prem_text, L, text=True), Planet(), L, 500.0, z[0], z[0], z, z)
    K, nc = mdl.K, ops.nc
    gen = torch.Generator(device="cuda").manual_seed(5)
    rnd = lambda *s: torch.randn(s, dtype=torch.complex128, device="cuda", generator=gen)  # noqa: E731
    R, P = rnd(1, K, nc), rnd(1, nc)
    Rb, vb, rb = rnd(1, K, nc), rnd(1, nc), rnd(1, nc)
    R1, v, r = ops.modal(mdl.tabs, R.clone(), P)
    lhs = rdot(R1, Rb) + rdot(v, vb) + rdot(r, rb)
    Rbar, Pbar = Rb.clone(), torch.zeros_like(P)
    ops.modal_reverse(mdl.tabs, Rbar, vb, rb, Pbar)
    rhs = rdot(R, Rbar) + rdot(P, Pbar)
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs), (lhs, rhs)


def _nominal(name, nsteps, eps=1e-12):
    """The device solver on workload `name` for nsteps steps and the frozen-mask model of it."""
    from slevolve_amd import SHT, LoveNumbers, SeaLevelSolver
    from slevolve_amd.adjoint import frozen_model_from_run

    w = syn.workload(name)
    ice = torch.stack([torch.from_numpy(w.ice(k)) for k in range(nsteps + 1)]).cuda()
    slv = SeaLevelSolver(SHT(w.nlat, w.nlon, w.lmax), LoveNumbers(w.love_text, w.lmax, text=True), w.topo, ice[0],
                         norder=w.lmax, dt=w.dt, epsilon1=eps)
    model, outs = frozen_model_from_run(slv, ice, nsteps, record_outputs=True)
    return w, ice, model, outs


def test_frozen_model_reproduces_the_solver():
    w, ice, model, outs = _nominal("lmax256", 12)
    dsl, G20, ck = model.forward(ice)
    tg = model.tinit - dsl[0]
    err = float((tg - outs["tgrid"][-1]).abs().max())
    print(f"[adjoint] frozen model vs solver after 12 steps: max |dtgrid| {err:.2e} m", flush=True)
    assert err < 1e-8
    # restart from the checkpoint of step 0 reproduces the same final state
    d2, _ = model.resume(ck[0], 0, ice[None], 12)
    assert torch.allclose(d2, dsl, rtol=0, atol=1e-12)


def test_gradient_vs_central_differences_lmax256():
    """The adjoint configuration: 300 steps, 50 seeded directions (batched forward runs).  The
    hand-written FP64 gradient against central differences of the model with 1 m perturbations
    (the model is affine in the ice up to the quadratic centrifugal terms, so the agreement is set
    by those and by the fixed-point tolerance: ~1e-8 measured); the FP32 gradient of the
    torch.autograd.Function (FP32 ice history in, FP32 gradient out) is the FP64 gradient of that
    input, rounded."""
    from slevolve_amd.adjoint import IceHistoryMisfit, misfit_points

    N, ndir = 300, 50
    w, ice, model, _ = _nominal("lmax256", N)
    pts = misfit_points(model, 200)
    assert pts.numel() == 200
    J64, g64 = model.gradient(ice, pts)
    ice32 = ice.float().requires_grad_(True)
    J = IceHistoryMisfit.apply(ice32, model, pts)
    J.backward()
    # the Function's FP32 gradient is the FP64 gradient of its (FP32-rounded) input, rounded
    J32in, g32in = model.gradient(ice32.detach().double(), pts)
    assert ice32.grad.dtype == torch.float32 and torch.equal(ice32.grad, g32in.float())
    assert float(J.detach()) == float(torch.tensor(float(J32in[0]), dtype=torch.float32))
    print(f"[adjoint] J = {float(J64[0]):.6e}, |grad| max {float(g64.abs().max()):.3e}", flush=True)
    seed = int(np.random.SeedSequence(20261203 + 5).spawn(5)[4].generate_state(1)[0])
    gen = torch.Generator(device="cuda").manual_seed(seed)
    eps = 1.0  # metres of ice
    dirs = torch.randn((ndir, N + 1, w.nlat, w.nlon), dtype=torch.float64, device="cuda", generator=gen)
    base = ice.double()
    batch = torch.cat([base[None] + eps * dirs, base[None] - eps * dirs])
    dslN, G20, _ = model.forward(batch, checkpoint=0)
    Jb = model.misfit(dslN, G20, pts)
    fd = (Jb[:ndir] - Jb[ndir:]) / (2 * eps)
    ad = torch.tensor([float((g64 * dirs[i]).sum()) for i in range(ndir)], dtype=torch.float64)
    rel = ((fd - ad).abs() / fd.abs().clamp_min(1e-300)).max()
    print(f"[adjoint] 50 directions: max |FD - adjoint| / |FD| = {float(rel):.2e}", flush=True)
    assert float(rel) < 1e-7


class _CpuOps:
    """Independent FP64 CPU transforms for the reference gradient: dense Legendre tables of the
    oracle (legendre_block), torch FFTs, differentiable end to end."""

    def __init__(self, nlat, nlon, L):
        from oracle.sht import SQRT2_LEGACY, gauss_legendre, legendre_block

        self.nlat, self.nlon, self.L, self.dev = nlat, nlon, L, torch.device("cpu")
        x, w = gauss_legendre(nlat)
        nh = nlat // 2
        blk = legendre_block(x[:nh], L)
        l_of = np.concatenate([np.arange(m, L + 1) for m in range(L + 1)])
        m_of = np.concatenate([np.full(L + 1 - m, m) for m in range(L + 1)])
        self.nc = l_of.size
        Pd = np.zeros((nlat, self.nc))
        for i, (l, m) in enumerate(zip(l_of, m_of)):
            p = blk[m][l - m]
            Pd[:nh, i] = p
            Pd[nlat - 1 - np.arange(nh), i] = p * (-1.0) ** (l - m)
        self.Pd = torch.from_numpy(Pd)
        self.w = torch.from_numpy(w)
        self.m_of = torch.from_numpy(m_of)
        self.l_of = torch.from_numpy(l_of)
        self.s2 = SQRT2_LEGACY

    def A(self, g):
        F = torch.fft.rfft(g, dim=-1)[..., : self.L + 1] / self.nlon
        Fm = F[:, :, self.m_of]
        return (Fm * (self.Pd * self.w[:, None])[None]).sum(dim=1) / self.s2

    def S(self, c):
        X = torch.zeros((c.shape[0], self.nlat, self.L + 1), dtype=torch.complex128)
        X = X.index_add(2, self.m_of, c[:, None, :] * self.Pd[None].to(torch.complex128))
        return torch.fft.irfft(X * (self.nlon * self.s2), n=self.nlon, dim=-1)

    def modal(self, tabs, R, P):
        d = tabs["dec"][:, self.l_of]
        o = tabs["omd"][:, self.l_of]
        Rn = d[None] * R + o[None] * P[:, None, :]
        return Rn, (tabs["asl"][:, self.l_of][None] * Rn).sum(dim=1), (tabs["arr"][:, self.l_of][None] * Rn).sum(dim=1)


def test_gradient_vs_cpu_autograd_reference():
    """FP64 reference gradient on the 64 x 128, lmax 63 workload (8 steps)."""
    from slevolve_amd.adjoint import FrozenStepModel

    N = 8
    w, ice, model, _ = _nominal("lmax63", N)
    pts = torch.arange(0, w.nlat * w.nlon, 97, device="cuda")
    pts = pts[model.cstar[N].reshape(-1)[pts] > 0.5][:40]
    J_gpu, g_gpu = model.gradient(ice, pts)
    cpu = FrozenStepModel(_CpuOps(w.nlat, w.nlon, w.lmax), model_love(w), model_planet(), w.lmax, w.dt,
                          model.tinit.cpu(), model.cstar0.cpu(), model.keep.cpu(),
                          model.cstar.cpu(), tol=1e-14)
    x = ice.double().cpu().requires_grad_(True)
    dsl, G20, _ = cpu.forward(x, checkpoint=0)
    J = cpu.misfit(dsl, G20, pts.cpu())[0]
    J.backward()
    rel_J = abs(float(J) - float(J_gpu[0])) / abs(float(J))
    err = float((x.grad - g_gpu.cpu()).abs().max() / x.grad.abs().max())
    print(f"[adjoint] CPU autograd reference: |dJ|/J {rel_J:.2e}, max |dgrad| / max |grad| {err:.2e}", flush=True)
    assert rel_J < 1e-10 and err < 1e-8


def model_love(w):
    from slevolve_amd import LoveNumbers

    return LoveNumbers(w.love_text, w.lmax, text=True)


def model_planet():
    from slevolve_amd.solver import Planet

    return Planet()
