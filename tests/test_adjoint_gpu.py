"""The force adjoint and the checkpointed reversed-leapfrog gradient.

Pins: the VJP kernel against an FP64 torch-autograd VJP of the same force
and a closed-form two-body Jacobian; trajectory gradients against FP64
central finite differences of the engine's own FP64 forward and an FP64
torch-autograd restatement of the KDK integrator.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _dev():
    return torch.device("cuda:0")


def _torch_accel(x, m, eps, G=1.0):
    d = x[None, :, :] - x[:, None, :]
    eye = torch.eye(x.shape[0], dtype=x.dtype, device=x.device)
    s = (d * d).sum(-1) + eps * eps + eye  # the diagonal is masked below
    f = s.pow(-1.5) * (1 - eye)
    return G * (f[:, :, None] * d * m[None, :, None]).sum(1)


def _vjp_engine(gpu, precision, x, m, lam, eps, G=1.0):
    n = x.shape[0]
    with gpu.Engine("hip-direct", precision, softening=eps, G=G) as eng:
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        xs = x.t().contiguous()
        eng.set_state_device(n, xs.data_ptr(), None, m.data_ptr())
        lam_s = lam.t().contiguous()
        xbar = torch.zeros(3, n, dtype=torch.float64, device=x.device)
        mbar = torch.zeros(n, dtype=torch.float64, device=x.device)
        eng.force_vjp(lam_s.data_ptr(), xbar.data_ptr(), mbar.data_ptr())
        torch.cuda.synchronize()
    return xbar.t(), mbar


@pytest.mark.parametrize("n,eps,G", [(2, 0.0, 1.0), (300, 0.05, 1.3), (1100, 0.02, 1.0)])
def test_vjp_against_torch_autograd(gpu, n, eps, G):
    g = torch.Generator().manual_seed(n)
    x = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    m = (torch.rand(n, generator=g, dtype=torch.float64) + 0.5).to(_dev())
    lam = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    xr, mr = x.clone().requires_grad_(), m.clone().requires_grad_()
    (_torch_accel(xr, mr, eps, G) * lam).sum().backward()
    xb, mb = _vjp_engine(gpu, "fp64", x, m, lam, eps, G)
    assert torch.allclose(xb, xr.grad, rtol=1e-10, atol=1e-10 * xr.grad.abs().max().item())
    assert torch.allclose(mb, mr.grad, rtol=1e-10, atol=1e-10 * mr.grad.abs().max().item())
    if eps > 0:
        xm, mm = _vjp_engine(gpu, "mixed", x, m, lam, eps, G)
        scale_x, scale_m = xr.grad.abs().max().item(), mr.grad.abs().max().item()
        assert (xm - xr.grad).abs().max().item() < 1e-5 * scale_x
        assert (mm - mr.grad).abs().max().item() < 1e-5 * scale_m


def test_vjp_two_body_closed_form(gpu):
    # xbar_0 = J(d) (m_0 lam_1 - m_1 lam_0), J = s^-3/2 I - 3 s^-5/2 d d^T
    x = torch.tensor([[0.0, 0, 0], [1.0, 0.5, -0.25]], dtype=torch.float64, device=_dev())
    m = torch.tensor([2.0, 3.0], dtype=torch.float64, device=_dev())
    lam = torch.tensor([[1.0, 0, 0], [0, 1.0, 0]], dtype=torch.float64, device=_dev())
    eps = 0.1
    xb, mb = _vjp_engine(gpu, "fp64", x, m, lam, eps)
    d = (x[1] - x[0]).cpu().numpy()
    s = d @ d + eps * eps
    J = s ** -1.5 * np.eye(3) - 3 * s ** -2.5 * np.outer(d, d)
    w = 2.0 * lam[1].cpu().numpy() - 3.0 * lam[0].cpu().numpy()
    assert np.allclose(xb[0].cpu().numpy(), J @ w, rtol=1e-13)
    assert np.allclose(xb[1].cpu().numpy(), -J @ w, rtol=1e-13)  # momentum-like balance
    # mbar_0 = -lam_1 . (x_0 - x_1) f... : d a_1 / d m_0 = (x_0 - x_1) s^-3/2
    assert np.isclose(mb[0].item(), lam[1].cpu().numpy() @ (-d) * s ** -1.5, rtol=1e-13)


def _two_disk(gpu, n):
    pos, vel, m = gpu.generators.two_disk(n // 2, softening=0.05, seed=5)
    t = lambda a: torch.tensor(a, dtype=torch.float64, device=_dev())  # noqa: E731
    return t(pos), t(vel), t(m)


def _loss(x, target):
    return ((x - target) ** 2).sum(1).mean()


def test_trajectory_gradient_against_autograd_reference(gpu):
    from gcsim import adjoint

    x0, v0, m = _two_disk(gpu, 1024)
    steps, dt = 60, 1e-3
    target = x0 + 0.01 * torch.randn(x0.shape, dtype=torch.float64, device=_dev(),
                                     generator=torch.Generator(device=_dev()).manual_seed(1))
    xr, vr, mr = (t.clone().requires_grad_() for t in (x0, v0, m))
    xT, _ = adjoint.reference_trajectory(xr, vr, mr, dt, steps, softening=0.05)
    _loss(xT, target).backward()
    xa, va, ma = (t.clone().requires_grad_() for t in (x0, v0, m))
    xT2, _ = adjoint.trajectory(xa, va, ma, dt, steps, checkpoint_every=20, softening=0.05)
    assert torch.allclose(xT2, xT.detach(), rtol=0, atol=1e-12)
    _loss(xT2, target).backward()
    for got, want in ((xa.grad, xr.grad), (va.grad, vr.grad), (ma.grad, mr.grad)):
        err = (got - want).abs().max().item() / want.abs().max().item()
        assert err < 1e-8, err


def test_trajectory_gradient_against_finite_differences(gpu):
    # C3-shaped check at reduced size: central differences of the engine's
    # own FP64 forward along random directions in (v0, m)
    from gcsim import adjoint

    x0, v0, m = _two_disk(gpu, 2048)
    steps, dt, h = 100, 1e-3, 1e-6
    gen = torch.Generator(device=_dev()).manual_seed(9)
    target = x0 + v0 * steps * dt + 0.005 * torch.randn(x0.shape, dtype=torch.float64,
                                                        device=_dev(), generator=gen)
    eng = gpu.Engine("hip-direct", "fp64", softening=0.05)
    va, ma = v0.clone().requires_grad_(), m.clone().requires_grad_()
    xT, _ = adjoint.trajectory(x0, va, ma, dt, steps, engine=eng, checkpoint_every=20)
    _loss(xT, target).backward()

    def L(v, mm):
        with torch.no_grad():
            x, _ = adjoint.trajectory(x0, v, mm, dt, steps, engine=eng, checkpoint_every=20)
            return _loss(x, target).item()

    for k in range(4):
        du = torch.randn(v0.shape, dtype=torch.float64, device=_dev(), generator=gen)
        dm = torch.randn(m.shape, dtype=torch.float64, device=_dev(), generator=gen) * 1e-3
        fd = (L(v0 + h * du, m + h * dm) - L(v0 - h * du, m - h * dm)) / (2 * h)
        ad = (va.grad * du).sum().item() + (ma.grad * dm).sum().item()
        assert abs(fd - ad) <= 1e-5 * abs(ad) + 1e-9, (k, fd, ad)
    eng.close()


def _vjp_mixed_form(gpu, x, m, lam, eps, pairs, ipt=0):
    n = x.shape[0]
    with gpu.Engine("hip-direct", "mixed", softening=eps) as eng:
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        eng.set_pairs(pairs)
        if ipt:
            eng.set_launch_shape(ipt, 0)
        eng.set_state_device(n, x.t().contiguous().data_ptr(), None, m.data_ptr())
        lam_s = lam.t().contiguous()
        out = []
        for _ in range(2):
            xbar = torch.zeros(3, n, dtype=torch.float64, device=x.device)
            mbar = torch.zeros(n, dtype=torch.float64, device=x.device)
            eng.force_vjp(lam_s.data_ptr(), xbar.data_ptr(), mbar.data_ptr())
            torch.cuda.synchronize()
            out.append((xbar.t().clone(), mbar.clone()))
    return out


@pytest.mark.parametrize("n", [65536, 4097, 300])
def test_pair_once_vjp_matches_per_target_vjp(gpu, n):
    """The pair-once VJP (the default for full-range mixed states) against
    the per-target one and the FP64 VJP: same 1e-5 bound, deterministic,
    every targets-per-lane shape agreeing."""
    pos, _, mass = gpu.generators.config1(n=65536)
    idx = np.random.default_rng(n).choice(65536, n, replace=False)
    x = torch.tensor(pos[idx], dtype=torch.float64, device=_dev())
    m = torch.tensor(mass[idx], dtype=torch.float64, device=_dev())
    g = torch.Generator().manual_seed(n)
    lam = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    eps = 0.02
    ref_x, ref_m = _vjp_engine(gpu, "fp64", x, m, lam, eps)
    sx, sm = ref_x.abs().max().item(), ref_m.abs().max().item()
    (sym_x, sym_m), (sym_x2, sym_m2) = _vjp_mixed_form(gpu, x, m, lam, eps, "symmetric")
    assert torch.equal(sym_x, sym_x2) and torch.equal(sym_m, sym_m2)  # deterministic
    (tgt_x, tgt_m), _ = _vjp_mixed_form(gpu, x, m, lam, eps, "targets")
    for gx, gm in ((sym_x, sym_m), (tgt_x, tgt_m)):
        assert (gx - ref_x).abs().max().item() < 1e-5 * sx
        assert (gm - ref_m).abs().max().item() < 1e-5 * sm
    for ipt in (1, 4):
        (ix, im), _ = _vjp_mixed_form(gpu, x, m, lam, eps, "symmetric", ipt)
        assert (ix - ref_x).abs().max().item() < 1e-5 * sx
        assert (im - ref_m).abs().max().item() < 1e-5 * sm


@pytest.mark.parametrize("ipt", [0, 1, 4])
def test_pair_once_vjp_equal_mass_tiles(gpu, ipt):
    """Tiles of one mass each (two masses over the 8 tiles): the pair-once
    VJP against torch autograd in FP64, at the mixed path's 1e-5 bound,
    and deterministic."""
    n = 2048
    g = torch.Generator().manual_seed(77)
    x = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    m = torch.full((n,), 3e-4, dtype=torch.float64)
    m[1024:] = 7e-4  # tiles 4-7 of another mass
    m = m.to(_dev())
    lam = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    xr, mr = x.clone().requires_grad_(), m.clone().requires_grad_()
    (_torch_accel(xr, mr, 0.05) * lam).sum().backward()
    (gx, gm), (gx2, gm2) = _vjp_mixed_form(gpu, x, m, lam, 0.05, "symmetric", ipt)
    assert torch.equal(gx, gx2) and torch.equal(gm, gm2)
    assert (gx - xr.grad).abs().max().item() < 1e-5 * xr.grad.abs().max().item()
    assert (gm - mr.grad).abs().max().item() < 1e-5 * mr.grad.abs().max().item()


def test_mixed_vjp_forms_without_softening(gpu):
    """eps = 0 runs the guarded kernels (a pair closer than 1e-15 adds
    nothing); both mixed forms against the FP64 VJP, and a coincident pair
    contributes exactly nothing."""
    n = 700
    g = torch.Generator().manual_seed(71)
    x = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    m = (torch.rand(n, generator=g, dtype=torch.float64) + 0.5).to(_dev())
    lam = torch.randn(n, 3, generator=g, dtype=torch.float64).to(_dev())
    ref_x, ref_m = _vjp_engine(gpu, "fp64", x, m, lam, 0.0)
    sx, sm = ref_x.abs().max().item(), ref_m.abs().max().item()
    for pairs in ("symmetric", "targets"):
        (gx, gm), _ = _vjp_mixed_form(gpu, x, m, lam, 0.0, pairs)
        assert torch.isfinite(gx).all() and torch.isfinite(gm).all()
        assert (gx - ref_x).abs().max().item() < 1e-5 * sx, pairs
        assert (gm - ref_m).abs().max().item() < 1e-5 * sm, pairs
    # two particles at one point and one apart: only the distant pairs count
    x2 = torch.tensor([[0.0, 0.0, 0.0], [0.0, 0.0, 0.0], [1.0, 0.5, -0.25]],
                      dtype=torch.float64, device=_dev())
    m2 = torch.tensor([1.0, 2.0, 0.5], dtype=torch.float64, device=_dev())
    l2 = torch.tensor([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]],
                      dtype=torch.float64, device=_dev())
    r64 = _vjp_engine(gpu, "fp64", x2, m2, l2, 0.0)
    (sx2, sm2), _ = _vjp_mixed_form(gpu, x2, m2, l2, 0.0, "symmetric")
    assert torch.allclose(sx2, r64[0], rtol=1e-5, atol=1e-6)
    assert torch.allclose(sm2, r64[1], rtol=1e-5, atol=1e-6)
