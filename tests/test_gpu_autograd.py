"""GPU: the differentiable relaxation behind a torch.autograd.Function (test harness only).

The membership loss is a function of the per-voxel weight scale, the Gaussian-RBF sigma and the
polarity alpha; forward() runs the library's checkpointed adjoint through the C-ABI (loss and
all three gradients in one call) and backward() scales the stored gradients by the incoming
one.  Checked: autograd's gradients equal the adjoint's own output; torch.autograd.gradcheck on
sigma and alpha with the FP64 adjoint (central differences of the same GPU loss); FP64 central
differences of a few per-voxel scale entries to 1e-4 relative; a descent step lowers the loss.
The discrete OIFT forest is an input here (the oracle's labels), not differentiated.
"""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P

pytestmark = pytest.mark.gpu
EPS = 1e-3


class RelaxLoss:
    """Wraps the adjoint as an autograd Function (built lazily so that importing this module
    needs no torch)."""

    fn = None

    @classmethod
    def get(cls):
        if cls.fn is None:
            import torch

            class _Fn(torch.autograd.Function):
                @staticmethod
                def forward(fctx, scale, sigma, alpha, rctx, problem, fp64):
                    image, seeds, label, target, niter, adj = problem
                    loss, gsc, gs, ga, _ = rctx.relax_adjoint(image, seeds, label, 1, niter, scale.detach(),
                                                              float(sigma), float(alpha), target, ckpt_every=10,
                                                              adj_r=adj, eps=EPS, fp64=fp64)
                    fctx.save_for_backward(gsc)
                    fctx.g_sigma, fctx.g_alpha = gs, ga
                    fctx.like = [(t.dtype, t.device) for t in (scale, sigma, alpha)]
                    return torch.tensor(loss, dtype=torch.float64, device=scale.device)

                @staticmethod
                def backward(fctx, g):
                    (gsc,) = fctx.saved_tensors
                    (dsc, vsc), (dsg, vsg), (dal, val) = fctx.like
                    return ((g * gsc).to(dtype=dsc, device=vsc), (g * fctx.g_sigma).to(dtype=dsg, device=vsg),
                            (g * fctx.g_alpha).to(dtype=dal, device=val), None, None, None)

            cls.fn = _Fn
        return cls.fn


def _problem(shape, seed, adj, niter):
    import torch
    c = P.small_case(shape, seed, per=50.0, adj_r=adj, top=255)
    lab = O.oift_image(c.image, c.seeds, c.label_init(), adj, 50.0)["label"]
    rng = np.random.default_rng(seed)
    scale = (1.0 + 0.2 * rng.standard_normal(c.image.shape)).astype(np.float32)
    target = np.clip(0.1 + 0.8 * (lab > 0) + 0.05 * rng.standard_normal(c.image.shape), 0.01, 0.99)
    dev = torch.device("cuda")
    prob = (torch.from_numpy(c.image.astype(np.int16)).to(dev), torch.from_numpy(c.seeds.astype(np.int32)).to(dev),
            torch.from_numpy(lab.astype(np.int32)).to(dev), torch.from_numpy(target.astype(np.float32)).to(dev),
            niter, adj)
    return prob, torch.from_numpy(scale).to(dev)


def test_autograd_gradients_equal_the_adjoint(ctx):
    import torch
    prob, scale0 = _problem((12, 13, 14), 31, 1.0, 20)
    scale = scale0.clone().requires_grad_(True)
    sigma = torch.tensor(28.0, dtype=torch.float64, device="cuda", requires_grad=True)
    alpha = torch.tensor(0.35, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = RelaxLoss.get().apply(scale, sigma, alpha, ctx, prob, False)
    loss.backward()
    image, seeds, label, target, niter, adj = prob
    l2, gsc, gs, ga, _ = ctx.relax_adjoint(image, seeds, label, 1, niter, scale0, 28.0, 0.35, target,
                                           ckpt_every=10, adj_r=adj, eps=EPS)
    assert float(loss.detach()) == pytest.approx(l2, rel=1e-12)
    assert torch.equal(scale.grad, gsc)
    assert float(sigma.grad) == pytest.approx(gs, rel=1e-9) and float(alpha.grad) == pytest.approx(ga, rel=1e-9)
    # the chain rule through a downstream op: d(3 L)/dsigma = 3 dL/dsigma
    sigma.grad = None
    (3.0 * RelaxLoss.get().apply(scale, sigma, alpha, ctx, prob, False)).backward()
    assert float(sigma.grad) == pytest.approx(3.0 * gs, rel=1e-9)


@pytest.mark.parametrize("adj", [1.0, 1.75])
def test_autograd_gradcheck_sigma_alpha_fp64(ctx, adj):
    import torch
    prob, scale = _problem((9, 10, 11), 32, adj, 12)
    sigma = torch.tensor(25.0, dtype=torch.float64, device="cuda", requires_grad=True)
    alpha = torch.tensor(0.4, dtype=torch.float64, device="cuda", requires_grad=True)
    fn = RelaxLoss.get()
    assert torch.autograd.gradcheck(lambda s, a: fn.apply(scale, s, a, ctx, prob, True), (sigma, alpha),
                                    eps=1e-5, atol=1e-9, rtol=1e-4)


def test_autograd_scale_gradient_fp64_central_differences(ctx):
    import torch
    prob, scale0 = _problem((10, 11, 12), 33, 1.0, 15)
    fn = RelaxLoss.get()
    sigma = torch.tensor(30.0, dtype=torch.float64, device="cuda")
    alpha = torch.tensor(0.3, dtype=torch.float64, device="cuda")
    scale = scale0.clone().requires_grad_(True)
    fn.apply(scale, sigma, alpha, ctx, prob, True).backward()
    g = scale.grad.double().cpu().numpy().ravel()
    for idx in np.argsort(-np.abs(g))[:4]:      # the entries that matter most
        up, dn = scale0.clone(), scale0.clone()
        up.view(-1)[idx] += 1e-3
        dn.view(-1)[idx] -= 1e-3
        h = float(up.view(-1)[idx]) - float(dn.view(-1)[idx])   # the float32 step actually taken
        lu = float(fn.apply(up, sigma, alpha, ctx, prob, True))
        ld = float(fn.apply(dn, sigma, alpha, ctx, prob, True))
        fd = (lu - ld) / h
        assert abs(g[idx] - fd) <= 1e-4 * abs(fd), (idx, g[idx], fd)


def test_autograd_descent_step_lowers_the_loss(ctx):
    import torch
    prob, scale0 = _problem((12, 12, 12), 34, 1.0, 20)
    fn = RelaxLoss.get()
    scale = scale0.clone().requires_grad_(True)
    sigma = torch.tensor(30.0, dtype=torch.float64, device="cuda", requires_grad=True)
    alpha = torch.tensor(0.4, dtype=torch.float64, device="cuda", requires_grad=True)
    loss = fn.apply(scale, sigma, alpha, ctx, prob, False)
    loss.backward()
    with torch.no_grad():
        gn2 = float((scale.grad.double() ** 2).sum() + sigma.grad ** 2 + alpha.grad ** 2)
        lr = 1e-2 * float(loss.detach()) / gn2   # a small fraction of the first-order decrease
        s2 = (scale - lr * scale.grad).detach()
        sg2, al2 = (sigma - lr * sigma.grad).detach(), (alpha - lr * alpha.grad).detach()
    assert float(fn.apply(s2, sg2, al2, ctx, prob, False)) < float(loss.detach())
