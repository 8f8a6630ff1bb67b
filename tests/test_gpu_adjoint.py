"""GPU: the checkpointed adjoint of the differentiable relaxation against the FP64 oracle's
reverse-mode gradients and against FP64 central finite differences, plus checkpoint
invariance.  FP32 states bound the agreement: memberships within 1e-5 absolute, loss within
1e-5 relative, gradients within 2e-3 relative (of the largest entry for the per-voxel scale).
The BCE clamp eps is 1e-3: with o = 1 - f near 0 the FP32 state's absolute error would be
amplified by 1/o.)"""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P

pytestmark = pytest.mark.gpu
EPS = 1e-3


def _setup(shape, seed, adj):
    c = P.small_case(shape, seed, per=50.0, adj_r=adj, top=255)
    lab = O.oift_image(c.image, c.seeds, c.label_init(), adj, 50.0)["label"]
    rng = np.random.default_rng(seed)
    scale = (1.0 + 0.2 * rng.standard_normal(c.image.shape)).astype(np.float32)
    target = np.clip(0.1 + 0.8 * (lab > 0) + 0.05 * rng.standard_normal(c.image.shape), 0.01, 0.99)
    return c, lab, scale, target.astype(np.float32)


def _gpu(ctx, c, lab, scale, target, niter, sigma, alpha, adj, ckpt):
    import torch
    dev = torch.device("cuda")
    return ctx.relax_adjoint(torch.from_numpy(c.image.astype(np.int16)).to(dev),
                             torch.from_numpy(c.seeds.astype(np.int32)).to(dev),
                             torch.from_numpy(lab.astype(np.int32)).to(dev), 1, niter,
                             torch.from_numpy(scale).to(dev), sigma, alpha, torch.from_numpy(target).to(dev),
                             ckpt_every=ckpt, adj_r=adj, eps=EPS, want_memb=True)


@pytest.mark.parametrize("shape,seed,adj,niter", [((12, 14, 16), 21, 1.0, 25), ((10, 11, 12), 22, 1.75, 12)])
def test_adjoint_matches_fp64_oracle_and_finite_differences(ctx, shape, seed, adj, niter):
    c, lab, scale, target = _setup(shape, seed, adj)
    sigma, alpha = 30.0, 0.4
    print(f"case shape={shape} adj={adj} niter={niter}", flush=True)
    loss, gsc, gs, ga, f = _gpu(ctx, c, lab, scale, target, niter, sigma, alpha, adj, 10)
    r = O.relax_diff(c.image, c.seeds, lab, 1, niter, scale.astype(np.float64), sigma, alpha,
                     target.astype(np.float64), adj, eps=EPS)
    assert np.abs(f.cpu().numpy() - r["f"]).max() <= 1e-5
    assert abs(loss - r["loss"]) <= 1e-5 * abs(r["loss"])
    assert abs(gs - r["g_sigma"]) <= 2e-3 * abs(r["g_sigma"])
    assert abs(ga - r["g_alpha"]) <= 2e-3 * abs(r["g_alpha"])
    g = gsc.cpu().numpy().astype(np.float64)
    assert np.abs(g - r["g_scale"]).max() <= 2e-3 * np.abs(r["g_scale"]).max()
    # FP64 central differences (h = 1e-4) of the loss itself
    h = 1e-4

    def L(s=sigma, a=alpha):
        return O.relax_diff(c.image, c.seeds, lab, 1, niter, scale.astype(np.float64), s, a,
                            target.astype(np.float64), adj, eps=EPS)["loss"]

    fd_s = (L(s=sigma + h) - L(s=sigma - h)) / (2 * h)
    fd_a = (L(a=alpha + h) - L(a=alpha - h)) / (2 * h)
    assert abs(gs - fd_s) <= 2e-3 * abs(fd_s)
    assert abs(ga - fd_a) <= 2e-3 * abs(fd_a)


def test_checkpoint_interval_does_not_change_the_gradient(ctx):
    c, lab, scale, target = _setup((12, 13, 14), 23, 1.0)
    ref = _gpu(ctx, c, lab, scale, target, 20, 25.0, 0.3, 1.0, 1)
    for ck in (3, 10, 20, 50):
        got = _gpu(ctx, c, lab, scale, target, 20, 25.0, 0.3, 1.0, ck)
        assert got[0] == pytest.approx(ref[0], rel=1e-12)
        assert np.array_equal(got[1].cpu().numpy(), ref[1].cpu().numpy())
        assert got[2] == pytest.approx(ref[2], rel=1e-9) and got[3] == pytest.approx(ref[3], rel=1e-9)


def _gpu64(ctx, c, lab, scale, target, niter, sigma, alpha, adj, ckpt=10):
    import torch
    dev = torch.device("cuda")
    return ctx.relax_adjoint(torch.from_numpy(c.image.astype(np.int16)).to(dev),
                             torch.from_numpy(c.seeds.astype(np.int32)).to(dev),
                             torch.from_numpy(lab.astype(np.int32)).to(dev), 1, niter,
                             torch.from_numpy(scale).to(dev), sigma, alpha, torch.from_numpy(target).to(dev),
                             ckpt_every=ckpt, adj_r=adj, eps=EPS, want_memb=True, fp64=True)


@pytest.mark.parametrize("shape,seed,adj,niter", [((12, 14, 16), 31, 1.0, 25), ((10, 11, 12), 32, 1.75, 12)])
def test_fp64_adjoint_meets_1e4_against_finite_differences(ctx, shape, seed, adj, niter):
    """The FP64 form of the GPU adjoint: equal to the FP64 oracle's reverse-mode gradient
    up to summation order, and within 1e-4 relative of FP64 central differences (h = 1e-4)
    of the GPU's own FP64 loss, for sigma, alpha and single entries of the scale map."""
    c, lab, scale, target = _setup(shape, seed, adj)
    sigma, alpha, h = 30.0, 0.4, 1e-4
    print(f"case shape={shape} adj={adj} niter={niter}", flush=True)
    loss, gsc, gs, ga, f = _gpu64(ctx, c, lab, scale, target, niter, sigma, alpha, adj)
    r = O.relax_diff(c.image, c.seeds, lab, 1, niter, scale.astype(np.float64), sigma, alpha,
                     target.astype(np.float64), adj, eps=EPS)
    g = gsc.cpu().numpy()
    df = np.abs(f.cpu().numpy() - r["f"]).max()
    print("fp64 vs oracle: f", df, "loss", abs(loss - r["loss"]) / abs(r["loss"]),
          "sigma", abs(gs - r["g_sigma"]) / abs(r["g_sigma"]), "alpha", abs(ga - r["g_alpha"]) / abs(r["g_alpha"]),
          "scale", np.abs(g - r["g_scale"]).max() / np.abs(r["g_scale"]).max(), flush=True)
    # last-ulp differences of exp and of contracted products compound over the iterations
    assert df <= 1e-12
    assert abs(loss - r["loss"]) <= 1e-12 * abs(r["loss"])
    assert abs(gs - r["g_sigma"]) <= 1e-10 * abs(r["g_sigma"])
    assert abs(ga - r["g_alpha"]) <= 1e-10 * abs(r["g_alpha"])
    assert np.abs(g - r["g_scale"]).max() <= 1e-10 * np.abs(r["g_scale"]).max()

    def L(s=sigma, a=alpha, sc=scale):
        return _gpu64(ctx, c, lab, sc, target, niter, s, a, adj)[0]

    fd_s = (L(s=sigma + h) - L(s=sigma - h)) / (2 * h)
    fd_a = (L(a=alpha + h) - L(a=alpha - h)) / (2 * h)
    assert abs(gs - fd_s) <= 1e-4 * abs(fd_s)
    assert abs(ga - fd_a) <= 1e-4 * abs(fd_a)
    # the scale map is an FP32 input: perturb it by a step it can represent and compare
    # with the FP64 gradient at the largest entries
    for i in np.argsort(-np.abs(g.ravel()))[:3]:
        hs = np.float32(1e-2)
        up, dn = scale.copy(), scale.copy()
        up.ravel()[i] += hs
        dn.ravel()[i] -= hs
        step = float(up.ravel()[i]) - float(dn.ravel()[i])
        fd = (L(sc=up) - L(sc=dn)) / step
        assert abs(g.ravel()[i] - fd) <= 1e-4 * abs(fd) + 1e-12


@pytest.mark.parametrize("seed", [100, 101])
def test_config5_headline_128_fp64_and_fp32(ctx, seed):
    """The differentiable-relaxation workload at its own size: a 128^3 headline phantom, the
    exact OIFT forest, 50 iterations, checkpoints every 10, BCE against the Gaussian
    (sigma 2)-blurred true mask, scale 1, alpha 0.5, RBF sigma = the image's global std.
    FP64 adjoint: equal to the FP64 oracle's reverse-mode gradient (1e-10 relative) and within
    1e-4 relative of FP64 central differences (h = 1e-4) for sigma, alpha and the three largest
    scale-map entries.  FP32 adjoint: within 2e-3 relative of the same FP64 reference."""
    import time
    import torch
    import scipy.ndimage as nd
    c = P.headline_case(128, seed=seed, niter=50)
    dev = torch.device("cuda")
    img = torch.from_numpy(c.image.astype(np.int16)).to(dev)
    seeds = torch.from_numpy(c.seeds.astype(np.int32)).to(dev)
    lab_t = torch.from_numpy(c.label_init()).to(dev)
    ctx.oift_fmax(img, seeds, lab_t, 1.0, c.per, want_cost=False)
    lab = lab_t.cpu().numpy()
    assert np.array_equal(lab, O.oift_image(c.image, c.seeds, c.label_init(), 1.0, c.per)["label"])
    target = nd.gaussian_filter(c.truth.astype(np.float32), 2.0).astype(np.float32)
    scale = np.ones(c.image.shape, np.float32)
    sigma = float(np.float32(np.sqrt(c.image.astype(np.float64).var())))
    alpha, niter, h = 0.5, 50, 1e-4
    print(f"case cfg5 seed={seed} sigma={sigma:.3f}", flush=True)
    t0 = time.time()
    r = O.relax_diff(c.image, c.seeds, lab, 1, niter, scale.astype(np.float64), sigma, alpha,
                     target.astype(np.float64), 1.0, eps=EPS)
    print(f"case cfg5 seed={seed}: FP64 oracle {time.time() - t0:.1f} s", flush=True)
    loss, gsc, gs, ga, f = _gpu64(ctx, c, lab, scale, target, niter, sigma, alpha, 1.0)
    g = gsc.cpu().numpy()
    # the loss and the sigma / alpha gradients are sums over 2M voxels (and 50 iterations),
    # reduced in another order than the oracle's sequential sum: ~1e-10 relative apart
    print(f"case cfg5 seed={seed}: fp64 vs oracle f {np.abs(f.cpu().numpy() - r['f']).max():.2e} "
          f"loss {abs(loss - r['loss']) / abs(r['loss']):.2e} sigma {abs(gs - r['g_sigma']) / abs(r['g_sigma']):.2e} "
          f"alpha {abs(ga - r['g_alpha']) / abs(r['g_alpha']):.2e}", flush=True)
    assert np.abs(f.cpu().numpy() - r["f"]).max() <= 1e-12
    assert abs(loss - r["loss"]) <= 1e-9 * abs(r["loss"])
    assert abs(gs - r["g_sigma"]) <= 1e-8 * abs(r["g_sigma"])
    assert abs(ga - r["g_alpha"]) <= 1e-8 * abs(r["g_alpha"])
    assert np.abs(g - r["g_scale"]).max() <= 1e-8 * np.abs(r["g_scale"]).max()

    def L(s=sigma, a=alpha, sc=scale):
        return _gpu64(ctx, c, lab, sc, target, niter, s, a, 1.0)[0]

    fd_s = (L(s=sigma + h) - L(s=sigma - h)) / (2 * h)
    fd_a = (L(a=alpha + h) - L(a=alpha - h)) / (2 * h)
    print(f"case cfg5 seed={seed}: sigma {gs:.6e} vs fd {fd_s:.6e}; alpha {ga:.6e} vs fd {fd_a:.6e}", flush=True)
    assert abs(gs - fd_s) <= 1e-4 * abs(fd_s)
    assert abs(ga - fd_a) <= 1e-4 * abs(fd_a)
    for i in np.argsort(-np.abs(g.ravel()))[:3]:
        hs = np.float32(1e-2)
        up, dn = scale.copy(), scale.copy()
        up.ravel()[i] += hs
        dn.ravel()[i] -= hs
        step = float(up.ravel()[i]) - float(dn.ravel()[i])
        fd = (L(sc=up) - L(sc=dn)) / step
        assert abs(g.ravel()[i] - fd) <= 1e-4 * abs(fd) + 1e-12
    # the FP32 form (the one bench.py times) against the same FP64 reference
    l32, g32, gs32, ga32, f32 = _gpu(ctx, c, lab, scale, target, niter, sigma, alpha, 1.0, 10)
    assert np.abs(f32.cpu().numpy() - r["f"]).max() <= 1e-5
    assert abs(l32 - r["loss"]) <= 1e-5 * abs(r["loss"])
    assert abs(gs32 - r["g_sigma"]) <= 2e-3 * abs(r["g_sigma"])
    assert abs(ga32 - r["g_alpha"]) <= 2e-3 * abs(r["g_alpha"])
    assert np.abs(g32.cpu().numpy().astype(np.float64) - r["g_scale"]).max() <= 2e-3 * np.abs(r["g_scale"]).max()
