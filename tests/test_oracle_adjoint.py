"""Pins for the FP64 differentiable relaxation (orc_relax_diff_f64): its reverse-mode
gradients against central finite differences of its own loss, and closed forms."""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P


def _setup(shape=(7, 8, 9), seed=7, adj=1.0):
    c = P.small_case(shape, seed, per=50.0, adj_r=adj, top=255)
    lab = O.oift_image(c.image, c.seeds, c.label_init(), adj, 50.0)["label"]
    rng = np.random.default_rng(seed)
    scale = 1.0 + 0.2 * rng.standard_normal(c.image.shape)
    target = np.clip(0.1 + 0.8 * (lab > 0) + 0.05 * rng.standard_normal(c.image.shape), 0.01, 0.99)
    return c, lab, scale, target


@pytest.mark.parametrize("adj,niter", [(1.0, 12), (1.75, 7)])
def test_adjoint_matches_central_differences(adj, niter):
    c, lab, scale, target = _setup(adj=adj)
    sigma, alpha, h = 25.0, 0.35, 1e-4
    r = O.relax_diff(c.image, c.seeds, lab, 1, niter, scale, sigma, alpha, target, adj)

    def loss(sc=scale, s=sigma, a=alpha):
        return O.relax_diff(c.image, c.seeds, lab, 1, niter, sc, s, a, target, adj)["loss"]

    fd_s = (loss(s=sigma + h) - loss(s=sigma - h)) / (2 * h)
    fd_a = (loss(a=alpha + h) - loss(a=alpha - h)) / (2 * h)
    assert abs(r["g_sigma"] - fd_s) <= 1e-4 * abs(fd_s) + 1e-12
    assert abs(r["g_alpha"] - fd_a) <= 1e-4 * abs(fd_a) + 1e-12
    g = r["g_scale"]
    idx = np.argsort(-np.abs(g.ravel()))[:6]   # the most sensitive scale entries
    for i in idx:
        sp = scale.copy().ravel(); sm = scale.copy().ravel()
        sp[i] += h; sm[i] -= h
        fd = (loss(sc=sp.reshape(scale.shape)) - loss(sc=sm.reshape(scale.shape))) / (2 * h)
        assert abs(g.ravel()[i] - fd) <= 1e-4 * abs(fd) + 1e-12


def test_zero_iterations_have_zero_parameter_gradient():
    c, lab, scale, target = _setup()
    r = O.relax_diff(c.image, c.seeds, lab, 1, 0, scale, 20.0, 0.3, target)
    assert r["g_sigma"] == 0.0 and r["g_alpha"] == 0.0 and not r["g_scale"].any()
    o = np.clip(1.0 - r["f"], 1e-6, 1 - 1e-6)
    bce = -np.mean(target * np.log(o) + (1 - target) * np.log(1 - o))
    assert abs(r["loss"] - bce) < 1e-12


def test_constant_image_makes_sigma_inert():
    """All differences are zero, so every Gaussian factor is exp(0) = 1 whatever sigma is."""
    c, lab, scale, target = _setup()
    img = np.full(c.image.shape, 40, np.int32)
    r = O.relax_diff(img, c.seeds, lab, 1, 9, scale, 13.0, 0.3, target)
    assert r["g_sigma"] == 0.0
    # and with alpha = 0 and a uniform scale each update is the plain neighbour mean
    r0 = O.relax_diff(img, c.seeds, lab, 1, 1, np.ones(img.shape), 13.0, 0.0, target)
    f0 = np.where(lab == 1, 0.0, 1.0)
    zs, ys, xs = img.shape
    seeds = set(int(s) for s in c.seeds)
    for p in range(0, img.size, 37):
        x, y, z = p % xs, (p // xs) % ys, p // (xs * ys)
        if p in seeds:
            continue
        vals = [f0.ravel()[(x + dx) + (y + dy) * xs + (z + dz) * xs * ys] for dx, dy, dz in O.spheric(1.0)[1:]
                if 0 <= x + dx < xs and 0 <= y + dy < ys and 0 <= z + dz < zs]
        assert abs(r0["f"].ravel()[p] - np.mean(vals)) < 1e-12
