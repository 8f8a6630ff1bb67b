"""GPU: the CLI's stages around the hot path are bit-exact against the oracle and the
reference's fixtures: Sobel gradient source (+ normalisation), BFS distance from the
object seeds, boundary seeds (both forms, raster order), conditional dilation."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ST = os.path.join(HERE, "golden", "stages")


def _dev(a, dtype=np.int32):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).astype(dtype)).cuda()


def _files(prefix):
    return sorted(glob.glob(os.path.join(ST, prefix + "_*.npz")))


# ---- Sobel ----------------------------------------------------------------------------

@pytest.mark.parametrize("path", _files("sobel"), ids=lambda p: os.path.basename(p)[:-4])
def test_sobel_matches_reference_fixture(ctx, path):
    f = np.load(path)
    print("fixture", os.path.basename(path), flush=True)
    g = ctx.sobel(_dev(f["image"]), normalize=False)
    assert np.array_equal(g.cpu().numpy(), f["grad"])


@pytest.mark.parametrize("shape", [(1, 1, 1), (2, 3, 4), (3, 3, 3), (1, 20, 30), (17, 33, 70), (40, 37, 130)],
                         ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [np.int16, np.int32])
def test_sobel_ragged(ctx, shape, dtype):
    print("shape", shape, dtype.__name__, flush=True)
    rng = np.random.default_rng(sum(shape))
    lo, hi = (0, 32767) if dtype == np.int16 else (-3_000_000, 3_000_000)
    img = rng.integers(lo, hi, size=shape).astype(np.int32)
    want = O.sobel(img)
    g = ctx.sobel(_dev(img, dtype), normalize=False)
    assert np.array_equal(g.cpu().numpy(), want)
    gn = ctx.sobel(_dev(img, dtype), normalize=True)
    assert np.array_equal(gn.cpu().numpy(), O.grad_normalize(want, int(img.max())))


def test_sobel_headline_volume(ctx):
    """The experiment's gradient source on the 256^3 phantom after two blurs."""
    c = P.headline_case(256)
    blurred = O.gaussian_blur(c.image, 2)
    want = O.grad_normalize(O.sobel(blurred), int(blurred.max()))
    d = ctx.gaussian_blur(_dev(c.image, np.int16), passes=2)
    g = ctx.sobel(d, normalize=True)
    assert np.array_equal(g.cpu().numpy(), want)


# ---- distance and boundary seeds -------------------------------------------------------

@pytest.mark.parametrize("path", _files("seeds"), ids=lambda p: os.path.basename(p)[:-4])
def test_seeds_match_reference_fixture(ctx, path):
    f = np.load(path)
    lab, seeds = f["label"], f["seeds"]
    print("fixture", os.path.basename(path), lab.shape, flush=True)
    assert np.array_equal(ctx.bfs_edt(_dev(lab), _dev(seeds)).cpu().numpy(), f["edt"])
    for stride in (1, 2, 3, 5, 8):
        for geo in (False, True):
            key = f"{'geo' if geo else 'uni'}{stride}"
            dl = _dev(lab)
            new = ctx.boundary_seeds(dl, stride, geodesic=geo, seeds=_dev(seeds))
            assert np.array_equal(dl.cpu().numpy(), f[key + "_label"]), key
            assert np.array_equal(np.concatenate([seeds, new.cpu().numpy()]), f[key + "_seeds"]), key


def test_edt_and_seeds_random(ctx):
    rng = np.random.default_rng(21)
    for shape in [(1, 1, 1), (5, 1, 9), (23, 31, 45), (64, 50, 70)]:
        n = int(np.prod(shape))
        lab = np.full(n, -1, np.int32)
        seeds = rng.choice(n, size=min(n, 5), replace=False).astype(np.int32)
        lab[seeds] = rng.integers(0, 3, size=seeds.size)
        lab = lab.reshape(shape)
        print("shape", shape, flush=True)
        want_d = O.bfs_edt(lab, seeds)
        assert np.array_equal(ctx.bfs_edt(_dev(lab), _dev(seeds)).cpu().numpy(), want_d)
        for stride in (1, 4, 7):
            wl, ws = O.boundary_seeds(lab, stride)
            dl = _dev(lab)
            assert np.array_equal(ctx.boundary_seeds(dl, stride).cpu().numpy(), ws)
            assert np.array_equal(dl.cpu().numpy(), wl)
            wl, ws = O.boundary_seeds_geodesic(lab, stride, want_d)
            dl = _dev(lab)
            assert np.array_equal(ctx.boundary_seeds(dl, stride, geodesic=True, seeds=_dev(seeds)).cpu().numpy(), ws)
            assert np.array_equal(dl.cpu().numpy(), wl)


def test_boundary_seeds_capacity_error_leaves_label(ctx):
    import roift_amd as R
    lab = np.full((6, 7, 8), -1, np.int32)
    dl = _dev(lab)
    with pytest.raises(R.RoiftError):
        ctx.boundary_seeds(dl, 1, capacity=10)
    assert np.array_equal(dl.cpu().numpy(), lab)
    with pytest.raises(R.RoiftError):
        ctx.boundary_seeds(dl, 0)


# ---- conditional dilation --------------------------------------------------------------

@pytest.mark.parametrize("path", _files("dilation"), ids=lambda p: os.path.basename(p)[:-4])
def test_dilation_matches_reference_fixture(ctx, path):
    f = np.load(path)
    print("fixture", os.path.basename(path), flush=True)
    for dtype in (np.int16, np.int32):
        for pc in (0.0, 25.0, 50.0, 90.0, 100.0):
            for r in (1.0, 1.5):
                dl = _dev(f["label"])
                ctx.dilation_conditional(_dev(f["image"], dtype), dl, pc, r)
                assert np.array_equal(dl.cpu().numpy(), f[f"p{int(pc)}_r{r}"]), (dtype, pc, r)


def test_dilation_ties_signed_and_empty(ctx):
    rng = np.random.default_rng(22)
    lab = (rng.random((20, 24, 28)) < 0.05).astype(np.int32)
    for img in (np.full(lab.shape, 7, np.int32),                                    # all ties
                rng.integers(-5, 5, size=lab.shape).astype(np.int32),               # signed, many ties
                rng.integers(-(2**31), 2**31 - 1, size=lab.shape, dtype=np.int64).astype(np.int32)):
        for pc in (0.0, 33.3, 50.0, 99.9, 100.0):
            for r in (1.0, 1.75):
                want = O.dilation_conditional(img, lab, r, pc)
                dl = _dev(lab)
                n = ctx.dilation_conditional(_dev(img), dl, pc, r)
                got = dl.cpu().numpy()
                assert np.array_equal(got, want), (pc, r)
                assert n == int((got != lab).sum())
    empty = np.zeros((4, 5, 6), np.int32)
    dl = _dev(empty)
    assert ctx.dilation_conditional(_dev(np.ones_like(empty)), dl, 50.0) == 0
    assert not dl.cpu().numpy().any()


def test_dilation_headline_after_pipeline(ctx):
    """The CLI's last step on the 256^3 phantom: OIFT + 50 relaxation iterations, then
    dilation_conditional(fscn, label, 1, 50)."""
    import torch
    c = P.headline_case(256)
    img = _dev(c.image, np.int16)
    lab = torch.from_numpy(c.label_init()).cuda()
    ctx.pipeline(img, _dev(c.seeds), lab, c.niter, 1.0, c.per)
    before = lab.cpu().numpy()
    n = ctx.dilation_conditional(img, lab, 50.0)
    want = O.dilation_conditional(c.image, before, 1.0, 50.0)
    assert np.array_equal(lab.cpu().numpy(), want)
    assert n == int((want != before).sum()) and n > 0
