"""GPU: roift_gaussian_blur_dev is bit-exact against the oracle (and the reference's
fixtures) for int16 and int32 volumes, any pass count, ragged sizes, in place,
every plane-chunk size, the host-pointer form and the gft mirror."""
import glob
import os

import numpy as np
import pytest

import oracle as O
from roift_amd import gft, phantom as P

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _dev(a, dtype):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a).astype(dtype)).cuda()


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "blur", "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_blur_matches_reference_fixture(ctx, path):
    f = np.load(path)
    print("fixture", os.path.basename(path), f["image"].shape, flush=True)
    img = _dev(f["image"], np.int32)
    for k in (1, 2, 3):
        out = ctx.gaussian_blur(img, passes=k)
        assert np.array_equal(out.cpu().numpy(), f[f"out{k}"]), f"passes={k}"


SHAPES = [(1, 1, 1), (1, 5, 7), (3, 4, 5), (2, 17, 65), (17, 33, 70), (40, 37, 130), (9, 1, 200)]


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
@pytest.mark.parametrize("dtype", [np.int16, np.int32])
def test_blur_ragged_shapes(ctx, shape, dtype):
    print("shape", shape, dtype.__name__, flush=True)
    rng = np.random.default_rng(sum(shape))
    lo, hi = (0, 32767) if dtype == np.int16 else (-2_000_000, 2_000_000)
    img = rng.integers(lo, hi, size=shape).astype(np.int32)
    d = _dev(img, dtype)
    for k in (0, 1, 2, 3, 4):
        out = ctx.gaussian_blur(d, passes=k)
        assert out.dtype == d.dtype
        assert np.array_equal(out.cpu().numpy().astype(np.int32), O.gaussian_blur(img, k)), f"passes={k}"


def test_blur_in_place(ctx):
    img = np.random.default_rng(3).integers(0, 4000, size=(21, 30, 70)).astype(np.int32)
    for k in (1, 2, 3):
        d = _dev(img, np.int16)
        ctx.gaussian_blur(d, passes=k, out=d)
        assert np.array_equal(d.cpu().numpy().astype(np.int32), O.gaussian_blur(img, k)), f"passes={k}"


@pytest.mark.parametrize("planes", [1, 2, 3, 5, 64])
def test_blur_plane_chunks(ctx, planes):
    """Every z-chunk size gives the same result (chunk edges recompute their halo planes)."""
    img = np.random.default_rng(4).integers(0, 30000, size=(23, 20, 66)).astype(np.int32)
    want = O.gaussian_blur(img, 2)
    ctx.set_param("blur_planes", planes)
    try:
        out = ctx.gaussian_blur(_dev(img, np.int16), passes=2)
    finally:
        ctx.set_param("blur_planes", 0)
    assert np.array_equal(out.cpu().numpy().astype(np.int32), want)


def test_blur_headline_volume(ctx):
    """The 256^3 int16 ellipsoid phantom, blurred twice as the CLI does before the OIFT."""
    c = P.headline_case(256)
    print("headline", c.image.shape, flush=True)
    out = ctx.gaussian_blur(_dev(c.image, np.int16), passes=2)
    assert np.array_equal(out.cpu().numpy().astype(np.int32), O.gaussian_blur(c.image, 2))
    t = ctx.kernel_timings()["blur_ms"]
    assert t > 0


def test_blur_host_form_and_gft_mirror(ctx):
    img = np.random.default_rng(8).integers(0, 2000, size=(6, 9, 11)).astype(np.int32)
    assert np.array_equal(ctx.gaussian_blur_host(img, 2), O.gaussian_blur(img, 2))
    scn = gft.Scene32.FromArray(img)
    out = gft.Scene32.GaussianBlur(gft.Scene32.GaussianBlur(scn))
    assert np.array_equal(out.volume(), O.gaussian_blur(img, 2))


def test_blur_rejects_bad_arguments(ctx):
    import roift_amd as R
    d = _dev(np.zeros((2, 3, 4)), np.int32)
    with pytest.raises(R.RoiftError):
        ctx.gaussian_blur(d, passes=-1)
    with pytest.raises(R.RoiftError):
        ctx.gaussian_blur(_dev(np.zeros((2, 3, 4)), np.float32), passes=1)
