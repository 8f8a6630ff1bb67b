"""The oracle's gft::Scene32::GaussianBlur restatement, pinned to the reference's
outputs (golden fixtures written by the compiled reference) and to what the
mathematics fixes: the normalised kernel's closed form, constant volumes,
linearity up to rounding, and a float64 evaluation of the same sum."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "blur", "*.npz")))


def test_fixtures_present():
    assert len(FIXTURES) == 5


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[:-4] for p in FIXTURES])
def test_oracle_matches_reference_fixture(path):
    f = np.load(path)
    assert np.array_equal(O.gauss_kernel(), f["taps"])
    for k in (1, 2, 3):
        assert np.array_equal(O.gaussian_blur(f["image"], k), f[f"out{k}"]), f"passes={k}"


def test_kernel_closed_form():
    """Normalised taps: centre 1/(1+6/e), faces (1/e)/(1+6/e), summing to 1."""
    k = O.gauss_kernel().astype(np.float64)
    assert k.shape == (7,)
    e = np.exp(-1.0)
    assert abs(k[0] - 1.0 / (1.0 + 6.0 * e)) < 1e-7
    assert np.allclose(k[1:], e / (1.0 + 6.0 * e), atol=1e-7)
    assert abs(k.sum() - 1.0) < 1e-6


def test_constant_volume():
    """Interior voxels of a constant volume keep their value; a voxel with m taps
    outside the volume gets round(c * (1 - m*k1))."""
    c = 1000
    out = O.gaussian_blur(np.full((5, 6, 7), c, np.int32), 1)
    k = O.gauss_kernel().astype(np.float64)
    assert (out[1:-1, 1:-1, 1:-1] == c).all()
    assert out[2, 2, 0] == int(np.floor(c * (1 - k[1]) + 0.5))     # one face outside
    assert out[0, 0, 0] == int(np.floor(c * (1 - 3 * k[1]) + 0.5))  # a corner: three outside


def test_float64_sum_within_one():
    """The FP32 sum in tap order differs from the exact sum by far less than one unit,
    so the rounded results agree except where the exact value sits next to .5."""
    rng = np.random.default_rng(5)
    img = rng.integers(0, 4096, size=(8, 9, 10)).astype(np.int32)
    k = O.gauss_kernel().astype(np.float64)
    p = np.pad(img.astype(np.float64), 1)
    exact = k[0] * p[1:-1, 1:-1, 1:-1]
    for dz, dy, dx in ((-1, 0, 0), (0, -1, 0), (0, 0, -1), (0, 0, 1), (0, 1, 0), (1, 0, 0)):
        exact = exact + k[1] * p[1 + dz:p.shape[0] - 1 + dz, 1 + dy:p.shape[1] - 1 + dy, 1 + dx:p.shape[2] - 1 + dx]
    want = np.floor(exact + 0.5).astype(np.int64)
    got = O.gaussian_blur(img, 1).astype(np.int64)
    assert np.abs(got - want).max() <= 1
    near_half = np.abs(exact - np.floor(exact) - 0.5) < 1e-3
    assert (got[~near_half] == want[~near_half]).all()


def test_negative_values_round_away_from_zero():
    """gft's ROUND: x < 0 rounds to (int)(x - 0.5), so -v blurs to -(blur of v)."""
    rng = np.random.default_rng(6)
    img = rng.integers(0, 30000, size=(6, 7, 8)).astype(np.int32)
    assert np.array_equal(O.gaussian_blur(-img, 2), -O.gaussian_blur(img, 2))


def test_zero_passes_is_identity_and_passes_compose():
    rng = np.random.default_rng(7)
    img = rng.integers(0, 30000, size=(4, 5, 6)).astype(np.int32)
    assert np.array_equal(O.gaussian_blur(img, 0), img)
    assert np.array_equal(O.gaussian_blur(O.gaussian_blur(img, 1), 2), O.gaussian_blur(img, 3))


def test_planted_mistake_is_caught():
    """A restatement that renormalised the kernel at the border would not match the
    reference fixture: the fixtures discriminate the border rule."""
    f = np.load(os.path.join(HERE, "golden", "blur", "ell.npz"))
    img = f["image"]
    ones = O.gaussian_blur(np.full(img.shape, 1 << 16, np.int32), 1).astype(np.float64) / (1 << 16)
    renorm = np.floor(O.gaussian_blur(img, 1) / ones + 0.5).astype(np.int32)
    assert not np.array_equal(renorm, f["out1"])


def test_round_identity_used_by_the_kernel():
    """The kernel rounds with trunc(x + copysign(0.5 - 2^-25, x)) in FP32; it must equal
    gft's ROUND ((x < 0) ? (int)(x - 0.5) : (int)(x + 0.5), in double) for every float.
    Checked on every float in [1, 4), every float within 4096 ulps of each power of two up
    to 2^31 and of each half-integer below 2^12, and 2M random floats, both signs."""
    h = np.nextafter(np.float32(0.5), np.float32(0))
    parts = [np.arange(np.float32(1).view(np.uint32), np.float32(4).view(np.uint32), dtype=np.uint32).view(np.float32)]
    for e in range(-3, 32):
        b = np.float32(2.0 ** e).view(np.uint32)
        parts.append(np.arange(b - 4096, b + 4096, dtype=np.uint32).view(np.float32))
    for n in range(0, 4096, 7):
        b = np.float32(n + 0.5).view(np.uint32)
        parts.append(np.arange(b - 64, b + 64, dtype=np.uint32).view(np.float32))
    rng = np.random.default_rng(9)
    parts.append((rng.random(2_000_000) * 2.0 ** rng.integers(-2, 31, 2_000_000)).astype(np.float32))
    x = np.concatenate(parts)
    x = np.concatenate([x, -x])
    got = np.trunc(x + np.copysign(h, x)).astype(np.float64)
    xd = x.astype(np.float64)
    want = np.where(xd < 0, np.trunc(xd - 0.5), np.trunc(xd + 0.5))
    assert np.array_equal(got, want)
