"""The oracle's restatements of the CLI's stages around the hot path (Sobel gradient
source and its normalisation, BFS distance from the object seeds, boundary seeds,
conditional dilation), pinned to fixtures written by the compiled reference and to
what the mathematics fixes: ramps and constants for Sobel, brute-force Manhattan
distances, brute-force predicates over the faces, a brute-force stable selection."""
import glob
import os

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
ST = os.path.join(HERE, "golden", "stages")


def _files(prefix):
    return sorted(glob.glob(os.path.join(ST, prefix + "_*.npz")))


def _id(p):
    return os.path.basename(p)[:-4]


def test_fixtures_present():
    assert len(_files("sobel")) == 4 and len(_files("seeds")) == 3 and len(_files("dilation")) == 2


# ---- Sobel ----------------------------------------------------------------------------

@pytest.mark.parametrize("path", _files("sobel"), ids=_id)
def test_sobel_matches_reference_fixture(path):
    f = np.load(path)
    assert np.array_equal(O.sobel(f["image"]), f["grad"])


def test_sobel_ramp_and_constant():
    """I = a*x: away from the wrapping x columns only the x-difference sum is non-zero,
    16 * 2a >> 4 = 2a, so G = ROUND(sqrt((2a)^2 / 3)); a constant volume gives 0."""
    zs, ys, xs = 6, 7, 9
    for a in (1, 3, 50, 700):
        img = np.broadcast_to(a * np.arange(xs, dtype=np.int32), (zs, ys, xs)).copy()
        g = O.sobel(img)
        want = int(np.floor(np.sqrt((2.0 * a) ** 2 / 3.0) + 0.5))
        assert (g[1:-1, 1:-1, 1:-1] == want).all()
    assert not O.sobel(np.full((5, 6, 7), 123, np.int32)).any()


def test_sobel_range_is_the_linear_interior():
    """Only linear indices in [1 + xs + xs*ys, N - 1 - xs - xs*ys) are written."""
    rng = np.random.default_rng(1)
    img = rng.integers(0, 1000, size=(5, 6, 7)).astype(np.int32)
    g = O.sobel(img).ravel()
    di = 1 + 7 + 42
    assert not g[:di].any() and not g[g.size - di:].any() and g[di:g.size - di].any()


def test_grad_normalize_closed_form():
    rng = np.random.default_rng(2)
    G = rng.integers(0, 5000, size=(4, 5, 6)).astype(np.int32)
    for wmax in (1, 255, 4095, 1 << 20):
        want = (G.astype(np.int64) * wmax // int(G.max())).astype(np.int32)
        assert np.array_equal(O.grad_normalize(G, wmax), want)
    assert np.array_equal(O.grad_normalize(G, 0), G)
    assert np.array_equal(O.grad_normalize(np.zeros_like(G), 10), np.zeros_like(G))


# ---- distance and boundary seeds -------------------------------------------------------

def _manhattan(label, seeds):
    zs, ys, xs = label.shape
    src = [int(p) for p in seeds if 0 <= p < label.size and label.ravel()[p] > 0]
    out = np.full(label.size, 2**31 - 1, np.int64)
    z, y, x = np.meshgrid(np.arange(zs), np.arange(ys), np.arange(xs), indexing="ij")
    for p in src:
        pz, py, px = p // (xs * ys), (p // xs) % ys, p % xs
        out = np.minimum(out, (np.abs(z - pz) + np.abs(y - py) + np.abs(x - px)).ravel())
    return out.reshape(label.shape).astype(np.int64)


@pytest.mark.parametrize("path", _files("seeds"), ids=_id)
def test_edt_matches_reference_and_manhattan(path):
    f = np.load(path)
    d = O.bfs_edt(f["label"], f["seeds"])
    assert np.array_equal(d, f["edt"])
    assert np.array_equal(d.astype(np.int64), _manhattan(f["label"], f["seeds"]))


def test_edt_without_object_seeds_is_unreached():
    lab = np.full((3, 4, 5), -1, np.int32)
    lab.ravel()[7] = 0
    assert (O.bfs_edt(lab, np.array([7], np.int32)) == 2**31 - 1).all()


def _faces(shape):
    zs, ys, xs = shape
    z, y, x = np.meshgrid(np.arange(zs), np.arange(ys), np.arange(xs), indexing="ij")
    return z, y, x, (x == 0) | (x == xs - 1) | (y == 0) | (y == ys - 1) | (z == 0) | (z == zs - 1)


def _brute_uniform(label, stride):
    z, y, x, face = _faces(label.shape)
    sel = face & (x % stride == 0) & (y % stride == 0) & (z % stride == 0) & (label == -1)
    return np.flatnonzero(sel.ravel())


def _brute_geodesic(label, stride, edt):
    z, y, x, face = _faces(label.shape)
    d = edt.astype(np.int64)
    es = np.where(d <= stride, max(1, stride // 2), np.where(d <= 3 * stride, stride, 2 * stride))
    sel = face & (label == -1) & (d > 2 * stride) & ((x % es == 0) | (y % es == 0) | (z % es == 0))
    return np.flatnonzero(sel.ravel())


@pytest.mark.parametrize("path", _files("seeds"), ids=_id)
def test_boundary_seeds_match_reference_and_brute_force(path):
    f = np.load(path)
    lab, seeds = f["label"], f["seeds"]
    for stride in (1, 2, 3, 5, 8):
        # the fixtures hold the reference's whole seed array: the caller's seeds, then the new ones
        l1, s1 = O.boundary_seeds(lab, stride)
        assert np.array_equal(l1, f[f"uni{stride}_label"])
        assert np.array_equal(np.concatenate([seeds, s1]), f[f"uni{stride}_seeds"])
        assert np.array_equal(s1, _brute_uniform(lab, stride))
        l2, s2 = O.boundary_seeds_geodesic(lab, stride, O.bfs_edt(lab, seeds))
        assert np.array_equal(l2, f[f"geo{stride}_label"])
        assert np.array_equal(np.concatenate([seeds, s2]), f[f"geo{stride}_seeds"])
        assert np.array_equal(s2, _brute_geodesic(lab, stride, f["edt"]))
        assert (l2.ravel()[s2] == 0).all() and (lab.ravel()[s2] == -1).all()


# ---- conditional dilation --------------------------------------------------------------

def _brute_dilation(img, lab, radius, pc):
    sph = O.spheric(radius)
    zs, ys, xs = lab.shape
    border = np.zeros(lab.shape, bool)
    for dx, dy, dz in sph[1:]:
        sh = np.zeros(lab.shape, bool)
        src = lab == 1
        zsl = slice(max(0, -dz), zs - max(0, dz)), slice(max(0, dz), zs - max(0, -dz))
        ysl = slice(max(0, -dy), ys - max(0, dy)), slice(max(0, dy), ys - max(0, -dy))
        xsl = slice(max(0, -dx), xs - max(0, dx)), slice(max(0, dx), xs - max(0, -dx))
        sh[zsl[1], ysl[1], xsl[1]] = src[zsl[0], ysl[0], xsl[0]]
        border |= sh & (lab == 0)
    pos = np.flatnonzero(border.ravel())
    out = lab.copy().ravel()
    if pos.size:
        order = np.argsort(img.ravel()[pos], kind="stable")
        idx = np.float32((pc / 100.0) * (pos.size - 1))
        out[pos[order[: int(idx) + 1]]] = 1
    return out.reshape(lab.shape)


@pytest.mark.parametrize("path", _files("dilation"), ids=_id)
def test_dilation_matches_reference_and_brute_force(path):
    f = np.load(path)
    for pc in (0.0, 25.0, 50.0, 90.0, 100.0):
        for r in (1.0, 1.5):
            got = O.dilation_conditional(f["image"], f["label"], r, pc)
            assert np.array_equal(got, f[f"p{int(pc)}_r{r}"]), (pc, r)
            assert np.array_equal(got, _brute_dilation(f["image"], f["label"], r, pc)), (pc, r)


def test_dilation_ties_keep_raster_order():
    """A constant image: every border voxel ties, so the kept ones are the first in
    raster order; an unstable order (reversed ties) would keep others."""
    lab = np.zeros((5, 6, 7), np.int32)
    lab[2, 3, 3] = 1
    img = np.full(lab.shape, 9, np.int32)
    got = O.dilation_conditional(img, lab, 1.0, 50.0)
    border = [2 * 42 + 2 * 7 + 3, 1 * 42 + 3 * 7 + 3, 2 * 42 + 3 * 7 + 2, 2 * 42 + 3 * 7 + 4,
              2 * 42 + 4 * 7 + 3, 3 * 42 + 3 * 7 + 3]
    keep = sorted(border)[: int(np.float32(0.5 * 5)) + 1]
    assert sorted(np.flatnonzero(got.ravel() != lab.ravel()).tolist()) == keep


def test_dilation_empty_border_changes_nothing():
    lab = np.zeros((3, 4, 5), np.int32)
    assert np.array_equal(O.dilation_conditional(np.ones_like(lab), lab, 1.0, 50.0), lab)
