"""Pins for the relaxation oracle (ORelax_1_Multi restatement) -- CPU only."""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P


@pytest.mark.parametrize("seed,adj,per", [(0, 1.0, 50.0), (1, 1.75, 50.0), (2, 1.0, -20.0)])
def test_masked_band_equals_dense_update(seed, adj, per):
    c = P.small_case((9, 10, 11), seed, per=per, adj_r=adj, top=4095)
    o = O.oift_image(c.image, c.seeds, c.label_init(), adj, per)
    for fp64 in (False, True):
        m = O.relax(c.image, c.seeds, o["label"], 7, adj, per, fp64=fp64)
        d = O.relax(c.image, c.seeds, o["label"], 7, adj, per, fp64=fp64, dense=True)
        assert np.array_equal(m["memb"], d["memb"])  # bit-identical
        assert np.array_equal(m["label"], d["label"])


def test_constant_image_one_iteration_closed_form():
    """Constant image: every arc weight is Wmax^8, so one iteration is the plain
    average of the neighbours' memberships (seeds then reset)."""
    shape = (4, 5, 6)
    zs, ys, xs = shape
    img = np.full(shape, 7, np.int32)
    lab = np.zeros(img.size, np.int32)
    lab[:img.size // 2] = 1
    seeds = np.array([0, img.size - 1], np.int64)
    lab0 = lab.reshape(shape)
    r = O.relax(img, seeds, lab0, 1, 1.0, 50.0, fp64=True)
    f_bg = (lab == 0).astype(np.float64)
    A = O.spheric(1.0)
    exp = np.empty(img.size)
    for p in range(img.size):
        x, y, z = p % xs, (p // xs) % ys, p // (xs * ys)
        vals = []
        for dx, dy, dz in A[1:]:
            qx, qy, qz = x + dx, y + dy, z + dz
            if 0 <= qx < xs and 0 <= qy < ys and 0 <= qz < zs:
                vals.append(f_bg[qx + qy * xs + qz * xs * ys])
        exp[p] = np.mean(vals)
    exp[seeds] = f_bg[seeds]
    assert np.abs(r["memb"][0].ravel() - exp).max() < 1e-12
    assert np.abs(r["memb"][1].ravel() - (1 - exp)).max() < 1e-12


@pytest.mark.parametrize("seed", range(3))
def test_relax_invariants_and_fp32_vs_fp64(seed):
    c = P.small_case((10, 12, 14), seed, per=50.0, top=4095)
    o = O.oift_image(c.image, c.seeds, c.label_init(), 1.0, 50.0)
    r32 = O.relax(c.image, c.seeds, o["label"], 20, 1.0, 50.0)
    r64 = O.relax(c.image, c.seeds, o["label"], 20, 1.0, 50.0, fp64=True)
    m = r64["memb"]
    assert m.min() >= 0.0 and m.max() <= 1.0
    assert np.abs(m.sum(0) - 1.0).max() < 1e-12
    S = c.seeds
    assert np.array_equal(m[0].ravel()[S], (c.seed_labels == 0).astype(float))
    assert np.abs(r32["memb"] - m).max() < 1e-5


def test_per_class_relax_matches_reference_when_built():
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built")
    c = P.small_case((9, 11, 13), 4, per=50.0, top=4095)
    pc = np.array([-50.0, 50.0], np.float32)
    o = O.oift_image(c.image, c.seeds, c.label_init(), 1.0, per_class=pc)
    r = O.relax(c.image, c.seeds, o["label"], 6, 1.0, per_class=pc)
    assert np.array_equal(r["label"], O.ref_relax(c.image, c.seeds, o["label"], 6, 1.0, per_class=pc))


def test_weight_identities_exhaustive():
    """(int)(d*1.5) == d + d//2 and (int)(d*0.5) == d//2 for every 16-bit d, and the
    general double formula for other polarities (Python floats are IEEE doubles)."""
    d = np.arange(0, 65536, dtype=np.int64)
    assert np.array_equal((d * 1.5).astype(np.int64), d + d // 2)
    assert np.array_equal((d * 0.5).astype(np.int64), d // 2)
    img = np.zeros((1, 1, 2), np.int32)
    for per in (50.0, 30.0, -12.5, 100.0, 0.0):
        for a, b in [(0, 1000), (4095, 17), (7, 7), (1, 0)]:
            img[0, 0] = (a, b)
            wa, wb = O.edge_weights(img, 1.0, per)
            dd = abs(a - b)
            up, dn = int(dd * (1.0 + per / 100.0)), int(dd * (1.0 - per / 100.0))
            assert wa[0, 0, 0, 0] == (up if a > b else dn if a < b else dd)   # slot 0 = +x: p=(0) -> q=(1)
            assert wb[0, 0, 0, 0] == (up if b > a else dn if b < a else dd)
            assert (wa[1:] == 0xFFFFFFFF).all() and wa[0, 0, 0, 1] == 0xFFFFFFFF  # no +y/+z/+x neighbour
