"""Pins for the oracle's ORelax_1 binary forms and the experiment variants
(OIFT_Multi_GradWeight, ORelax_1_Multi_GradWeight, ORelax_1_Multi_GaussRBF) -- CPU only.

Reference-produced fixtures live in tests/golden/variants (tests/golden/make_golden.py);
where oracle/_ref is built the restatements are also compared with the reference's own
code on fresh random inputs.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P

HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = sorted(glob.glob(os.path.join(HERE, "golden", "variants", "*.npz")))


def load_variant(path):
    z = np.load(path, allow_pickle=False)
    c = {k: z[k] for k in z.files}
    c["name"] = os.path.splitext(os.path.basename(path))[0]
    c["kind"] = str(c["kind"])
    for k in ("adj_r", "per", "alpha", "sigma"):
        c[k] = float(c[k])
    c["niter"] = int(c["niter"])
    lab0 = np.full(c["image"].size, -1, np.int32)
    lab0[c["seeds"]] = c["seed_labels"]
    c["label0"] = lab0.reshape(c["image"].shape)
    return c


def oracle_variant(c):
    """The oracle's answer for one fixture: (oift_label, relax_label, memberships)."""
    I = c["image"].astype(np.int32)
    if c["kind"] == "grad":
        o = O.oift_gradweight(I, c["G"], c["seeds"], c["label0"], c["adj_r"], c["alpha"], c["per"])["label"]
        r = O.relax_model(I, c["seeds"], o, c["niter"], "grad", G=c["G"], alpha=c["alpha"], adj_r=c["adj_r"],
                          per=c["per"])
        return o, r["label"], r["memb"]
    if c["kind"] == "rbf":
        o = O.oift_image(I, c["seeds"], c["label0"], c["adj_r"], c["per"])["label"]
        sigma = c["sigma"] if c["sigma"] > 0 else O.estimate_sigma(I)
        r = O.relax_model(I, c["seeds"], o, c["niter"], "rbf", sigma=sigma, adj_r=c["adj_r"], per=c["per"])
        return o, r["label"], r["memb"]
    form = str(c["form"])
    lb = c["oift_label"].astype(np.int32)
    W = c["G"] if form == "W" else None
    Wxyz = (c["Wx"], c["Wy"], c["Wz"]) if form == "Wxyz" else None
    seeds = np.zeros(0, np.int32) if form == "noseeds" else c["seeds"]
    r = O.relax_binary_w(I, seeds, lb, c["niter"], c["adj_r"], c["per"], W=W, Wxyz=Wxyz)
    return lb, r["label"], r["memb"]


@pytest.mark.parametrize("path", VARIANTS, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_reference_fixture(path):
    c = load_variant(path)
    o, r, _ = oracle_variant(c)
    assert np.array_equal(o, c["oift_label"].astype(np.int32))
    assert np.array_equal(r, c["relax_label"].astype(np.int32))


def test_variant_fixtures_present():
    kinds = {load_variant(p)["kind"] for p in VARIANTS}
    assert kinds == {"grad", "rbf", "binary"}


def _grad(img):
    g = np.zeros(img.shape, np.int64)
    I = img.astype(np.int64)
    for ax in range(3):
        g += np.abs(np.roll(I, -1, ax) - np.roll(I, 1, ax))
    return (g * max(int(I.max()), 1) // max(int(g.max()), 1)).astype(np.int32)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,adj,alpha,per", [(0, 1.0, 0.5, 50.0), (1, 1.75, 0.2, -40.0), (2, 1.0, 0.9, 0.0)])
def test_gradweight_vs_compiled_reference(seed, adj, alpha, per):
    c = P.small_case((9, 11, 13), 100 + seed, per=per, adj_r=adj, top=1023)
    I = c.image.astype(np.int32)
    G = _grad(I)
    lab0 = c.label_init()
    ref = O.ref_oift_gradweight(I, G, c.seeds, lab0, adj, alpha, per)
    mine = O.oift_gradweight(I, G, c.seeds, lab0, adj, alpha, per)["label"]
    assert np.array_equal(ref, mine)
    rr = O.ref_relax_gradweight(I, G, c.seeds, ref, 6, adj, alpha, per)
    rm = O.relax_model(I, c.seeds, ref, 6, "grad", G=G, alpha=alpha, adj_r=adj, per=per)["label"]
    assert np.array_equal(rr, rm)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("seed,adj,sigma", [(0, 1.0, 0.0), (1, 1.75, 0.0), (2, 1.0, 25.0)])
def test_rbf_vs_compiled_reference(seed, adj, sigma):
    c = P.small_case((9, 11, 13), 200 + seed, per=50.0, adj_r=adj, top=1023)
    I = c.image.astype(np.int32)
    assert O.ref_estimate_sigma(I) == O.estimate_sigma(I)
    s = sigma if sigma > 0 else O.estimate_sigma(I)
    o = O.oift_image(I, c.seeds, c.label_init(), adj, 50.0)["label"]
    assert np.array_equal(O.ref_relax_rbf(I, c.seeds, o, 7, s, adj),
                          O.relax_model(I, c.seeds, o, 7, "rbf", sigma=s, adj_r=adj)["label"])


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("form", ["image", "image26", "W", "Wxyz", "noseeds"])
def test_binary_forms_vs_compiled_reference(form):
    c = P.small_case((9, 11, 13), 300, per=50.0, top=1023)
    I = c.image.astype(np.int32)
    adj = 1.75 if form == "image26" else 1.0
    lb = (O.oift_image(I, c.seeds, c.label_init(), adj, 50.0)["label"] > 0).astype(np.int32)
    W = _grad(I) if form == "W" else None
    Wxyz = tuple(_grad(np.roll(I, k, k)) for k in (2, 1, 0)) if form == "Wxyz" else None
    seeds = None if form == "noseeds" else c.seeds
    ref = O.ref_relax_binary(I, seeds, lb, 5, adj, 50.0, W=W, Wxyz=Wxyz)
    mine = O.relax_binary_w(I, np.zeros(0, np.int32) if seeds is None else seeds, lb, 5, adj, 50.0, W=W,
                            Wxyz=Wxyz)["label"]
    assert np.array_equal(ref, mine)


def test_gradweight_with_alpha_one_equals_oift_multi_weights():
    """alpha = 1 and per = +-50 make the float GradWeight arc (int)(|d| * 1.5f) / (int)(|d| * 0.5f),
    which equals OIFT_Multi's double-precision weight for every 16-bit difference."""
    rng = np.random.default_rng(5)
    I = rng.integers(0, 4096, size=(6, 7, 8), dtype=np.int32)
    G = rng.integers(0, 4096, size=I.shape, dtype=np.int32)
    for per in (50.0, -50.0):
        wa, wb = O.gradweight_edges(I, G, 1.75, 1.0, per)
        ea, eb = O.edge_weights(I, 1.75, per)
        assert np.array_equal(wa, ea) and np.array_equal(wb, eb)


def test_gradweight_edge_symmetry():
    """The blended base weight is symmetric in p and q, so w_bg(p->q) == w_obj(q->p) still holds and
    swapping the roles of p and q swaps the two stored orientations."""
    rng = np.random.default_rng(6)
    I = rng.integers(0, 1000, size=(5, 6, 7), dtype=np.int32)
    G = rng.integers(0, 1000, size=I.shape, dtype=np.int32)
    wa, wb = O.gradweight_edges(I, G, 1.0, 0.37, 30.0)
    # slot 2 = +z neighbour: reversing z maps slot 2 of p to the arc pair of q read backwards
    wa_r, wb_r = O.gradweight_edges(I[::-1].copy(), G[::-1].copy(), 1.0, 0.37, 30.0)
    inner = np.s_[:-1]
    assert np.array_equal(wa[2][inner], wb_r[2][::-1][1:])
    assert np.array_equal(wb[2][inner], wa_r[2][::-1][1:])


def test_estimate_sigma_closed_form():
    rng = np.random.default_rng(7)
    I = rng.integers(0, 3000, size=(5, 6, 7), dtype=np.int32)
    assert abs(O.estimate_sigma(I) - float(np.float32(np.sqrt(I.astype(np.float64).var())))) <= 1e-3
    assert O.estimate_sigma(np.full((3, 3, 3), 9, np.int32)) == 1.0   # variance floor


def test_rbf_constant_image_is_plain_averaging():
    """Constant image: every RBF weight is exp(0) = 1, so one iteration is the plain neighbour mean."""
    shape = (4, 5, 6)
    zs, ys, xs = shape
    img = np.full(shape, 3, np.int32)
    lab = np.zeros(img.size, np.int32)
    lab[: img.size // 3] = 1
    seeds = np.array([0, img.size - 1], np.int64)
    r = O.relax_model(img, seeds, lab.reshape(shape), 1, "rbf", sigma=5.0, fp64=True)
    f_bg = (lab == 0).astype(np.float64)
    exp = np.empty(img.size)
    for p in range(img.size):
        x, y, z = p % xs, (p // xs) % ys, p // (xs * ys)
        vals = [f_bg[(x + dx) + (y + dy) * xs + (z + dz) * xs * ys] for dx, dy, dz in O.spheric(1.0)[1:]
                if 0 <= x + dx < xs and 0 <= y + dy < ys and 0 <= z + dz < zs]
        exp[p] = np.mean(vals)
    exp[seeds] = f_bg[seeds]
    assert np.abs(r["memb"][0].ravel() - exp).max() < 1e-12


@pytest.mark.parametrize("model", ["grad", "rbf"])
def test_model_masked_equals_dense_and_fp32_close_to_fp64(model):
    c = P.small_case((9, 10, 11), 400, per=50.0, top=1023)
    I = c.image.astype(np.int32)
    G = _grad(I)
    o = O.oift_image(I, c.seeds, c.label_init(), 1.0, 50.0)["label"]
    kw = dict(G=G, alpha=0.4, sigma=60.0, per=50.0)
    m = O.relax_model(I, c.seeds, o, 7, model, **kw)
    d = O.relax_model(I, c.seeds, o, 7, model, dense=True, **kw)
    assert np.array_equal(m["memb"], d["memb"])
    f64 = O.relax_model(I, c.seeds, o, 7, model, fp64=True, **kw)
    assert np.abs(m["memb"] - f64["memb"]).max() < 1e-5


def test_binary_constant_image_closed_form_and_dense():
    """Constant image, image form: every weight is Wmax^8, so one iteration is the neighbour mean of
    f = label/Imax; the dense update equals the band for 0/1 labels."""
    shape = (4, 5, 6)
    zs, ys, xs = shape
    img = np.full(shape, 11, np.int32)
    lab = np.zeros(img.size, np.int32)
    lab[img.size // 2:] = 1
    r = O.relax_binary_w(img, np.zeros(0, np.int32), lab.reshape(shape), 1, 1.0, 50.0)
    f = lab.astype(np.float64)
    exp = np.empty(img.size)
    for p in range(img.size):
        x, y, z = p % xs, (p // xs) % ys, p // (xs * ys)
        vals = [f[(x + dx) + (y + dy) * xs + (z + dz) * xs * ys] for dx, dy, dz in O.spheric(1.0)[1:]
                if 0 <= x + dx < xs and 0 <= y + dy < ys and 0 <= z + dz < zs]
        exp[p] = np.mean(vals)
    assert np.abs(r["memb"].ravel() - exp).max() < 1e-6
    c = P.small_case((9, 10, 11), 401, per=50.0, top=1023)
    lb = (O.oift_image(c.image, c.seeds, c.label_init(), 1.0, 50.0)["label"] > 0).astype(np.int32)
    m = O.relax_binary_w(c.image, c.seeds, lb, 6, 1.0, 50.0)
    d = O.relax_binary_w(c.image, c.seeds, lb, 6, 1.0, 50.0, dense=True)
    assert np.array_equal(m["memb"], d["memb"])


def test_planted_mistake_in_grad_model_is_caught():
    """A dropped (1-alpha) blend term would change the fixture labels: the fixtures pin the formula."""
    c = load_variant([p for p in VARIANTS if "grad_0" in p][0])
    I = c["image"].astype(np.int32)
    bad = O.oift_gradweight(I, np.zeros_like(c["G"]), c["seeds"], c["label0"], c["adj_r"], c["alpha"],
                            c["per"])["label"]
    assert not np.array_equal(bad, c["oift_label"].astype(np.int32))
