"""GPU parity of the ORelax_1 binary forms and the experiment variants against the
reference-produced fixtures (tests/golden/variants) and the oracle.

Bars: OIFT_Multi_GradWeight labels and costs bit-exact; GradWeight / RBF memberships within
1e-5 of the FP64 oracle, labels equal except where the FP64 top two are within 2e-5;
ORelax_1 memberships equal to the FP32 oracle (same operation order) within 1e-6.
"""
import glob
import os

import numpy as np
import pytest

import oracle as O
from test_oracle_variants import load_variant, oracle_variant

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
VARIANTS = sorted(glob.glob(os.path.join(HERE, "golden", "variants", "*.npz")))
TOL = 1e-5


def _dev(a, dtype=None):
    import torch
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda()


def _near_tie(m64, tol=2e-5):
    s = np.sort(m64, axis=0)
    return (s[-1] - s[-2]) < tol


@pytest.mark.parametrize("path", VARIANTS, ids=lambda p: os.path.basename(p)[:-4])
def test_variant_matches_fixture_and_oracle(ctx, path):
    import torch
    c = load_variant(path)
    print(f"case {c['name']} kind={c['kind']} shape={c['image'].shape}", flush=True)
    img = _dev(c["image"].astype(np.int16))
    seeds = _dev(c["seeds"].astype(np.int32))
    I = c["image"].astype(np.int32)
    if c["kind"] == "grad":
        G = _dev(c["G"], torch.int32)
        lab = _dev(c["label0"].astype(np.int32))
        cost = ctx.oift_gradweight(img, G, seeds, lab, c["alpha"], c["adj_r"], c["per"])
        wa, wb = O.gradweight_edges(I, c["G"], c["adj_r"], c["alpha"], c["per"])
        o = O.oift_edges(wa, wb, c["seeds"], c["label0"], c["adj_r"], pol_sign=np.array([-1], np.int8))
        assert np.array_equal(lab.cpu().numpy(), c["oift_label"].astype(np.int32))
        assert np.array_equal(cost.cpu().numpy(), o["cost"])
        memb = ctx.relax_gradweight(img, G, seeds, lab, c["niter"], c["alpha"], c["adj_r"], c["per"])
        r64 = O.relax_model(I, c["seeds"], o["label"], c["niter"], "grad", G=c["G"], alpha=c["alpha"],
                            adj_r=c["adj_r"], per=c["per"], fp64=True)
    elif c["kind"] == "rbf":
        lab = _dev(c["label0"].astype(np.int32))
        ctx.oift_fmax(img, seeds, lab, c["adj_r"], c["per"])
        assert np.array_equal(lab.cpu().numpy(), c["oift_label"].astype(np.int32))
        memb, used = ctx.relax_rbf(img, seeds, lab, c["niter"], c["sigma"], c["adj_r"])
        assert np.float32(used) == c["sigma_used"]
        r64 = O.relax_model(I, c["seeds"], c["oift_label"].astype(np.int32), c["niter"], "rbf",
                            sigma=float(c["sigma_used"]), adj_r=c["adj_r"], fp64=True)
    else:
        form = str(c["form"])
        lab = _dev(c["oift_label"].astype(np.int32))
        W = _dev(c["G"], torch.int32) if form == "W" else None
        Wxyz = tuple(_dev(c[k], torch.int32) for k in ("Wx", "Wy", "Wz")) if form == "Wxyz" else None
        s = None if form == "noseeds" else seeds
        f = ctx.relax_binary(img, s, lab, c["niter"], c["adj_r"], c["per"], W=W, Wxyz=Wxyz)
        _, rl, rm = oracle_variant(c)
        assert np.array_equal(lab.cpu().numpy(), c["relax_label"].astype(np.int32))
        assert np.abs(f.cpu().numpy() - rm).max() <= 1e-6
        return
    m = memb.cpu().numpy().astype(np.float64)
    assert m.shape == r64["memb"].shape
    assert np.abs(m - r64["memb"]).max() <= TOL
    near = _near_tie(r64["memb"])
    got = lab.cpu().numpy()
    assert np.array_equal(got[~near], c["relax_label"].astype(np.int32)[~near])


def test_relax_without_seeds_runs_orelax_1(ctx):
    """ORelax_1_Multi with an empty S is ORelax_1 (gft_ift.cpp:8678-8682)."""
    import torch
    c = load_variant([p for p in VARIANTS if p.endswith("binary_4.npz")][0])
    img = _dev(c["image"].astype(np.int16))
    lab_a = _dev(c["oift_label"].astype(np.int32))
    lab_b = lab_a.clone()
    empty = torch.zeros(0, dtype=torch.int32, device="cuda")
    ctx.relax(img, empty, lab_a, c["niter"], c["adj_r"], c["per"], want_memb=False)
    ctx.relax_binary(img, None, lab_b, c["niter"], c["adj_r"], c["per"], want_memb=False)
    assert torch.equal(lab_a, lab_b)
    assert np.array_equal(lab_a.cpu().numpy(), c["relax_label"].astype(np.int32))


def test_rbf_auto_sigma_matches_estimate(ctx):
    import torch
    rng = np.random.default_rng(3)
    I = rng.integers(0, 4000, size=(20, 21, 22)).astype(np.int16)
    lab = torch.zeros(I.shape, dtype=torch.int32, device="cuda")
    seeds = torch.tensor([0, 5], dtype=torch.int32, device="cuda")
    _, used = ctx.relax_rbf(_dev(I), seeds, lab, 0, 0.0)
    assert np.float32(used) == np.float32(O.estimate_sigma(I.astype(np.int32)))


def test_gft_mirror_variants(ctx):
    """The gft mirror's ORelax_1(W, ...) and exp.* reproduce the reference fixtures."""
    from roift_amd import gft
    c = load_variant([p for p in VARIANTS if p.endswith("binary_2.npz")][0])
    A = gft.AdjRel3.Spheric(c["adj_r"])
    scn = gft.Scene32.FromArray(c["image"].astype(np.int32))
    W = gft.Scene32.FromArray(c["G"])
    label = gft.Scene32.FromArray(c["oift_label"].astype(np.int32))
    S = np.concatenate([[c["seeds"].size], c["seeds"]]).astype(np.int32)
    gft.ift.ORelax_1(W, A, scn, c["per"], S, label, c["niter"])
    assert np.array_equal(label.volume(), c["relax_label"].astype(np.int32))
    g = load_variant([p for p in VARIANTS if p.endswith("grad_0.npz")][0])
    A = gft.AdjRel3.Spheric(g["adj_r"])
    scn = gft.Scene32.FromArray(g["image"].astype(np.int32))
    G = gft.Scene32.FromArray(g["G"])
    label = gft.Scene32.FromArray(g["label0"])
    S = np.concatenate([[g["seeds"].size], g["seeds"]]).astype(np.int32)
    gft.exp.OIFT_Multi_GradWeight(A, scn, G, g["per"], g["alpha"], S, label)
    assert np.array_equal(label.volume(), g["oift_label"].astype(np.int32))
    r = gft.exp.estimate_sigma(scn)
    assert np.float32(r) == np.float32(O.estimate_sigma(g["image"].astype(np.int32)))
