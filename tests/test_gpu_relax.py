"""GPU parity of the fuzzy relaxation (ORelax_1_Multi / _PerClass).

Bar: FP32 memberships within 1e-5 max-abs of the FP64 oracle; thresholded labels
identical except voxels whose FP64 membership lies within 1e-5 of the decision
boundary.
"""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P
from conftest import golden_cases

pytestmark = pytest.mark.gpu

TOL = 1e-5


def dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def gpu_relax(ctx, image, seeds, label, niter, adj, per, per_class=None):
    import torch
    lab = dev(label.astype(np.int32))
    memb = ctx.relax(dev(image.astype(np.int16)), dev(seeds.astype(np.int32)), lab, niter, adj, per, per_class)
    torch.cuda.synchronize()
    return lab.cpu().numpy(), memb.cpu().numpy()


def check(ctx, image, seeds, oift_label, niter, adj, per, per_class=None):
    lab, memb = gpu_relax(ctx, image, seeds, oift_label, niter, adj, per, per_class)
    r = O.relax(image, seeds, oift_label, niter, adj, per, per_class, fp64=True)
    assert memb.shape == r["memb"].shape
    err = np.abs(memb.astype(np.float64) - r["memb"]).max()
    assert err <= TOL, err
    # voxels where the FP64 argmax is within the tolerance of a tie may flip
    top2 = np.sort(r["memb"], axis=0)[-2:]
    near = (top2[1] - top2[0]) < 2 * TOL
    assert np.array_equal(lab[~near], r["label"][~near])
    return err


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["name"])
def test_relax_matches_reference_golden(ctx, case):
    oift = case["oift_label"].astype(np.int32)
    lab, memb = gpu_relax(ctx, case["image"], case["seeds"], oift, case["niter"], case["adj_r"], case["per"],
                          case["per_class"])
    r = O.relax(case["image"], case["seeds"], oift, case["niter"], case["adj_r"], case["per"], case["per_class"],
                fp64=True)
    top2 = np.sort(r["memb"], axis=0)[-2:]
    near = (top2[1] - top2[0]) < 2 * TOL
    assert np.array_equal(lab[~near], case["relax_label"].astype(np.int32)[~near])
    assert np.abs(memb.astype(np.float64) - r["memb"]).max() <= TOL


@pytest.mark.parametrize("shape,seed,adj,per,niter", [((12, 13, 14), 0, 1.0, 50.0, 10), ((9, 21, 17), 1, 1.0, -50.0, 25),
                                                       ((11, 12, 13), 2, 1.75, 50.0, 12), ((1, 40, 50), 3, 1.0, 50.0, 30),
                                                       ((14, 14, 14), 4, 1.0, 0.0, 5), ((10, 10, 10), 5, 1.0, 50.0, 0),
                                                       ((10, 12, 9), 6, 1.0, 100.0, 1), ((31, 1, 45), 9, 1.0, 50.0, 20),
                                                       ((17, 23, 1), 10, 1.0, -50.0, 20), ((1, 30, 40), 11, 1.75, 50.0, 15)])
def test_relax_matches_fp64_oracle(ctx, shape, seed, adj, per, niter):
    c = P.small_case(shape, seed, per=per, adj_r=adj)
    o = O.oift_image(c.image, c.seeds, c.label_init(), adj, per)
    check(ctx, c.image, c.seeds, o["label"], niter, adj, per)


def test_relax_three_classes_and_per_class(ctx):
    c = P.small_case((16, 18, 20), 8, per=50.0)
    zs, ys, xs = c.image.shape
    extra = np.array([3 + 4 * xs + 5 * xs * ys, 16 + 14 * xs + 12 * xs * ys], np.int64)
    keep = ~np.isin(c.seeds, extra)
    seeds = np.concatenate([c.seeds[keep], extra])
    labels = np.concatenate([c.seed_labels[keep], [2, 5]]).astype(np.int32)
    lab0 = np.full(c.image.size, -1, np.int32)
    lab0[seeds] = labels
    lab0 = lab0.reshape(c.image.shape)
    o = O.oift_image(c.image, seeds, lab0, 1.0, 50.0)
    check(ctx, c.image, seeds, o["label"], 15, 1.0, 50.0)
    pc = np.array([-50, 50, 20, 0, 0, 70], np.float32)
    check(ctx, c.image, seeds, o["label"], 15, 1.0, 0.0, pc)
    pc2 = np.array([-40, 40], np.float32)   # two classes, per-class form
    o2 = O.oift_image(c.image, c.seeds, c.label_init(), 1.0, per_class=pc2)
    check(ctx, c.image, c.seeds, o2["label"], 15, 1.0, 0.0, pc2)


@pytest.mark.parametrize("n", [64, 128, 256])
def test_relax_headline_phantom_50_iterations(ctx, n):
    """The headline workload at 64^3, 128^3 and its full 256^3: the z-marching kernel's 50
    iterations within 1e-5 of the FP64 oracle, labels equal outside near-ties."""
    import time
    c = P.headline_case(n)
    t0 = time.time()
    o = O.oift_image(c.image, c.seeds, c.label_init(), 1.0, 50.0)
    print(f"case headline n={n}: oracle OIFT {time.time() - t0:.1f} s", flush=True)
    t0 = time.time()
    err = check(ctx, c.image, c.seeds, o["label"], 50, 1.0, 50.0)
    print(f"case headline n={n}: relax max|err| vs FP64 {err:.2e} ({time.time() - t0:.1f} s)", flush=True)


def test_pipeline_equals_separate_stages(ctx):
    import torch
    c = P.headline_case(48, niter=20)
    img = dev(c.image)
    seeds = dev(c.seeds.astype(np.int32))
    lab = dev(c.label_init())
    cost = torch.empty_like(lab)
    memb = torch.empty((2,) + c.image.shape, dtype=torch.float32, device="cuda:0")
    ctx.pipeline(img, seeds, lab, 20, 1.0, 50.0, cost=cost, memb=memb)
    torch.cuda.synchronize()
    lab2 = dev(c.label_init())
    cost2 = ctx.oift_fmax(img, seeds, lab2, 1.0, 50.0)
    memb2 = ctx.relax(img, seeds, lab2, 20, 1.0, 50.0)
    assert torch.equal(cost, cost2) and torch.equal(lab, lab2) and torch.equal(memb, memb2)
    t = ctx.timings()
    assert t["relax_ms"] > 0
