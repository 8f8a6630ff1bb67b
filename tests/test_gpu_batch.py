"""GPU: the batched pipeline (roift_batch_pipeline_dev) gives every volume exactly the
result of the plain per-volume pipeline, including on HIP-graph replays interleaved with
other eager work, and the OIFT part equals the oracle."""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P

pytestmark = pytest.mark.gpu


def _upload(cases):
    import torch
    imgs = [torch.from_numpy(c.image.astype(np.int16)).cuda() for c in cases]
    seeds = [torch.from_numpy(c.seeds.astype(np.int32)).cuda() for c in cases]
    labs = [torch.from_numpy(c.label_init()).cuda() for c in cases]
    return imgs, seeds, labs


def _plain(ctx, cases, niter):
    import torch
    imgs, seeds, labs = _upload(cases)
    out = []
    for img, s, lab in zip(imgs, seeds, labs):
        cost = torch.empty_like(lab)
        ctx.pipeline(img, s, lab, niter, 1.0, cases[0].per, cost=cost)
        out.append((lab.cpu().numpy(), cost.cpu().numpy()))
    return out


def _batched(ctx, cases, niter, use_graph):
    imgs, seeds, labs = _upload(cases)
    costs = ctx.batch_pipeline(imgs, seeds, labs, niter, cases[0].per, use_graph=use_graph)
    return [(l.cpu().numpy(), c.cpu().numpy()) for l, c in zip(labs, costs)]


def _same(a, b):
    for i, ((la, ca), (lb, cb)) in enumerate(zip(a, b)):
        assert np.array_equal(la, lb), f"labels differ in volume {i}"
        assert np.array_equal(ca, cb), f"costs differ in volume {i}"


@pytest.mark.parametrize("use_graph", [False, True])
def test_batch_equals_plain_small(ctx, use_graph):
    cases = P.batch_cases(count=7, seed=11, lo=10, hi=36, niter=6)
    print("volumes", [c.image.shape for c in cases], flush=True)
    plain = _plain(ctx, cases, 6)
    _same(_batched(ctx, cases, 6, use_graph), plain)


def test_batch_oift_equals_oracle(ctx):
    cases = P.batch_cases(count=4, seed=12, lo=10, hi=28, niter=0)
    got = _batched(ctx, cases, 0, False)
    for c, (lab, cost) in zip(cases, got):
        o = O.oift_image(c.image, c.seeds, c.label_init(), 1.0, c.per)
        assert np.array_equal(lab, o["label"]) and np.array_equal(cost, o["cost"])


def test_batch_graph_replays_with_eager_work_between(ctx):
    """Replay the cached graph several times with other eager launches and syncs in between;
    every replay must equal the eager per-volume result."""
    import torch
    cases = P.batch_cases(count=5, seed=13, lo=12, hi=30, niter=5)
    plain = _plain(ctx, cases, 5)
    for rep in range(3):
        _same(_batched(ctx, cases, 5, True), plain)
        # unrelated eager work on the same context and stream
        other = P.small_case((14, 15, 16), 50 + rep)
        lab = torch.from_numpy(other.label_init()).cuda()
        ctx.pipeline(torch.from_numpy(other.image.astype(np.int16)).cuda(),
                     torch.from_numpy(other.seeds.astype(np.int32)).cuda(), lab, 3, 1.0, other.per)
        torch.cuda.synchronize()


def test_batch_three_labels_falls_back_per_volume(ctx):
    """Seed labels {0, 1, 2} take the per-volume relaxation inside the batch call."""
    cases = P.batch_cases(count=3, seed=14, lo=12, hi=26, niter=4)
    for c in cases:   # relabel the first internal seed as label 2
        c.seed_labels = c.seed_labels.copy()
        c.seed_labels[0] = 2
    plain = _plain(ctx, cases, 4)
    _same(_batched(ctx, cases, 4, True), plain)


def test_batch_rejects_bad_seed(ctx):
    import roift_amd as R
    cases = P.batch_cases(count=2, seed=15, lo=10, hi=14, niter=2)
    imgs, seeds, labs = _upload(cases)
    seeds[1][0] = cases[1].image.size + 5   # outside its own volume
    with pytest.raises(R.RoiftError):
        ctx.batch_pipeline(imgs, seeds, labs, 2, cases[0].per, use_graph=False)


def test_batch_headline_config_64_volumes(ctx):
    """64 uneven volumes, dims log-uniform in [64, 192], 30 relaxation iterations: batch == plain."""
    cases = P.batch_cases(count=64, seed=4, lo=64, hi=192, niter=30)
    print("total voxels", sum(c.image.size for c in cases), flush=True)
    plain = _plain(ctx, cases, 30)
    _same(_batched(ctx, cases, 30, True), plain)


def test_batch_graph_object_label_change_rebuilds(ctx):
    """Two same-shaped graph calls whose object label changes (1, then 2): the second must
    not replay launches captured with the first call's label (each equals the eager path)."""
    cases = P.batch_cases(count=4, seed=16, lo=12, hi=24, niter=4)
    for lab_obj in (1, 2, 1):
        for c in cases:
            c.seed_labels = np.where(c.seed_labels > 0, lab_obj, 0).astype(c.seed_labels.dtype)
        eager = _batched(ctx, cases, 4, False)
        graphed = _batched(ctx, cases, 4, True)
        _same(graphed, eager)
        for lab, _ in graphed:
            assert set(np.unique(lab)) <= {0, lab_obj}, f"labels {np.unique(lab)} with object label {lab_obj}"


def test_batch_graph_replays_and_reports(ctx):
    """The first call of a key runs eagerly and captures the whole pipeline (cooperative OIFT
    kernel included) as one graph; the calls that repeat the key replay it."""
    cases = P.batch_cases(count=6, seed=21, lo=12, hi=30, niter=5)
    plain = _plain(ctx, cases, 5)
    c0, r0, e0, _ = ctx.batch_graph_info()
    for rep in range(3):
        print("call", rep, flush=True)
        _same(_batched(ctx, cases, 5, True), plain)
    c1, r1, e1, err = ctx.batch_graph_info()
    print("graph info", (c1 - c0, r1 - r0, e1 - e0, err), flush=True)
    assert err == "", err
    assert c1 == c0 + 1 and r1 == r0 + 2 and e1 == e0


def _flip(cases):
    """Same shapes and seed counts, new contents: images mirrored in x (seeds kept)."""
    import copy
    out = copy.deepcopy(cases)
    for c in out:
        c.image = np.ascontiguousarray(c.image[:, :, ::-1])
    return out


def test_batch_graph_replay_takes_new_contents(ctx):
    """Replays read the inputs' current contents: two data sets of the same shapes alternate."""
    a = P.batch_cases(count=4, seed=23, lo=12, hi=28, niter=4)
    b = _flip(a)
    pa, pb = _plain(ctx, a, 4), _plain(ctx, b, 4)
    _, r0, e0, _ = ctx.batch_graph_info()
    for cases, plain in ((a, pa), (b, pb), (a, pa), (b, pb)):
        _same(_batched(ctx, cases, 4, True), plain)
    _, r1, e1, _ = ctx.batch_graph_info()
    assert r1 - r0 == 3 and e1 == e0


def test_batch_graph_bad_seed_on_replay(ctx):
    """A seed outside its volume found after a replay gives the eager path's error, and the
    library leaves the label buffer as it was passed in."""
    import roift_amd as R
    import torch
    cases = P.batch_cases(count=3, seed=24, lo=12, hi=20, niter=3)
    _batched(ctx, cases, 3, True)   # eager run and capture
    imgs, seeds, labs = _upload(cases)
    seeds[2][0] = cases[2].image.size + 7
    with pytest.raises(R.RoiftError, match="out of its volume"):
        ctx.batch_pipeline(imgs, seeds, labs, 3, cases[0].per, use_graph=True)
    lab_buf = ctx._batch_bufs[3]
    want = torch.cat([l.reshape(-1) for l in labs])
    assert torch.equal(lab_buf, want), "labels changed by a failed replay"
    # and the next valid call still replays correctly
    _same(_batched(ctx, cases, 3, True), _plain(ctx, cases, 3))


def test_batch_graph_three_labels_on_replay_reruns(ctx):
    """Seed labels {0, 1, 2} arriving at a graph captured for {0, 1}: the call reruns eagerly
    (per-volume relaxation) and equals the plain pipeline."""
    cases = P.batch_cases(count=3, seed=25, lo=12, hi=24, niter=4)
    _batched(ctx, cases, 4, True)   # eager run and capture with labels {0, 1}
    for c in cases:
        c.seed_labels = c.seed_labels.copy()
        c.seed_labels[0] = 2
    _, _, e0, _ = ctx.batch_graph_info()
    _same(_batched(ctx, cases, 4, True), _plain(ctx, cases, 4))
    _, _, e1, _ = ctx.batch_graph_info()
    assert e1 == e0 + 1


def test_batch_graph_wide_weights_on_replay_rerun(ctx):
    """int32 volumes whose arc weights outgrow 16 bits after the capture: the replay sees the
    builder's overflow flag and reruns eagerly with 32-bit weights."""
    import torch
    cases = P.batch_cases(count=3, seed=26, lo=12, hi=22, niter=3)

    def run(scale):
        imgs = [torch.from_numpy(c.image.astype(np.int32) * scale).cuda() for c in cases]
        seeds = [torch.from_numpy(c.seeds.astype(np.int32)).cuda() for c in cases]
        labs = [torch.from_numpy(c.label_init()).cuda() for c in cases]
        costs = ctx.batch_pipeline(imgs, seeds, labs, 3, cases[0].per, use_graph=True)
        got = [(l.cpu().numpy(), k.cpu().numpy()) for l, k in zip(labs, costs)]
        ref = []
        for img, s, c in zip(imgs, seeds, cases):
            lab = torch.from_numpy(c.label_init()).cuda()
            cost = torch.empty_like(lab)
            ctx.pipeline(img, s, lab, 3, 1.0, c.per, cost=cost)
            ref.append((lab.cpu().numpy(), cost.cpu().numpy()))
        return got, ref

    run(1)   # eager run and capture
    _, _, e0, _ = ctx.batch_graph_info()
    got, ref = run(40)   # intensities up to ~160k: weights beyond 16 bits
    _same(got, ref)
    _, _, e1, _ = ctx.batch_graph_info()
    assert e1 == e0 + 1
