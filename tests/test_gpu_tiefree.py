"""GPU: the local-step OIFT for tie-free maps (ROIFT_PARAM_OIFT_MODE 1) over explicit
oriented weights.  Bar: labels and the f_max cost map bit-exact against the oracle's OIFT over
the same weights (orc_oift_edges: the reference's pop order with a keyed heap).  Tie-free maps
are dense ranks of (weight, bijective hash of the arc); a map with ties must fall back to the
level-synchronous kernel and stay exact."""
import numpy as np
import pytest

import oracle as O
from roift_amd import phantom as P
from test_gpu_config3 import _rank_weights

pytestmark = pytest.mark.gpu


def _run(ctx, ra, rb, seeds, lab0, adj, pol_sign=None, mode=1):
    import torch
    ctx.set_param("oift_mode", mode)
    try:
        lab = torch.from_numpy(lab0).cuda()
        cost = ctx.oift_edges(ra, rb, torch.from_numpy(seeds.astype(np.int32)).cuda(), lab, adj, pol_sign=pol_sign)
        torch.cuda.synchronize()
        return lab.cpu().numpy(), cost.cpu().numpy(), ctx.local_profile(), ctx.timings()
    finally:
        ctx.set_param("oift_mode", 0)


@pytest.mark.parametrize("shape,seed,adj,per", [((20, 24, 28), 1, 1.0, 50.0), ((18, 22, 26), 2, 1.75, 50.0),
                                                ((1, 96, 80), 3, 1.0, -30.0), ((30, 31, 33), 4, 1.75, 0.0),
                                                ((40, 44, 48), 5, 1.0, 80.0)])
def test_local_steps_bit_exact_on_rank_maps(ctx, shape, seed, adj, per):
    import torch
    c = P.small_case(shape, seed, per=per, adj_r=adj, top=4095)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    ra, rb = _rank_weights(ctx, img, per) if adj == 1.75 else _rank_weights6(ctx, img, per)
    print(f"case shape={shape} adj={adj} per={per}", flush=True)
    before = ctx.local_profile()["tie_fallbacks"]
    lab, cost, prof, t = _run(ctx, ra, rb, c.seeds, c.label_init(), adj)
    o = O.oift_edges(ra.cpu().numpy().view(np.uint32), rb.cpu().numpy().view(np.uint32), c.seeds,
                     c.label_init(), adj)
    print("profile", prof, flush=True)
    assert prof["tie_fallbacks"] == before
    assert np.array_equal(lab, o["label"])
    assert np.array_equal(cost, o["cost"])
    assert prof["rounds"] < prof["levels"]


def _rank_weights6(ctx, img, per):
    """Dense ranks of (w, hash(arc)) for every in-volume oriented arc, 6-adjacency."""
    import torch
    wa, wb, _, _ = ctx.weights(img, 1.0, per, wdtype="u32")
    w = torch.stack([wa, wb]).reshape(-1).to(torch.int64) & 0xFFFFFFFF
    valid = w != 0xFFFFFFFF
    idx = torch.arange(w.numel(), device=w.device, dtype=torch.int64)
    h = (idx * 2654435761) & 0xFFFFFFFF
    h = h ^ (h >> 15)
    key = torch.where(valid, (w << 32) | h, torch.full_like(w, 2**63 - 1))
    order = torch.argsort(key)
    rank = torch.empty_like(order)
    rank[order] = torch.arange(order.numel(), device=w.device, dtype=torch.int64)
    r = torch.where(valid, rank, torch.full_like(rank, 0xFFFFFFFF))
    r32 = torch.where(r >= 2**31, r - 2**32, r).to(torch.int32).reshape((2,) + tuple(wa.shape))
    return r32[0].contiguous(), r32[1].contiguous()


def test_local_steps_pol_sign_and_three_labels(ctx):
    """Three labels with a per-label class table (label 2 on the transposed weights)."""
    import torch
    c = P.small_case((24, 26, 28), 7, per=50.0, adj_r=1.0, top=4095)
    zs, ys, xs = c.image.shape
    lab0 = c.label_init()
    extra = 3 + 3 * xs + 3 * xs * ys
    assert extra not in set(c.seeds.tolist())
    lab0.ravel()[extra] = 2
    seeds = np.concatenate([c.seeds, [extra]]).astype(np.int32)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    ra, rb = _rank_weights6(ctx, img, 50.0)
    ps = np.array([-1, 1, -1], np.int8)
    before = ctx.local_profile()["tie_fallbacks"]
    lab, cost, prof, _ = _run(ctx, ra, rb, seeds, lab0, 1.0, pol_sign=ps)
    o = O.oift_edges(ra.cpu().numpy().view(np.uint32), rb.cpu().numpy().view(np.uint32), seeds, lab0, 1.0,
                     pol_sign=ps)
    assert prof["tie_fallbacks"] == before
    assert np.array_equal(lab, o["label"]) and np.array_equal(cost, o["cost"])


def test_local_steps_fall_back_on_ties(ctx):
    """Image-derived weights are full of ties: the run falls back and stays exact."""
    import torch
    c = P.small_case((20, 22, 24), 8, per=50.0, top=255)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    wa, wb, _, _ = ctx.weights(img, 1.0, 50.0, wdtype="u32")
    before = ctx.local_profile()["tie_fallbacks"]
    lab, cost, prof, _ = _run(ctx, wa, wb, c.seeds, c.label_init(), 1.0)
    o = O.oift_edges(wa.cpu().numpy().view(np.uint32), wb.cpu().numpy().view(np.uint32), c.seeds, c.label_init(), 1.0)
    assert prof["tie_fallbacks"] == before + 1
    assert np.array_equal(lab, o["label"]) and np.array_equal(cost, o["cost"])


def test_local_steps_config3_tie_free_96(ctx):
    """The 26-adjacency three-object volume at 96 x 96 x 64 with tie-free rank weights."""
    import time
    import torch
    c = P.multi_object_case(shape=(64, 96, 96), seed=3)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    ra, rb = _rank_weights(ctx, img, c.per)
    _run(ctx, ra, rb, c.seeds, c.label_init(), 1.75)   # warm-up (buffers, code objects)
    before = ctx.local_profile()["tie_fallbacks"]
    t0 = time.time()
    lab, cost, prof, t = _run(ctx, ra, rb, c.seeds, c.label_init(), 1.75)
    wall = time.time() - t0
    print(f"tie-free 96x96x64: {wall * 1e3:.1f} ms wall, oift kernel {ctx.kernel_timings()['oift_kernel_ms']:.1f} ms,"
          f" {prof}", flush=True)
    o = O.oift_edges(ra.cpu().numpy().view(np.uint32), rb.cpu().numpy().view(np.uint32), c.seeds, c.label_init(), 1.75)
    assert prof["tie_fallbacks"] == before
    assert np.array_equal(lab, o["label"]) and np.array_equal(cost, o["cost"])


def test_local_steps_config3_tie_free_192(ctx):
    """The same three-object 26-adjacency volume at 192 x 192 x 128 (4.7M voxels, eight times the
    case above): about 605k levels in about 10k grid rounds.  The full 384 x 384 x 256 map does not
    finish within the kernel's 120 s limit (the round count grows about tenfold per doubling of
    the edge length), so this is the largest tie-free case tested."""
    import time
    import torch
    c = P.multi_object_case(shape=(128, 192, 192), seed=3)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    ra, rb = _rank_weights(ctx, img, c.per)
    del img
    before = ctx.local_profile()["tie_fallbacks"]
    print("tie-free 192x192x128 start", flush=True)
    t0 = time.time()
    lab, cost, prof, t = _run(ctx, ra, rb, c.seeds, c.label_init(), 1.75)
    print(f"tie-free 192x192x128: {(time.time() - t0) * 1e3:.1f} ms wall, {prof}", flush=True)
    wa = ra.cpu().numpy().view(np.uint32)
    wb = rb.cpu().numpy().view(np.uint32)
    del ra, rb
    o = O.oift_edges(wa, wb, c.seeds, c.label_init(), 1.75)
    assert prof["tie_fallbacks"] == before
    assert np.array_equal(lab, o["label"]) and np.array_equal(cost, o["cost"])
