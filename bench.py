"""Benchmark: OIFT (f_max) + 50-iteration fuzzy relaxation on a 256^3 int16 volume, 6-adjacency.

One step = one pass of the hot path over one synthetic volume already resident
in HBM: oriented weights + exact OIFT + 50 FP32 relaxation iterations + final
labels (roift_pipeline_dev), preceded by restoring the seed-only label map.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 256] [--no-cpu] [--no-batch] [--no-adjoint]
                    [--no-config3]

Prints one JSON line.  Its "blur" object times the CLI's preprocessing (two Gaussian-blur
passes, one fused launch) on the same volume, outside the step.  Its "batched" object times the 64-volume batch (volumes of
64-192 voxels a side) as 64 plain calls and as one batched call; its "adjoint" object
times the differentiable relaxation's forward + checkpointed adjoint on 16 x 128^3; its
"config3" object times the 384x384x256 26-adjacency three-object volume (OIFT + 100
Gaussian-RBF relaxation iterations); "hbm_copy_measured" is a timed 1 GiB device copy, the
measured bandwidth beside the 8 TB/s specification the rooflines use.  The
CPU baseline runs the reference's own
OIFT_Multi + ORelax_1_Multi (compiled from its sources into oracle/_ref) on one
core over a bounded sample, or the oracle restatement when that build is absent.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic HBM bytes per voxel (int16 image, uint16 weights, FP32 single-channel memberships):
ALG = dict(weights=2 + 3 * 2 * 2, oift=12 + 4 + 4 + 4, relax_iter=2 + 4 + 4, labels=4,
           # the fused OIFT set-up (k_prep): image read, the packed 32-byte arc slot, and the
           # OIFT state it initialises (word 8, run 8, flags 1, two stamps 8, cost 4, key stamps 8,
           # label read 4) -- it replaces the planar weights, k_pack_arcs and k_init
           prep=2 + 32 + 8 + 8 + 1 + 8 + 4 + 8 + 4,
           blur2=2 + 2,   # both Gaussian-blur passes fused: read the int16 image once, write it once
           sobel=2 + 4)   # Sobel: read the int16 image, write the int32 gradient
# HBM bytes per voxel per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE on this
# workload, scaled by the same counters on a 1 GiB copy (profiles/pmc_traffic_headline256.md).
TRAFFIC = {"k_oift_persistent": 1503.80, "k_relax2_march": 10.51, "k_weights": 16.13, "k_prep": 80.06,
           "k_blur": 4.71, "k_sobel": 7.34}
# What limits each kernel, from its profiles (rocprofv3 counters and A/B timings):
LIMITED_BY = {
    "k_oift_persistent": "latency: ~2,100 grid barriers (2.4 us of atomics plus 1.6-3.5 us of agent-scope "
                         "L2 write-back/invalidate each) and the flood's dependent chains (1.7 us per "
                         "single-wave hop); HBM traffic is a by-product of the barriers' write-back/invalidate",
    "k_relax2_march": "VALU and HBM together: 86% VALU-busy at ~88 lane-instructions per voxel (the eighth "
                      "power, the division, six neighbours) with ~4 TB/s of algorithmic bytes",
    "k_weights": "HBM",
    "k_prep": "HBM (stores): 69 of its 75 bytes per voxel are writes",
    "k_blur": "VALU: the reference's FP32 tap order, ~50 instructions per output voxel",
    "k_sobel": "L1/TA: 54 taps per voxel shared 4-wide",
}


def copy_peak(dev, nbytes=1 << 30, reps=5):
    """Measured HBM bandwidth of a plain device copy (torch's copy kernel, 1 GiB read + 1 GiB
    written), best of `reps`, beside the 8 TB/s specification the rooflines are priced on."""
    import torch
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        dst.copy_(src)
        e1.record()
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        best = ms if best is None else min(best, ms)
    del src, dst
    gbs = 2 * nbytes / (best / 1e3) / 1e9
    return {"GB/s": round(gbs, 1), "bytes_moved": 2 * nbytes, "ms": round(best, 4),
            "frac_of_spec": round(gbs / HBM_PEAK_GBS, 4)}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(size=256, runs=1, second=(128, 3)):
    """The reference's own OIFT_Multi + ORelax_1_Multi (oracle/_ref, compiled -O3 from its
    sources; the oracle restatement where that build is absent) on the headline phantom, on
    one host core: the process is pinned to one core with sched_setaffinity for the timed
    calls.  `runs` runs at `size` (the median is the baseline) and, as a second figure,
    second[1] runs at second[0]^3."""
    import oracle as O
    from roift_amd import phantom as P
    kind = "reference" if O.ref_available() else "port"
    allowed = sorted(os.sched_getaffinity(0))
    core = allowed[-1]
    os.sched_setaffinity(0, {core})

    def timed(n, k):
        c = P.headline_case(n)
        lab0 = c.label_init()
        ts = []
        for _ in range(k):
            if kind == "reference":
                _, t1, t2 = O.ref_pipeline(c.image, c.seeds, lab0, c.niter, 1.0, c.per)
                ts.append(t1 + t2)
            else:
                t0 = time.perf_counter()
                o = O.oift_image(c.image, c.seeds, lab0, 1.0, c.per)
                O.relax(c.image, c.seeds, o["label"], c.niter, 1.0, c.per)
                ts.append(time.perf_counter() - t0)
        return float(np.median(ts)), ts, c.image.size

    try:
        t, ts, nvox = timed(size, runs)
        t2, ts2, nvox2 = timed(*second)
    finally:
        os.sched_setaffinity(0, set(allowed))
    impl = "reference gft sources -O3" if kind == "reference" else "oracle restatement -O3"
    return {"value": round(nvox / t / 1e6, 4), "unit": "Mvoxels/s", "cores": 1, "kind": kind,
            "sample": f"{size}^3 headline phantom (the bench workload, same generator), OIFT_Multi + "
                      f"{50}-iteration ORelax_1_Multi, {runs} run(s) of {', '.join(f'{x:.2f}' for x in ts)} s "
                      f"(median), pinned to host core {core} of {len(allowed)} allowed "
                      f"({os.cpu_count()} online), CPU {_cpu_model()}, {impl}",
            "second_figure": {"size": f"{second[0]}^3", "runs": second[1], "median_s": round(t2, 3),
                              "Mvoxels/s": round(nvox2 / t2 / 1e6, 4)}}


def reset_params(ctx):
    for name in ("oift_blocks", "flood_budget", "flood_mode", "relax_tile"):
        ctx.set_param(name, 0)


def batched_bench(ctx, reps=3):
    """64 uneven volumes (dims log-uniform in [64, 192]), OIFT + 30 relaxation iterations each:
    64 plain pipeline calls one after another vs one batched call (HIP-graph relaxation).
    Best of `reps` runs each: the batched OIFT's time varies by about 15% from run to run
    (64 volumes' fronts race for the same workgroups)."""
    import torch
    from roift_amd import phantom as P
    from roift_amd.tune import autotune_oift
    cases = P.batch_cases(count=64, seed=4, lo=64, hi=192, niter=30)
    nvox = int(sum(c.image.size for c in cases))
    imgs = [torch.from_numpy(c.image.astype(np.int16)).cuda() for c in cases]
    seeds = [torch.from_numpy(c.seeds.astype(np.int32)).cuda() for c in cases]
    lab0 = [torch.from_numpy(c.label_init()).cuda() for c in cases]
    labs = [l.clone() for l in lab0]
    per, niter = cases[0].per, cases[0].niter

    def plain():
        for img, s, l0, l in zip(imgs, seeds, lab0, labs):
            l.copy_(l0)
            ctx.pipeline(img, s, l, niter, 1.0, per)

    def batched():
        for l0, l in zip(lab0, labs):
            l.copy_(l0)
        ctx.batch_pipeline(imgs, seeds, labs, niter, per, use_graph=True, want_cost=False)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        best = None
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        return best

    reset_params(ctx)
    t_plain = timed(plain)
    tune = autotune_oift(ctx, "batch64x[64,192]^3", batched, reps=3)
    t_batch = timed(batched)
    graph = ctx.batch_graph_info()
    reset_params(ctx)
    return {"workload": "64 volumes, dims log-uniform in [64,192], oift_fmax + 30 relax iterations, 6-adj",
            "voxels": nvox, "plain_mvox_s": round(nvox / t_plain / 1e6, 2),
            "batched_mvox_s": round(nvox / t_batch / 1e6, 2), "speedup": round(t_plain / t_batch, 2),
            "plain_ms": round(t_plain * 1e3, 1), "batched_ms": round(t_batch * 1e3, 1),
            "tuned": tune["setting"], "graph": {"captures": graph[0], "replays": graph[1], "eager_reruns": graph[2],
                                               "capture_error": graph[3] or None}}


def config3_bench(ctx, reps=2):
    """384 x 384 x 256 voxels (z, y, x = 256, 384, 384), 26-adjacency, three objects: the exact
    OIFT followed by the configuration's 100 Gaussian-RBF relaxation iterations (sigma 0.8);
    best of `reps` after one warm-up."""
    import torch
    from roift_amd import phantom as P
    c = P.multi_object_case()
    n = int(c.image.size)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    seeds = torch.from_numpy(c.seeds.astype(np.int32)).cuda()
    lab0 = torch.from_numpy(c.label_init()).cuda()
    best, ko, kr = None, None, None
    for r in range(reps + 1):
        lab = lab0.clone()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.oift_fmax(img, seeds, lab, 1.75, c.per, want_cost=False)
        oift_k = ctx.kernel_timings()["oift_kernel_ms"]
        ctx.relax_rbf(img, seeds, lab, 100, 0.8, 1.75, want_memb=False)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        if r and (best is None or dt < best):
            best, ko, kr = dt, oift_k, ctx.timings()["relax_ms"]
    return {"workload": "3 objects 384x384x256 (z 256), 26-adj, oift_fmax + 100 Gaussian-RBF relax iterations "
                        "(sigma 0.8)", "voxels": n, "mvox_s": round(n / best / 1e6, 2), "ms": round(best * 1e3, 1),
            "oift_kernel_ms": round(ko, 1), "relax_ms": round(kr, 1)}


def adjoint_bench(ctx, count=16, size=128, niter=50, ckpt=10):
    """16 volumes of 128^3: exact OIFT, then the differentiable relaxation (50 iterations)
    forward + checkpointed adjoint (every 10) against the Gaussian(sigma=2)-blurred true mask;
    learnable per-voxel scale 1.0, alpha 0.5, RBF sigma = the image's global std: sigma is in
    raw intensity units, and with intensities in [0, 4095] a sigma of 1.0 would make nearly
    every FP32 weight exp(-d^2 / sigma^2) underflow to 0 (|d| >= 11)."""
    import torch
    import scipy.ndimage as nd
    from roift_amd import phantom as P
    vols = []
    for i in range(count):
        c = P.headline_case(size, seed=100 + i, niter=niter)
        img = torch.from_numpy(c.image.astype(np.int16)).cuda()
        seeds = torch.from_numpy(c.seeds.astype(np.int32)).cuda()
        lab = torch.from_numpy(c.label_init()).cuda()
        ctx.oift_fmax(img, seeds, lab, 1.0, c.per, want_cost=False)
        target = torch.from_numpy(nd.gaussian_filter(c.truth.astype(np.float32), 2.0)).cuda()
        scale = torch.ones(c.image.shape, dtype=torch.float32, device="cuda")
        sigma = float(np.float32(np.sqrt(c.image.astype(np.float64).var())))
        vols.append((img, seeds, lab, scale, sigma, target))

    def run(fp64=False, nv=count):
        out = []
        for img, seeds, lab, scale, sigma, target in vols[:nv]:
            out.append(ctx.relax_adjoint(img, seeds, lab, 1, niter, scale, sigma, 0.5, target, ckpt_every=ckpt,
                                         fp64=fp64)[0])
        return out

    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    losses = run()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    nvox = count * size ** 3
    # the FP64 form (the one checked against FP64 finite differences to 1e-4) on 4 of them
    n64 = min(4, count)
    run(True, n64)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    run(True, n64)
    torch.cuda.synchronize()
    dt64 = time.perf_counter() - t0
    return {"workload": f"{count} x {size}^3, differentiable relax {niter} iterations fwd + adjoint, "
                        f"checkpoint every {ckpt}, BCE vs blurred truth, sigma = image std, alpha 0.5, scale 1",
            "voxels": nvox, "fwd_bwd_mvox_s": round(nvox / dt / 1e6, 2), "ms": round(dt * 1e3, 1),
            "mean_loss": round(float(np.mean(losses)), 6),
            "fp64_fwd_bwd_mvox_s": round(n64 * size ** 3 / dt64 / 1e6, 2)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=256)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--probe-copy", action="store_true", help="also run a 1 GiB device copy (profiler scaling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-volume batched-vs-plain measurement")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the 16 x 128^3 relaxation-adjoint measurement")
    ap.add_argument("--no-config3", action="store_true", help="skip the 384x384x256 26-adjacency measurement")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the headline pipeline's launch parameters (grid, budget, flood mode, relaxation "
                         "tile) before the timed steps (default: the library's defaults)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "batches would shard across GPUs by volume, but no multi-GPU path is built.",
              file=sys.stderr)
        sys.exit(2)

    import torch
    import roift_amd as R
    from roift_amd import phantom as P

    case = P.headline_case(args.size)
    dev = torch.device("cuda:0")
    ctx = R.Context(0)
    img = torch.from_numpy(case.image.astype(np.int16)).to(dev)
    seeds = torch.from_numpy(case.seeds.astype(np.int32)).to(dev)
    lab0 = torch.from_numpy(case.label_init()).to(dev)
    lab = torch.empty_like(lab0)
    cost = torch.empty_like(lab0)
    memb = torch.empty((2,) + tuple(img.shape), dtype=torch.float32, device=dev)
    n = img.numel()

    def step():
        lab.copy_(lab0)
        ctx.pipeline(img, seeds, lab, case.niter, 1.0, case.per, cost=cost, memb=memb)

    tuned = None
    if args.tune:
        from roift_amd.tune import STAGES_PLAIN, autotune
        tuned = autotune(ctx, f"headline{args.size}", step, STAGES_PLAIN, reps=3)["setting"]
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    stage = dict(weights_ms=[], oift_ms=[], relax_ms=[], oift_steps=0, flood_rounds=0)
    kern = dict(weights_ms=[], oift_kernel_ms=[], relax_loop_ms=[])
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        t = ctx.timings()
        for k in ("weights_ms", "oift_ms", "relax_ms"):
            stage[k].append(t[k])
        stage["oift_steps"], stage["flood_rounds"] = int(t["oift_steps"]), int(t["flood_rounds"])
        kt = ctx.kernel_timings()
        for k in kern:
            kern[k].append(kt[k])
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms = dt / args.steps * 1e3
    mvox = n * args.steps / dt / 1e6

    # Rooflines from device events recorded on the library's stream around each kernel:
    # algorithmic bytes per launch / average launch duration.
    def roof(name, alg_per_voxel, launch_ms):
        # "bound" names the roof the kernel is priced against (HBM: nothing here is a
        # contraction); "limited_by" is what the profiles show actually limits it
        gbs = alg_per_voxel * n / (launch_ms / 1e3) / 1e9
        out = {"kernel": name, "bound": "hbm", "limited_by": LIMITED_BY.get(name), "achieved": round(gbs, 2),
               "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 5),
               "launch_ms": round(launch_ms, 4), "alg_bytes_per_launch": alg_per_voxel * n}
        tr = TRAFFIC.get(name)
        out["traffic"] = int(tr * n) if tr is not None else None
        return out

    # one more pass with the OIFT's detailed counters on (flood critical path, batch sizes,
    # single-wave hops); they cost a few device atomics per flood round, so the timed steps
    # above run without them
    ctx.set_param("oift_profile", 1)
    step()
    torch.cuda.synchronize()
    prof = ctx.oift_profile()
    ctx.set_param("oift_profile", 0)
    # the CLI's preprocessing (oiftrelax.cpp:368-375): two Gaussian-blur passes, outside the step
    blurred = torch.empty_like(img)
    blur_ms = []
    for i in range(6):
        ctx.gaussian_blur(img, passes=2, out=blurred)
        if i:
            blur_ms.append(ctx.kernel_timings()["blur_ms"])
    blur_k = float(np.mean(blur_ms))
    # the CLI's other stages on the same volume (device time of each call, median of 5):
    # Sobel gradient source (+ normalisation), boundary seeds (stride 8, both forms), and the
    # final dilation_conditional(fscn, label, 1, 50) of the step's labels
    def stage_time(fn):
        t = []
        for _ in range(5):
            fn()
            t.append(ctx.kernel_timings()["stage_ms"])
        return float(np.median(t))
    grad = torch.empty(tuple(img.shape), dtype=torch.int32, device=dev)
    sobel_k = stage_time(lambda: ctx.sobel(blurred, normalize=False, out=grad))
    sobel_norm_ms = stage_time(lambda: ctx.sobel(blurred, normalize=True, out=grad))
    seed_lab = torch.empty_like(lab0)
    obj_seeds = seeds[torch.from_numpy(case.seed_labels > 0).to(dev)]
    def seeds_call(geo):
        seed_lab.copy_(lab0)
        ctx.boundary_seeds(seed_lab, 8, geodesic=geo, seeds=obj_seeds)
    seeds_uni_ms = stage_time(lambda: seeds_call(False))
    seeds_geo_ms = stage_time(lambda: seeds_call(True))
    dil_lab = torch.empty_like(lab)
    def dil_call():
        dil_lab.copy_(lab)
        ctx.dilation_conditional(blurred, dil_lab, 50.0)
    dil_ms = stage_time(dil_call)
    oift_k = float(np.mean(kern["oift_kernel_ms"]))
    relax_k = float(np.mean(kern["relax_loop_ms"])) / case.niter
    weights_k = float(np.mean(kern["weights_ms"]))
    pipe_bytes = (ALG["weights"] + ALG["oift"] + ALG["relax_iter"] * case.niter + ALG["labels"]) * n
    res = {
        "metric": "Mvoxels/s OIFT+relax at 256^3 6-adj (and % HBM roofline); batched vs plain",
        "value": round(mvox, 3),
        "unit": "Mvoxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16/u64 int (OIFT) + f32 (relax)",
        "data": "synthetic",
        "config": {"workload": case.meta["workload"], "shape": list(img.shape), "image": "int16",
                   "adjacency": 6, "alpha": 0.5, "relax_iters": case.niter, "seeds": int(seeds.numel()),
                   "launch_params": tuned or "library defaults"},
        "stages_ms": {k: round(float(np.mean(v)), 3) for k, v in stage.items() if isinstance(v, list)},
        "kernel_ms": {"prep_weights_and_state": round(weights_k, 4), "oift_persistent": round(oift_k, 3),
                      "relax_iteration": round(relax_k, 4)},
        "oift_phase_ms": {k: v[0] for k, v in prof.items() if k in ctx.PHASES},
        "oift_phase_count": {k: v[1] for k, v in prof.items() if k in ctx.PHASES},
        "oift_flood_items": prof["flood_items"][0],
        "oift_touched_slots": prof["touched_slots"][0],
        "oift_flood_batches": {"sum": prof["flood_batches"][0], "max_per_workgroup": prof["flood_batches"][1],
                               "critical_path": prof["flood_critical_batches"][0],
                               "critical_path_items": prof["flood_critical_batches"][1],
                               "single_wave_hops": prof["flood_wave_hops"][0],
                               "single_wave_hop_ms_summed": prof["flood_wave_hops"][1]},
        "oift_level_steps": stage["oift_steps"],
        "oift_flood_rounds": stage["flood_rounds"],
        "pipeline_roofline": {"bound": "hbm", "algorithmic_bytes_per_voxel": pipe_bytes // n,
                              "achieved": round(pipe_bytes / (dt / args.steps) / 1e9, 2),
                              "peak": HBM_PEAK_GBS, "unit": "GB/s",
                              "frac": round(pipe_bytes / (dt / args.steps) / 1e9 / HBM_PEAK_GBS, 5)},
        # the dominant kernel (most of the step's time): the persistent OIFT kernel
        "roofline": roof("k_oift_persistent", ALG["oift"], oift_k),
        "other_rooflines": [roof("k_relax2_march", ALG["relax_iter"], relax_k),
                            roof("k_prep", ALG["prep"], weights_k),
                            roof("k_blur", ALG["blur2"], blur_k),
                            roof("k_sobel", ALG["sobel"], sobel_k)],
        "blur": {"passes": 2, "launches": 1, "ms": round(blur_k, 4),
                 "Mvoxels_per_s": round(n / (blur_k / 1e3) / 1e6, 1)},
        "cli_stages_ms": {"sobel": round(sobel_k, 4), "sobel_normalized": round(sobel_norm_ms, 4),
                          "boundary_seeds_stride8": round(seeds_uni_ms, 4),
                          "boundary_seeds_geodesic8": round(seeds_geo_ms, 4),
                          "dilation_conditional_p50": round(dil_ms, 4)},
    }
    if not args.no_batch:
        res["batched"] = batched_bench(ctx)
    if not args.no_adjoint:
        res["adjoint"] = adjoint_bench(ctx)
    if not args.no_config3:
        res["config3"] = config3_bench(ctx)
    res["hbm_copy_measured"] = copy_peak(dev)
    if args.probe_copy:
        # a plain copy of known size, for scaling the profiler's FETCH_SIZE / WRITE_SIZE
        src = torch.empty(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(size=args.size)
        res["speedup_vs_cpu"] = round(res["value"] / res["cpu_baseline"]["value"], 1)
    torch.cuda.synchronize()
    ctx.close()   # release the device memory before the interpreter (and a profiler) shut down
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
