"""Writes the golden fixtures in this directory from the compiled reference.

Runs gft::ift::OIFT_Multi[_PerClass] and ORelax_1_Multi[_PerClass] of the
reference, and (into variants/) its ORelax_1 binary forms and the experiment
variants OIFT_Multi_GradWeight / ORelax_1_Multi_GradWeight / ORelax_1_Multi_GaussRBF, and (into
blur/) gft::Scene32::GaussianBlur applied 1-3 times, and (into stages/) SobelFilter,
compute_bfs_edt, the CLI's and the geodesic experiment's boundary seeds and the CLI's
dilation_conditional (oracle/_ref/libroift_ref.so, built by `make -C oracle ref` from the
reference sources) on small deterministic inputs and stores inputs and
outputs as .npz.  The reference holds no fixtures of its own, so these are
the reference-produced known answers the oracle and the GPU path are pinned
to.  Re-run only if the inputs below change:

    python tests/golden/make_golden.py          # everything
    python tests/golden/make_golden.py blur     # only blur/
    python tests/golden/make_golden.py stages   # only stages/
    python tests/golden/make_golden.py cli      # only cli/ (the reference CLI run on files)
    python tests/golden/make_golden.py oiftw    # only oiftw/ (OIFT(A, ...), OIFT(W, ...), OIFT(Wx, Wy, Wz, ...))
    python tests/golden/make_golden.py spacing  # only spacing/ (anisotropic voxel spacing)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from roift_amd import phantom as P  # noqa: E402


def border_seeds_2d(n):
    idx = [x + y * n for y in range(n) for x in range(n) if x in (0, n - 1) or y in (0, n - 1)]
    return np.array(idx, np.int64)


def cases():
    # 1. the reference's CPU-runnable 2-D case: 512^2, 4-adjacency (zsize=1), per=+50, 10 iterations
    c = P.plumbing_case_2d()
    yield "plumbing2d_512", dict(image=c.image, seeds=c.seeds, seed_labels=c.seed_labels, adj_r=1.0, per=50.0,
                                 per_class=None, niter=10)
    # 2. small 3-D ellipsoid phantoms, 6-adjacency, natural ties
    for i, (shape, seed, per) in enumerate([((12, 14, 16), 11, 50.0), ((16, 18, 20), 12, 50.0),
                                            ((10, 20, 24), 13, -30.0), ((14, 14, 14), 14, 0.0)]):
        c = P.small_case(shape, seed, per=per, top=255)
        yield f"ell6_{i}", dict(image=c.image, seeds=c.seeds, seed_labels=c.seed_labels, adj_r=1.0, per=per,
                                per_class=None, niter=8)
    # 3. 26-adjacency
    c = P.small_case((12, 12, 14), 21, per=50.0, top=255)
    yield "ell26_0", dict(image=c.image, seeds=c.seeds, seed_labels=c.seed_labels, adj_r=1.75, per=50.0,
                          per_class=None, niter=6)
    # 4. three labels, per-class polarity {-50, +50, +50}; the third label seeds a corner blob
    c = P.small_case((14, 16, 18), 31, per=50.0, top=255)
    zs, ys, xs = c.image.shape
    extra = np.array([3 + 3 * xs + 3 * xs * ys, 4 + 3 * xs + 3 * xs * ys], np.int64)
    keep = ~np.isin(c.seeds, extra)
    seeds = np.concatenate([c.seeds[keep][:27], extra, c.seeds[keep][27:]])
    labels = np.concatenate([c.seed_labels[keep][:27], [2, 2], c.seed_labels[keep][27:]]).astype(np.int32)
    yield "perclass3", dict(image=c.image, seeds=seeds, seed_labels=labels, adj_r=1.0, per=0.0,
                            per_class=np.array([-50.0, 50.0, 50.0], np.float32), niter=5)
    # 5. constant image (all weights zero): labels follow the FIFO order of S
    img = np.full((6, 7, 8), 100, np.int16)
    s = np.array([0, 8 * 7 * 6 - 1, 3 + 3 * 8 + 3 * 56], np.int64)
    yield "constant", dict(image=img, seeds=s, seed_labels=np.array([0, 0, 1], np.int32), adj_r=1.0, per=50.0,
                           per_class=None, niter=3)


def grad_image(img):
    """Deterministic gradient-magnitude image scaled to [0, max I] as the GradWeight experiment
    scales its Sobel magnitude (exp_gradient_weight.cpp:593-607); central differences here."""
    I = img.astype(np.int64)
    g = np.zeros_like(I)
    for ax in range(3):
        if I.shape[ax] > 2:
            d = np.abs(np.roll(I, -1, ax) - np.roll(I, 1, ax))
            g += d
    gmax, imax = int(g.max()), int(I.max())
    if gmax > 0 and imax > 0:
        g = g * imax // gmax
    return g.astype(np.int32)


def variant_cases():
    """Experiment variants and the ORelax_1 binary forms on small phantoms."""
    for i, (shape, seed, adj, alpha) in enumerate([((12, 14, 16), 41, 1.0, 0.5), ((10, 12, 14), 42, 1.75, 0.3),
                                                   ((14, 12, 10), 43, 1.0, 0.0)]):
        c = P.small_case(shape, seed, per=50.0, top=255)
        yield f"grad_{i}", dict(kind="grad", image=c.image, seeds=c.seeds, seed_labels=c.seed_labels, adj_r=adj,
                                per=50.0, alpha=alpha, sigma=0.0, niter=6)
    for i, (shape, seed, adj, sig) in enumerate([((12, 14, 16), 51, 1.0, 0.0), ((10, 12, 14), 52, 1.75, 0.0),
                                                 ((12, 12, 12), 53, 1.0, 40.0)]):
        c = P.small_case(shape, seed, per=50.0, top=255)
        yield f"rbf_{i}", dict(kind="rbf", image=c.image, seeds=c.seeds, seed_labels=c.seed_labels, adj_r=adj,
                               per=50.0, alpha=0.0, sigma=sig, niter=8)
    for i, (shape, seed, adj, form) in enumerate([((12, 14, 16), 61, 1.0, "image"), ((10, 12, 14), 62, 1.75, "image"),
                                                  ((12, 14, 16), 63, 1.0, "W"), ((12, 12, 14), 64, 1.0, "Wxyz"),
                                                  ((12, 14, 16), 65, 1.0, "noseeds")]):
        c = P.small_case(shape, seed, per=50.0, top=255)
        yield f"binary_{i}", dict(kind="binary", form=form, image=c.image, seeds=c.seeds, seed_labels=c.seed_labels,
                                  adj_r=adj, per=50.0, alpha=0.0, sigma=0.0, niter=5)


def write_variants():
    out_dir = os.path.join(HERE, "variants")
    os.makedirs(out_dir, exist_ok=True)
    for name, c in variant_cases():
        img = np.ascontiguousarray(c["image"], np.int32)
        lab0 = np.full(img.size, -1, np.int32)
        lab0[c["seeds"]] = c["seed_labels"]
        lab0 = lab0.reshape(img.shape)
        G = grad_image(img)
        extra = {}
        if c["kind"] == "grad":
            oift = O.ref_oift_gradweight(img, G, c["seeds"], lab0, c["adj_r"], c["alpha"], c["per"])
            relax = O.ref_relax_gradweight(img, G, c["seeds"], oift, c["niter"], c["adj_r"], c["alpha"], c["per"])
        else:
            oift = O.ref_oift(img, c["seeds"], lab0, c["adj_r"], c["per"])
            if c["kind"] == "rbf":
                sigma = c["sigma"] if c["sigma"] > 0 else O.ref_estimate_sigma(img)
                extra["sigma_used"] = np.float32(sigma)
                relax = O.ref_relax_rbf(img, c["seeds"], oift, c["niter"], sigma, c["adj_r"], c["per"])
            else:
                form = c["form"]
                lb = (oift > 0).astype(np.int32)
                W = G if form == "W" else None
                Wxyz = None
                if form == "Wxyz":
                    Wxyz = tuple(grad_image(np.roll(img, k, k)) for k in (2, 1, 0))
                    extra.update(Wx=Wxyz[0], Wy=Wxyz[1], Wz=Wxyz[2])
                seeds = None if form == "noseeds" else c["seeds"]
                oift = lb
                relax = O.ref_relax_binary(img, seeds, lb, c["niter"], c["adj_r"], c["per"], W=W, Wxyz=Wxyz)
                extra["form"] = np.array(form)
        np.savez_compressed(
            os.path.join(out_dir, f"{name}.npz"), kind=np.array(c["kind"]), image=img.astype(np.int16), G=G,
            seeds=c["seeds"].astype(np.int32), seed_labels=c["seed_labels"].astype(np.int32),
            adj_r=np.float32(c["adj_r"]), per=np.float32(c["per"]), alpha=np.float32(c["alpha"]),
            sigma=np.float32(c["sigma"]), niter=np.int32(c["niter"]), oift_label=oift.astype(np.int8),
            relax_label=relax.astype(np.int8), **extra)
        print(name, img.shape, "obj voxels", int((oift > 0).sum()), int((relax > 0).sum()))


def oiftw_cases():
    """The two-label OIFT(A, ...) and the per-voxel weight forms OIFT(W, A, ...) and
    OIFT(Wx, Wy, Wz, ...) on small phantoms: (name, form, shape, seed, adj, per, W scale,
    image offset).  W is a deterministic gradient image times the scale; a scale of 150
    pushes the arc weights past 16 bits, an offset makes the image partly negative (the
    W forms compare intensities only)."""
    yield "a_0", "A", (12, 14, 16), 71, 1.0, 50.0, 1, 0
    yield "a_1", "A", (10, 12, 14), 72, 1.75, -30.0, 1, 0
    yield "w_0", "W", (12, 14, 16), 73, 1.0, 50.0, 1, 0
    yield "w_1", "W", (10, 12, 14), 74, 1.75, -30.0, 1, 0
    yield "w_2", "W", (12, 12, 12), 75, 1.0, 70.0, 150, -300
    yield "w_3", "W", (12, 13, 14), 76, 1.0, 0.0, 1, 0
    yield "wxyz_0", "Wxyz", (12, 14, 16), 77, 1.0, 50.0, 1, 0
    yield "wxyz_1", "Wxyz", (12, 12, 14), 78, 1.0, -50.0, 40, -100


def write_oiftw():
    out_dir = os.path.join(HERE, "oiftw")
    os.makedirs(out_dir, exist_ok=True)
    for name, form, shape, seed, adj, per, scale, off in oiftw_cases():
        c = P.small_case(shape, seed, per=per, top=255, adj_r=adj)
        img = np.ascontiguousarray(c.image, np.int32) + off
        lab0 = c.label_init()
        rec = dict(form=np.array(form), image=img, seeds=c.seeds.astype(np.int32),
                   seed_labels=c.seed_labels.astype(np.int32), adj_r=np.float32(adj), per=np.float32(per))
        W = Wxyz = None
        if form == "W":
            W = grad_image(img - off) * scale
            rec["W"] = W
        if form == "Wxyz":
            Wxyz = tuple(grad_image(np.roll(img - off, k, k)) * scale for k in (2, 1, 0))
            rec.update(Wx=Wxyz[0], Wy=Wxyz[1], Wz=Wxyz[2])
        rec["oift_label"] = O.ref_oift_w(img, c.seeds, lab0, adj, per, W=W, Wxyz=Wxyz).astype(np.int8)
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), **rec)
        print("oiftw", name, form, img.shape, "obj voxels", int((rec["oift_label"] > 0).sum()))


def spacing_cases():
    """Anisotropic voxel spacing: (name, spacing, kind, adj).  It reaches the relaxations'
    Dpq and the blur taps only."""
    for i, sp in enumerate([(0.8, 1.2, 2.5), (1.0, 0.5, 0.5), (2.0, 2.0, 2.0)]):
        yield f"relax6_{i}", sp, "relax", 1.0
    yield "relax26_0", (0.8, 1.2, 2.5), "relax", 1.75
    yield "perclass_0", (0.7, 0.9, 1.6), "perclass", 1.0
    yield "rbf_0", (0.8, 1.2, 2.5), "rbf", 1.0
    yield "rbf26_0", (1.0, 0.5, 0.5), "rbf", 1.75
    yield "grad_0", (0.8, 1.2, 2.5), "grad", 1.0
    yield "binary_0", (0.8, 1.2, 2.5), "binary", 1.0
    yield "binary26_0", (1.0, 0.5, 0.5), "binary", 1.75


def write_spacing():
    out_dir = os.path.join(HERE, "spacing")
    os.makedirs(out_dir, exist_ok=True)
    for k, (name, sp, kind, adj) in enumerate(spacing_cases()):
        c = P.small_case((12, 14, 16), 90 + k, per=50.0, top=255, adj_r=adj)
        img = np.ascontiguousarray(c.image, np.int32)
        lab0 = c.label_init()
        seeds, slab = c.seeds, c.seed_labels
        pc = None
        if kind == "perclass":
            zs, ys, xs = img.shape
            extra = np.array([3 + 3 * xs + 3 * xs * ys], np.int64)
            keep = ~np.isin(seeds, extra)
            seeds = np.concatenate([seeds[keep][:10], extra, seeds[keep][10:]])
            slab = np.concatenate([slab[keep][:10], [2], slab[keep][10:]]).astype(np.int32)
            lab0 = np.full(img.size, -1, np.int32)
            lab0[seeds] = slab
            lab0 = lab0.reshape(img.shape)
            pc = np.array([-50.0, 50.0, 50.0], np.float32)
        oift = O.ref_oift(img, seeds, lab0, adj, 50.0, pc)
        rec = dict(kind=np.array(kind), spacing=np.array(sp, np.float32), image=img.astype(np.int16),
                   seeds=seeds.astype(np.int32), seed_labels=slab.astype(np.int32), adj_r=np.float32(adj),
                   per=np.float32(50.0), oift_label=oift.astype(np.int8), niter=np.int32(8),
                   per_class=np.zeros(0, np.float32) if pc is None else pc)
        if kind in ("relax", "perclass"):
            rec["relax_label"] = O.ref_relax(img, seeds, oift, 8, adj, 50.0, pc, spacing=sp)
        elif kind == "rbf":
            sigma = O.ref_estimate_sigma(img)
            rec["sigma"] = np.float32(sigma)
            rec["relax_label"] = O.ref_relax_rbf(img, seeds, oift, 8, sigma, adj, 50.0, spacing=sp)
        elif kind == "grad":
            G = grad_image(img)
            rec["G"] = G
            rec["alpha"] = np.float32(0.5)
            go = O.ref_oift_gradweight(img, G, seeds, lab0, adj, 0.5, 50.0)
            rec["oift_label"] = go.astype(np.int8)
            rec["relax_label"] = O.ref_relax_gradweight(img, G, seeds, go, 8, adj, 0.5, 50.0, spacing=sp)
        else:
            lb = (oift > 0).astype(np.int32)
            rec["oift_label"] = lb.astype(np.int8)
            rec["relax_label"] = O.ref_relax_binary(img, seeds, lb, 6, adj, 50.0, spacing=sp)
            rec["niter"] = np.int32(6)
        rec["relax_label"] = rec["relax_label"].astype(np.int8)
        blur_in = img * 16 - 1000
        rec["blur_in"] = blur_in
        rec["blur1"] = O.ref_gaussian_blur(blur_in, 1, spacing=sp)
        rec["blur2"] = O.ref_gaussian_blur(blur_in, 2, spacing=sp)
        rec["taps"] = O.ref_gauss_kernel(sp)
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), **rec)
        print("spacing", name, sp, kind, "obj voxels", int((rec["relax_label"] > 0).sum()))


def blur_cases():
    rng = np.random.default_rng(77)
    yield "ell", P.small_case((12, 14, 16), 11, per=50.0, top=255).image
    yield "rand_signed", rng.integers(-5000, 70000, size=(9, 11, 13)).astype(np.int32)
    yield "plane2d", rng.integers(0, 4096, size=(1, 40, 50)).astype(np.int32)
    yield "thin", rng.integers(0, 30000, size=(3, 1, 70)).astype(np.int32)
    yield "big_values", rng.integers(-(1 << 30), 1 << 30, size=(5, 6, 7)).astype(np.int32)


def write_blur():
    out_dir = os.path.join(HERE, "blur")
    os.makedirs(out_dir, exist_ok=True)
    for name, img in blur_cases():
        img = np.ascontiguousarray(img, np.int32)
        outs = {f"out{k}": O.ref_gaussian_blur(img, k) for k in (1, 2, 3)}
        np.savez_compressed(os.path.join(out_dir, f"{name}.npz"), image=img, taps=O.ref_gauss_kernel(), **outs)
        print("blur", name, img.shape)


def write_stages():
    out_dir = os.path.join(HERE, "stages")
    os.makedirs(out_dir, exist_ok=True)
    rng = np.random.default_rng(78)
    # Sobel: random, a blurred phantom, signed values (negative sums shift arithmetically)
    sob = {"rand": rng.integers(0, 4000, size=(9, 10, 11)).astype(np.int32),
           "phantom": O.ref_gaussian_blur(P.small_case((12, 14, 16), 11, per=50.0).image, 2),
           "signed": rng.integers(-20000, 20000, size=(6, 7, 8)).astype(np.int32),
           "plane": rng.integers(0, 255, size=(1, 12, 13)).astype(np.int32)}
    for name, img in sob.items():
        np.savez_compressed(os.path.join(out_dir, f"sobel_{name}.npz"), image=img, grad=O.ref_sobel(img))
        print("sobel", name, img.shape)
    # seeds: object seeds (label 1), background seeds (label 0) and an out-of-range entry
    for name, shape in (("vol", (14, 17, 19)), ("flat", (1, 20, 30)), ("thin", (3, 2, 40))):
        n = int(np.prod(shape))
        lab = np.full(n, -1, np.int32)
        obj = rng.choice(n, size=3, replace=False)
        bg = rng.choice(np.setdiff1d(np.arange(n), obj), size=4, replace=False)
        lab[obj] = 1
        lab[bg] = 0
        lab = lab.reshape(shape)
        seeds = np.concatenate([obj, bg, [n + 3]]).astype(np.int32)
        rec = dict(label=lab, seeds=seeds, edt=O.ref_bfs_edt(lab, seeds))
        for stride in (1, 2, 3, 5, 8):
            l1, s1, _ = O.ref_boundary_seeds(lab, seeds, stride)
            l2, s2, _ = O.ref_boundary_seeds_geodesic(lab, seeds, stride)
            rec[f"uni{stride}_label"], rec[f"uni{stride}_seeds"] = l1, s1
            rec[f"geo{stride}_label"], rec[f"geo{stride}_seeds"] = l2, s2
        np.savez_compressed(os.path.join(out_dir, f"seeds_{name}.npz"), **rec)
        print("seeds", name, shape)
    # conditional dilation of an OIFT result; the quantised image has many ties
    c = P.small_case((16, 18, 20), 12, per=50.0)
    img = O.ref_gaussian_blur(c.image, 2)
    lab = (O.ref_oift(img, c.seeds, c.label_init(), 1.0, c.per) > 0).astype(np.int32)
    for name, im in (("smooth", img), ("ties", (img // 64) * 64)):
        rec = dict(image=im, label=lab)
        for pc in (0.0, 25.0, 50.0, 90.0, 100.0):
            for r in (1.0, 1.5):
                rec[f"p{int(pc)}_r{r}"] = O.ref_dilation_conditional(im, lab, r, pc)
        np.savez_compressed(os.path.join(out_dir, f"dilation_{name}.npz"), **rec)
        print("dilation", name)


def cli_cases():
    """(name, volume file, volume, seed lines, polarity lines or None, argv tail, output file)."""
    c = P.small_case((24, 28, 32), 5, per=50.0)
    img = c.image.astype(np.int32)
    zs, ys, xs = img.shape
    xyz = lambda p: (int(p) % xs, (int(p) // xs) % ys, int(p) // (xs * ys))
    base = [(*xyz(p), 0, int(lb)) for p, lb in zip(c.seeds[:27], c.seed_labels[:27])]   # object cube
    bg = [(*xyz(p), 0, 0) for p in c.seeds[27::97][:12]]                                # a few faces
    yield "binary_scn", "vol.scn", img, base + bg, None, ["0.5", "10", "50"], "out.scn", []
    yield "negative_nii", "vol.nii.gz", img - 500, base + bg, None, ["-0.3", "5", "30"], "out.nii.gz", ["4"]
    three = base[:13] + [(3, 3, 3, 0, 2), (4, 3, 3, 0, 2)] + base[13:] + bg
    yield "perclass", "vol.nii.gz", img, three, ["3", "-0.5 0.5 0.5"], ["0.5", "5", "50"], "out.nii", ["0"]
    odd = base[:10] + [(1, 1, 1, 0, -3), (xs + 5, 2, 2, 0, 1)] + base[10:] + bg
    yield "clamped", "vol.scn", img, odd, None, ["0.0", "0", "80"], "out.nii.gz", ["6"]


def cli_spacing():
    """Voxel spacing the CLI cases' volume files carry (gft's default elsewhere)."""
    return {"aniso_nii": (0.7, 0.9, 2.0), "aniso_scn": (1.5, 1.0, 0.6)}


def cli_cases_spacing():
    c = P.small_case((24, 28, 32), 5, per=50.0)
    img = c.image.astype(np.int32)
    zs, ys, xs = img.shape
    xyz = lambda p: (int(p) % xs, (int(p) // xs) % ys, int(p) // (xs * ys))
    base = [(*xyz(p), 0, int(lb)) for p, lb in zip(c.seeds[:27], c.seed_labels[:27])]
    bg = [(*xyz(p), 0, 0) for p in c.seeds[27::97][:12]]
    yield "aniso_nii", "vol.nii.gz", img, base + bg, None, ["0.5", "10", "50"], "out.nii.gz", ["4"]
    yield "aniso_scn", "vol.scn", img, base + bg, None, ["-0.4", "8", "40"], "out.scn", []


def write_cli():
    import shutil
    root = os.path.join(HERE, "cli")
    import itertools
    for name, vname, vol, seeds, pol, args, oname, tail in itertools.chain(cli_cases(), cli_cases_spacing()):
        d = os.path.join(root, name)
        shutil.rmtree(d, ignore_errors=True)
        os.makedirs(d)
        with O._spacing(cli_spacing().get(name), ref=True):
            O.ref_scene_write(os.path.join(d, vname), vol)
        with open(os.path.join(d, "seeds.txt"), "w") as f:
            f.write(f"{len(seeds)}\n" + "".join(f"{x} {y} {z} {i} {lb}\n" for x, y, z, i, lb in seeds))
        extra = []
        if pol is not None:
            with open(os.path.join(d, "pol.txt"), "w") as f:
                f.write("\n".join(pol) + "\n")
            extra = ["pol.txt"]
        argv = [vname, "seeds.txt"] + args + [oname] + tail + extra
        cwd = os.getcwd()
        os.chdir(d)
        try:
            assert O.ref_cli(argv) == 0
            labels = O.ref_scene_read(oname)
            os.remove(oname)
        finally:
            os.chdir(cwd)
        with open(os.path.join(d, "argv.txt"), "w") as f:
            f.write(" ".join(argv) + "\n")
        np.savez_compressed(os.path.join(d, "expected.npz"), label=labels)
        print("cli", name, argv, np.bincount(labels.ravel()))


def main():
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libroift_ref.so is not built (make -C oracle ref)")
    if sys.argv[1:] == ["blur"]:
        write_blur()
        return
    if sys.argv[1:] == ["stages"]:
        write_stages()
        return
    if sys.argv[1:] == ["cli"]:
        write_cli()
        return
    if sys.argv[1:] == ["oiftw"]:
        write_oiftw()
        return
    if sys.argv[1:] == ["spacing"]:
        write_spacing()
        return
    for name, c in cases():
        img = np.ascontiguousarray(c["image"], np.int32)
        lab0 = np.full(img.size, -1, np.int32)
        lab0[c["seeds"]] = c["seed_labels"]
        lab0 = lab0.reshape(img.shape)
        oift = O.ref_oift(img, c["seeds"], lab0, c["adj_r"], c["per"], c["per_class"])
        relax = O.ref_relax(img, c["seeds"], oift, c["niter"], c["adj_r"], c["per"], c["per_class"])
        np.savez_compressed(
            os.path.join(HERE, f"{name}.npz"), image=img.astype(np.int16), seeds=c["seeds"].astype(np.int32),
            seed_labels=c["seed_labels"].astype(np.int32), adj_r=np.float32(c["adj_r"]), per=np.float32(c["per"]),
            per_class=(np.zeros(0, np.float32) if c["per_class"] is None else c["per_class"]),
            niter=np.int32(c["niter"]), oift_label=oift.astype(np.int8 if oift.max() < 127 else np.int32),
            relax_label=relax.astype(np.int8))
        print(name, img.shape, "obj voxels", int((oift > 0).sum()), int((relax > 0).sum()))
    write_variants()
    write_blur()
    write_stages()
    write_cli()
    write_oiftw()
    write_spacing()


if __name__ == "__main__":
    main()
