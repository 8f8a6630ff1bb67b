"""Deterministic synthetic volumes and seed sets for the OIFT + relaxation workloads.

No image files exist for these workloads, so every volume is generated from a
seeded numpy Generator.  Volumes are int arrays of shape (zs, ys, xs); the
linear voxel index is x + y*xs + z*xs*ys (gft_scene32.h:173-179).  Seeds are
returned in gft S order: user/internal seeds first, then boundary seeds in the
raster order inject_boundary_seeds (oiftrelax.cpp:175-224) appends them.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np
from scipy import ndimage

NIL = -1


@dataclass
class Case:
    image: np.ndarray            # (zs, ys, xs) int16/int32, values >= 0
    seeds: np.ndarray            # int64 linear indices in S order
    seed_labels: np.ndarray      # int32 labels, same order
    adj_r: float                 # 1.0 -> 6-adjacency, 1.75 -> 26-adjacency
    per: float                   # polarity x100 (alpha=0.5 -> 50)
    niter: int
    truth: np.ndarray | None = None   # boolean object mask used to build the volume
    meta: dict = field(default_factory=dict)

    def label_init(self) -> np.ndarray:
        lab = np.full(self.image.size, NIL, np.int32)
        lab[self.seeds] = self.seed_labels
        return lab.reshape(self.image.shape)


def _face_seeds(shape) -> np.ndarray:
    """Voxels on the volume's faces; a dimension of size 1 has no faces (a 2-D image's
    border is its four edges -- seeding z == 0 == zsize-1 would seed every pixel)."""
    zs, ys, xs = shape
    z, y, x = np.meshgrid(np.arange(zs), np.arange(ys), np.arange(xs), indexing="ij")
    face = np.zeros(shape, bool)
    for coord, size in ((x, xs), (y, ys), (z, zs)):
        if size > 1:
            face |= (coord == 0) | (coord == size - 1)
    return np.flatnonzero(face.ravel())


def _rescale(I: np.ndarray, top: int) -> np.ndarray:
    lo, hi = float(I.min()), float(I.max())
    if hi <= lo:
        return np.zeros(I.shape, np.int16)
    return np.rint(top * (I - lo) / (hi - lo)).astype(np.int16)


def ellipsoid_volume(shape, center, axes, inside=1200.0, outside=800.0, noise=80.0, bias=0.1,
                     blur=1.5, rng=None, top=4095):
    """Ellipsoid phantom: N(inside, noise) in N(outside, noise), radial bias
    1 + bias*cos(pi*r/rmax), Gaussian blur, rounded, rescaled to [0, top] int16."""
    rng = np.random.default_rng(rng)
    zs, ys, xs = shape
    cx, cy, cz = center
    ax, ay, az = axes
    z, y, x = np.ogrid[0:zs, 0:ys, 0:xs]
    r2 = ((x - cx) / ax) ** 2 + ((y - cy) / ay) ** 2 + ((z - cz) / az) ** 2
    truth = r2 <= 1.0
    vol = np.where(truth, inside, outside) + rng.standard_normal(shape, dtype=np.float32) * noise
    rr = np.sqrt((x - (xs - 1) / 2.0) ** 2 + (y - (ys - 1) / 2.0) ** 2 + (z - (zs - 1) / 2.0) ** 2)
    vol = vol * (1.0 + bias * np.cos(np.pi * rr / rr.max()))
    if blur > 0:
        vol = ndimage.gaussian_filter(vol.astype(np.float32), blur, truncate=5.0 / blur if blur else 4.0)
    vol = np.clip(np.rint(vol), -32768, 32767)
    return _rescale(vol, top), truth


def cube_seeds(shape, center_voxel, half: int) -> np.ndarray:
    zs, ys, xs = shape
    cx, cy, cz = center_voxel
    idx = [x + y * xs + z * xs * ys
           for z in range(max(0, cz - half), min(zs, cz + half + 1))
           for y in range(max(0, cy - half), min(ys, cy + half + 1))
           for x in range(max(0, cx - half), min(xs, cx + half + 1))]
    return np.array(idx, np.int64)


def headline_case(n: int = 256, seed: int = 2, niter: int = 50) -> Case:
    """256^3 int16, 6-adjacency, alpha=0.5, ellipsoid semi-axes (90,70,60) scaled to n,
    internal 5^3 seed cube at the centre, every face voxel a background seed, 50 iterations."""
    s = n / 256.0
    c = (n - 1) / 2.0
    img, truth = ellipsoid_volume((n, n, n), (c, c, c), (90 * s, 70 * s, 60 * s), rng=seed)
    cv = n // 2
    internal = cube_seeds((n, n, n), (cv, cv, cv), 2)
    faces = _face_seeds((n, n, n))
    faces = faces[~np.isin(faces, internal)]
    seeds = np.concatenate([internal, faces])
    labels = np.concatenate([np.ones(internal.size, np.int32), np.zeros(faces.size, np.int32)])
    return Case(img, seeds, labels, 1.0, 50.0, niter, truth,
                dict(workload=f"oift_fmax+relax {n}^3 int16 6-adj alpha=0.5 relax={niter}"))


def plumbing_case_2d(seed: int = 1) -> Case:
    """512x512 disk (r=120) N(170,12) on N(100,12), blur sigma 2, clipped to [0,255],
    rescaled to [0,1000]; internal 7x7 centre square, all border pixels background."""
    rng = np.random.default_rng(seed)
    n = 512
    y, x = np.ogrid[0:n, 0:n]
    truth = (x - 255.5) ** 2 + (y - 255.5) ** 2 <= 120 ** 2
    img = np.where(truth, 170.0, 100.0) + rng.standard_normal((n, n)) * 12.0
    img = ndimage.gaussian_filter(img, 2.0)
    img = np.clip(np.rint(img), 0, 255)
    img = np.rint(img * 1000.0 / 255.0).astype(np.int32)[None]
    internal = np.array([x_ + y_ * n for y_ in range(253, 260) for x_ in range(253, 260)], np.int64)
    yy, xx = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    border = np.flatnonzero(((xx == 0) | (xx == n - 1) | (yy == 0) | (yy == n - 1)).ravel())
    border = border[~np.isin(border, internal)]
    seeds = np.concatenate([internal, border])
    labels = np.concatenate([np.ones(internal.size, np.int32), np.zeros(border.size, np.int32)])
    return Case(img, seeds, labels, 1.0, 50.0, 10, truth[None], dict(workload="2d 512^2 4-adj"))


def small_case(shape=(24, 28, 32), seed=0, per=50.0, adj_r=1.0, niter=10, top=4095, noise=80.0,
               n_internal_half=1) -> Case:
    """Scaled-down ellipsoid case used by the parity tests."""
    zs, ys, xs = shape
    rng = np.random.default_rng(seed)
    center = ((xs - 1) / 2.0 + rng.uniform(-1, 1), (ys - 1) / 2.0 + rng.uniform(-1, 1),
              (zs - 1) / 2.0 + rng.uniform(-1, 1))
    axes = (xs * 0.33, ys * 0.3, zs * 0.28)
    img, truth = ellipsoid_volume(shape, center, axes, noise=noise, rng=rng, top=top)
    cv = (xs // 2, ys // 2, zs // 2)
    internal = cube_seeds(shape, cv, n_internal_half)
    faces = _face_seeds(shape)
    faces = faces[~np.isin(faces, internal)]
    seeds = np.concatenate([internal, faces])
    labels = np.concatenate([np.ones(internal.size, np.int32), np.zeros(faces.size, np.int32)])
    return Case(img, seeds, labels, adj_r, per, niter, truth, dict(workload=f"small {shape}"))


def multi_object_case(shape=(256, 384, 384), seed=3, n_internal=20, niter=100) -> Case:
    """Three nested/adjacent ellipsoids (labels: target=1, the others and the border=0),
    26-adjacency; 20 internal seeds sampled inside the target, external seeds sampled
    on the other objects plus every border voxel."""
    rng = np.random.default_rng(seed)
    zs, ys, xs = shape
    z, y, x = np.ogrid[0:zs, 0:ys, 0:xs]

    def ell(c, a):
        return ((x - c[0]) / a[0]) ** 2 + ((y - c[1]) / a[1]) ** 2 + ((z - c[2]) / a[2]) ** 2 <= 1

    A = ell((xs * 0.38, ys * 0.5, zs * 0.5), (xs * 0.22, ys * 0.25, zs * 0.3))
    B = ell((xs * 0.38, ys * 0.5, zs * 0.5), (xs * 0.1, ys * 0.11, zs * 0.13))   # nested in A
    Cc = ell((xs * 0.72, ys * 0.5, zs * 0.5), (xs * 0.15, ys * 0.2, zs * 0.25))  # adjacent
    vol = np.full(shape, 800.0, np.float32)
    vol[A] = 1150.0
    vol[B] = 1450.0
    vol[Cc] = 1000.0
    vol += rng.standard_normal(shape, dtype=np.float32) * 80.0
    vol = ndimage.gaussian_filter(vol, 1.5, truncate=5.0 / 1.5)
    img = _rescale(np.rint(vol), 4095)
    target = A & ~B
    tidx = np.flatnonzero(target.ravel())
    internal = np.sort(rng.choice(tidx, n_internal, replace=False))
    other = np.flatnonzero((B | Cc).ravel())
    ext = np.sort(rng.choice(other, min(200, other.size), replace=False))
    faces = _face_seeds(shape)
    ext = np.unique(np.concatenate([ext, faces]))
    ext = ext[~np.isin(ext, internal)]
    seeds = np.concatenate([internal, ext])
    labels = np.concatenate([np.ones(internal.size, np.int32), np.zeros(ext.size, np.int32)])
    return Case(img, seeds, labels, 1.75, 50.0, niter, target,
                dict(workload=f"3 objects {shape} 26-adj"))


def batch_cases(count=64, seed=4, lo=64, hi=192, niter=30):
    """Uneven-size ellipsoid volumes, each dimension log-uniform in [lo, hi]."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        dims = np.exp(rng.uniform(np.log(lo), np.log(hi), 3)).astype(int)
        xs, ys, zs = (int(d) for d in dims)
        shape = (zs, ys, xs)
        c = (xs / 2 + rng.uniform(-4, 4), ys / 2 + rng.uniform(-4, 4), zs / 2 + rng.uniform(-4, 4))
        a = (xs * rng.uniform(0.2, 0.35), ys * rng.uniform(0.2, 0.35), zs * rng.uniform(0.2, 0.35))
        contrast = rng.uniform(200, 600)
        img, truth = ellipsoid_volume(shape, c, a, inside=800 + contrast, outside=800,
                                      noise=rng.uniform(40, 100), rng=rng)
        cv = (int(round(c[0])), int(round(c[1])), int(round(c[2])))
        internal = cube_seeds(shape, cv, 1)
        faces = _face_seeds(shape)
        faces = faces[~np.isin(faces, internal)]
        seeds = np.concatenate([internal, faces])
        labels = np.concatenate([np.ones(internal.size, np.int32), np.zeros(faces.size, np.int32)])
        out.append(Case(img, seeds, labels, 1.0, 50.0, niter, truth, dict(workload=f"batch {shape}")))
    return out
