"""A mirror of the gft API surface the reference's callers use, backed by libroift.so.

Same names, argument meaning and in-place behaviour as the C++ originals:
  gft::Scene32::Create / Destroy               (gft_scene32.h:32-40)
  gft::Scene32::GaussianBlur                   (gft_filtering3.h:54, gft_filtering3.cpp:385-397)
  gft::AdjRel3::Spheric(r)                     (gft_adjrel3.cpp:47-85)
  gft::ift::OIFT_Multi / OIFT_Multi_PerClass   (gft_ift.h:108-122)
  gft::ift::ORelax_1_Multi / _PerClass         (gft_ift.h:451-465)
  gft::ift::ORelax_1 (image / W / Wx,Wy,Wz)     (gft_ift.h:443-482)
  and, in `exp`, the experiments' OIFT_Multi_GradWeight, ORelax_1_Multi_GradWeight,
  ORelax_1_Multi_GaussRBF and estimate_sigma
The gft seed array convention is kept: S[0] = count, S[1..S[0]] voxel indices.
Where the reference calls exit(-1) this raises RoiftError.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import Context, NIL

_ctx = None


def _context() -> Context:
    global _ctx
    if _ctx is None:
        _ctx = Context(0)
    return _ctx


@dataclass
class sScene32:
    data: np.ndarray          # int32, length n, x-fastest
    xsize: int
    ysize: int
    zsize: int
    dx: float = 1.0
    dy: float = 1.0
    dz: float = 1.0

    @property
    def n(self) -> int:
        return self.xsize * self.ysize * self.zsize

    def volume(self) -> np.ndarray:
        return self.data.reshape(self.zsize, self.ysize, self.xsize)


@dataclass
class sAdjRel3:
    d: np.ndarray   # (n, 3) displacements (x, y, z), centre first
    r: float

    @property
    def n(self) -> int:
        return int(self.d.shape[0])


class Scene32:
    @staticmethod
    def Create(xsize, ysize=None, zsize=None) -> sScene32:
        if isinstance(xsize, sScene32):
            s = xsize
            return sScene32(np.zeros(s.n, np.int32), s.xsize, s.ysize, s.zsize, s.dx, s.dy, s.dz)
        return sScene32(np.zeros(xsize * ysize * zsize, np.int32), xsize, ysize, zsize)

    @staticmethod
    def FromArray(vol: np.ndarray) -> sScene32:
        zs, ys, xs = vol.shape
        return sScene32(np.ascontiguousarray(vol, np.int32).ravel().copy(), xs, ys, zs)

    @staticmethod
    def Fill(scn: sScene32, value: int) -> None:
        scn.data[:] = value

    @staticmethod
    def GetMaximumValue(scn: sScene32) -> int:
        return int(scn.data.max())

    @staticmethod
    def GaussianBlur(scn: sScene32) -> sScene32:
        """A new scene: the 7-tap spherical Gaussian of radius 1, computed on the GPU."""
        out = _context().gaussian_blur_host(scn.volume(), passes=1)
        return sScene32(out.ravel(), scn.xsize, scn.ysize, scn.zsize, scn.dx, scn.dy, scn.dz)

    @staticmethod
    def Destroy(scn) -> None:
        return None


class AdjRel3:
    @staticmethod
    def Spheric(r: float) -> sAdjRel3:
        """Raster dz, dy, dx order over the ball with the centre moved to index 0."""
        r0, r2 = int(r), int(r * r)
        pts = [(x, y, z) for z in range(-r0, r0 + 1) for y in range(-r0, r0 + 1) for x in range(-r0, r0 + 1)
               if x * x + y * y + z * z <= r2]
        i0 = pts.index((0, 0, 0))
        pts = [pts[i0]] + pts[:i0] + pts[i0 + 1:]
        return sAdjRel3(np.array(pts, np.int32), float(r))

    @staticmethod
    def Destroy(A) -> None:
        return None


def _seeds_of(S, label: sScene32) -> np.ndarray:
    if S is None:  # every labelled voxel, raster order (gft_ift.cpp:944-948)
        return np.flatnonzero(label.data != NIL).astype(np.int32)
    S = np.asarray(S, np.int32)
    return S[1:1 + int(S[0])]


def _shape(scn: sScene32):
    return (scn.zsize, scn.ysize, scn.xsize)


class ift:
    @staticmethod
    def OIFT_Multi(A: sAdjRel3, scn: sScene32, per: float, S, label: sScene32) -> None:
        lab, _ = _context().oift_fmax_host(scn.volume(), _seeds_of(S, label), label.volume(), A.r, per, None, False)
        label.data[:] = lab.ravel()

    @staticmethod
    def OIFT_Multi_PerClass(A, scn, per_class, max_label, S, label) -> None:
        pc = np.asarray(per_class, np.float32)[: max_label + 1]
        lab, _ = _context().oift_fmax_host(scn.volume(), _seeds_of(S, label), label.volume(), A.r, 0.0, pc, False)
        label.data[:] = lab.ravel()

    @staticmethod
    def ORelax_1_Multi(A, scn, per, S, label, ntimes) -> None:
        if S is None or int(np.asarray(S)[0]) <= 0:   # binary fallback (gft_ift.cpp:8678-8682)
            return ift.ORelax_1(A, scn, per, S, label, ntimes)
        lab, _ = _context().relax_host(scn.volume(), _seeds_of(S, label), label.volume(), ntimes, A.r, per)
        label.data[:] = lab.ravel()

    @staticmethod
    def ORelax_1(*args) -> None:
        """ORelax_1(A, scn, per, S, label, ntimes), ORelax_1(W, A, scn, per, S, label, ntimes) or
        ORelax_1(Wx, Wy, Wz, scn, per, S, label, ntimes) (gft_ift.h:443-482)."""
        W = Wxyz = None
        if len(args) == 6:
            A, scn, per, S, label, ntimes = args
        elif len(args) == 7:
            W, A, scn, per, S, label, ntimes = args
            W = W.volume()
        elif len(args) == 8:
            wx, wy, wz, scn, per, S, label, ntimes = args
            Wxyz = (wx.volume(), wy.volume(), wz.volume())
            A = AdjRel3.Spheric(1.0)
        else:
            raise TypeError("ORelax_1 takes 6, 7 or 8 arguments")
        seeds = None if S is None else np.asarray(S, np.int32)[1:1 + int(np.asarray(S)[0])]
        lab, _ = _context().relax_binary_host(scn.volume(), seeds, label.volume(), ntimes, A.r, per, W, Wxyz)
        label.data[:] = lab.ravel()

    @staticmethod
    def ORelax_1_Multi_PerClass(A, scn, per_class, max_label, S, label, ntimes) -> None:
        if S is None or int(np.asarray(S)[0]) <= 0:
            return  # the reference returns without change (gft_ift.cpp:8942-8943)
        pc = np.asarray(per_class, np.float32)[: max_label + 1]
        lab, _ = _context().relax_host(scn.volume(), _seeds_of(S, label), label.volume(), ntimes, A.r, 0.0, pc)
        label.data[:] = lab.ravel()


class exp:
    """The experiment variants of the hot path (experiments/exp_gradient_weight.cpp,
    experiments/exp_gaussian_rbf_relax.cpp), with the experiments' argument order."""

    @staticmethod
    def OIFT_Multi_GradWeight(A, scn, grad_norm, per, alpha, S, label) -> None:
        import torch
        ctx = _context()
        dev = torch.device("cuda", ctx.device)
        img = torch.from_numpy(scn.volume()).to(dev)
        g = torch.from_numpy(grad_norm.volume()).to(dev)
        lab = torch.from_numpy(label.volume().copy()).to(dev)
        seeds = torch.from_numpy(_seeds_of(S, label)).to(dev)
        ctx.oift_gradweight(img, g, seeds, lab, alpha, A.r, per, want_cost=False)
        label.data[:] = lab.cpu().numpy().ravel()

    @staticmethod
    def ORelax_1_Multi_GradWeight(A, scn, grad_norm, per, alpha, S, label, ntimes) -> None:
        import torch
        if S is None or int(np.asarray(S)[0]) <= 0:
            return   # exp_gradient_weight.cpp:296
        ctx = _context()
        dev = torch.device("cuda", ctx.device)
        img = torch.from_numpy(scn.volume()).to(dev)
        g = torch.from_numpy(grad_norm.volume()).to(dev)
        lab = torch.from_numpy(label.volume().copy()).to(dev)
        seeds = torch.from_numpy(_seeds_of(S, label)).to(dev)
        ctx.relax_gradweight(img, g, seeds, lab, ntimes, alpha, A.r, per, want_memb=False)
        label.data[:] = lab.cpu().numpy().ravel()

    @staticmethod
    def ORelax_1_Multi_GaussRBF(A, scn, per, sigma, S, label, ntimes) -> None:
        import torch
        if S is None or int(np.asarray(S)[0]) <= 0:
            return   # exp_gaussian_rbf_relax.cpp:235
        ctx = _context()
        dev = torch.device("cuda", ctx.device)
        img = torch.from_numpy(scn.volume()).to(dev)
        lab = torch.from_numpy(label.volume().copy()).to(dev)
        seeds = torch.from_numpy(_seeds_of(S, label)).to(dev)
        ctx.relax_rbf(img, seeds, lab, ntimes, sigma, A.r, want_memb=False)
        label.data[:] = lab.cpu().numpy().ravel()

    @staticmethod
    def estimate_sigma(scn) -> float:
        """Global standard deviation (exp_gaussian_rbf_relax.cpp:194-208), computed on the device."""
        import torch
        ctx = _context()
        img = torch.from_numpy(scn.volume()).to(torch.device("cuda", ctx.device))
        lab = torch.zeros(img.shape, dtype=torch.int32, device=img.device)
        seeds = torch.zeros(0, dtype=torch.int32, device=img.device)
        _, used = ctx.relax_rbf(img, seeds, lab, 0, 0.0, want_memb=False)
        return used
