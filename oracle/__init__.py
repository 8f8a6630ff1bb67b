"""ctypes bindings for the CPU oracle (oracle.c) and, when built, the compiled
reference (oracle/_ref/libroift_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product package (roift_amd) never imports this module.

Volumes are numpy arrays of shape (zs, ys, xs) in C order, so the linear index
is x + y*xs + z*xs*ys, gft's addressing (gft_scene32.h:173-179).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

NIL = -1
INT32_MAX = 2**31 - 1


def build(reference: str = "/root/reference") -> None:
    """Compile liboracle.so and, if the reference checkout exists, _ref/."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    if os.path.isdir(os.path.join(reference, "gft", "src")):
        subprocess.run(["make", "-s", "-j8", "-C", _HERE, "ref", f"REF={reference}"], check=True)


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        i32p, f32p, i8p = C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_int8)
        u32p, i64p, f64p = C.POINTER(C.c_uint32), C.POINTER(C.c_int64), C.POINTER(C.c_double)
        L.orc_spheric.argtypes = [C.c_float, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int]
        L.orc_oift_image.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, f32p,
                                     C.c_int, i32p, i32p, i32p, i8p, i32p, i32p, i64p]
        L.orc_edge_weights.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, u32p, u32p]
        L.orc_oift_edges.argtypes = [u32p, u32p, C.c_int, C.c_int, C.c_int, C.c_float, i8p, C.c_int,
                                     i32p, i32p, i32p, i8p, u32p, i32p, i64p]
        L.orc_relax_f32.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, f32p,
                                    C.c_int, i32p, i32p, C.c_int, C.c_int, f32p, i32p, i32p]
        L.orc_relax_f64.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, f32p,
                                    C.c_int, i32p, i32p, C.c_int, C.c_int, f64p, i32p, i32p]
        L.orc_relax_binary_f32.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, i32p, i32p,
                                           C.c_int, f32p]
        L.orc_relax_binary_w_f32.argtypes = [i32p, i32p, i32p, i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                             C.c_float, i32p, i32p, C.c_int, C.c_int, f32p]
        for name, mp in (("orc_relax_model_f32", f32p), ("orc_relax_model_f64", f64p)):
            getattr(L, name).argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float,
                                         C.c_float, C.c_int, C.c_float, f32p, C.c_int, i32p, i32p, C.c_int,
                                         C.c_int, mp, i32p, i32p]
        L.orc_gradweight_edges.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                           u32p, u32p]
        L.orc_estimate_sigma.argtypes = [i32p, C.c_int64]
        L.orc_gauss_kernel.argtypes = [f32p]
        L.orc_gaussian_blur.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p]
        L.orc_sobel.argtypes = [i32p, C.c_int, C.c_int, C.c_int, i32p]
        L.orc_grad_normalize.argtypes = [i32p, C.c_int64, C.c_int32]
        L.orc_bfs_edt.argtypes = [i32p, C.c_int, C.c_int, C.c_int, i32p, C.c_int, i32p]
        L.orc_boundary_seeds.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p]
        L.orc_boundary_seeds.restype = C.c_int64
        L.orc_boundary_seeds_geodesic.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p, i32p]
        L.orc_boundary_seeds_geodesic.restype = C.c_int64
        L.orc_dilation_conditional.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
        L.orc_dilation_conditional.restype = C.c_int64
        L.orc_relax_diff_f64.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, i32p, i32p, C.c_int, C.c_int,
                                         f64p, C.c_double, C.c_double, f64p, C.c_double, f64p, f64p, f64p, f64p,
                                         f64p]
        L.orc_estimate_sigma.restype = C.c_float
        L.orc_set_spacing.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_w_edges.argtypes = [i32p, i32p, i32p, i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                  u32p, u32p]
        _LIB = L
    return _LIB


def ref_available():
    """Whether oracle/_ref was built, without loading it (so that collecting the tests
    never maps the compiled reference into a process that does not call it)."""
    return os.path.exists(os.path.join(_HERE, "_ref", "libroift_ref.so"))


def ref_lib():
    """The compiled reference, or None when oracle/_ref was not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libroift_ref.so")
        if not os.path.exists(path):
            return None
        L = C.CDLL(path)
        i32p, f32p, f64p = C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_double)
        L.ref_oift.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, f32p, C.c_int,
                               i32p, i32p]
        L.ref_relax.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_int, C.c_float, f32p, C.c_int,
                                i32p, i32p, C.c_int]
        L.ref_pipeline.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, i32p, i32p, C.c_int,
                                   f64p, f64p]
        L.ref_relax_binary.argtypes = [i32p, i32p, i32p, i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float,
                                       C.c_float, i32p, i32p, C.c_int]
        L.ref_oift_gradweight.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                          i32p, i32p]
        L.ref_relax_gradweight.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float,
                                           i32p, i32p, C.c_int]
        L.ref_relax_rbf.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, i32p, i32p,
                                    C.c_int]
        L.ref_estimate_sigma.argtypes = [i32p, C.c_int, C.c_int, C.c_int]
        L.ref_gaussian_blur.argtypes = [i32p, C.c_int, C.c_int, C.c_int, C.c_int, i32p]
        L.ref_gauss_kernel.argtypes = [f32p, C.c_int]
        L.ref_sobel.argtypes = [i32p, C.c_int, C.c_int, C.c_int, i32p]
        L.ref_bfs_edt.argtypes = [i32p, C.c_int, C.c_int, C.c_int, i32p, i32p]
        L.ref_boundary_seeds.argtypes = [i32p, C.c_int, C.c_int, C.c_int, i32p, C.c_int, i32p, C.c_int]
        L.ref_boundary_seeds_geodesic.argtypes = [i32p, C.c_int, C.c_int, C.c_int, i32p, C.c_int, i32p, C.c_int]
        L.ref_dilation_conditional.argtypes = [i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float]
        L.ref_cli_main.argtypes = [C.c_int, C.POINTER(C.c_char_p)]
        L.ref_scene_dims.argtypes = [C.c_char_p, C.POINTER(C.c_int)]
        L.ref_scene_read.argtypes = [C.c_char_p, i32p, C.c_longlong]
        L.ref_scene_write.argtypes = [C.c_char_p, i32p, C.c_int, C.c_int, C.c_int]
        L.ref_estimate_sigma.restype = C.c_float
        L.ref_set_spacing.argtypes = [C.c_float, C.c_float, C.c_float]
        L.ref_oift_w.argtypes = [i32p, i32p, i32p, i32p, i32p, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                 i32p, i32p]
        _REF = L
    return _REF


class _spacing:
    """Sets the oracle's (and, with ref=True, the compiled reference's) voxel spacing for
    the duration of a call; None leaves unit spacing."""

    def __init__(self, spacing, ref=False):
        self.sp = None if spacing is None else tuple(float(v) for v in spacing)
        self.fn = (lambda: ref_lib().ref_set_spacing) if ref else (lambda: lib().orc_set_spacing)

    def __enter__(self):
        if self.sp is not None:
            self.fn()(*self.sp)

    def __exit__(self, *exc):
        if self.sp is not None:
            self.fn()(1.0, 1.0, 1.0)
        return False


def _dims(vol):
    zs, ys, xs = vol.shape
    return xs, ys, zs


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _gft_S(seeds):
    s = _i32(seeds).ravel()
    return np.concatenate([np.array([s.size], np.int32), s])


def spheric(r: float):
    dx = (C.c_int * 125)(); dy = (C.c_int * 125)(); dz = (C.c_int * 125)()
    n = lib().orc_spheric(C.c_float(r), dx, dy, dz, 125)
    assert n > 0
    return np.array([(dx[i], dy[i], dz[i]) for i in range(n)], dtype=np.int32)


def oift_image(I, seeds, label, adj_r=1.0, per=0.0, per_class=None, want_order=False):
    """OIFT_Multi (per) or OIFT_Multi_PerClass (per_class) restated.

    Returns dict(label, cost, pred, value[, order]).  seeds: 1-D linear indices
    in S order (or None for the S==NULL form).
    """
    I = _i32(I)
    xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    n = I.size
    cost = np.empty(n, np.int32); pred = np.empty(n, np.int8); val = np.empty(n, np.int32)
    order = np.empty(n, np.int32) if want_order else None
    npop = np.zeros(1, np.int64)
    S = None if seeds is None else _gft_S(seeds)
    pc = None if per_class is None else np.ascontiguousarray(per_class, np.float32)
    rc = lib().orc_oift_image(_p(I, C.c_int32), xs, ys, zs, adj_r, int(pc is not None), per,
                              _p(pc, C.c_float), -1 if pc is None else pc.size - 1,
                              _p(S, C.c_int32), _p(lab, C.c_int32), _p(cost, C.c_int32), _p(pred, C.c_int8),
                              _p(val, C.c_int32), _p(order, C.c_int32), _p(npop, C.c_int64))
    if rc != 0:
        raise ValueError(f"orc_oift_image failed rc={rc}")
    out = dict(label=lab.reshape(I.shape), cost=cost.reshape(I.shape), pred=pred.reshape(I.shape),
               value=val.reshape(I.shape))
    if want_order:
        out["order"] = order[: int(npop[0])]
    return out


def edge_weights(I, adj_r=1.0, per=0.0):
    """(wa, wb) of shape (E, zs, ys, xs): wa = w_obj(p->p+d), wb = w_obj(p+d->p)."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    E = (spheric(adj_r).shape[0] - 1) // 2
    wa = np.empty((E,) + I.shape, np.uint32); wb = np.empty((E,) + I.shape, np.uint32)
    rc = lib().orc_edge_weights(_p(I, C.c_int32), xs, ys, zs, adj_r, per, _p(wa, C.c_uint32), _p(wb, C.c_uint32))
    assert rc == 0
    return wa, wb


def oift_edges(wa, wb, seeds, label, adj_r=1.0, pol_sign=None, want_order=False):
    wa = np.ascontiguousarray(wa, np.uint32); wb = np.ascontiguousarray(wb, np.uint32)
    shape = wa.shape[1:]
    zs, ys, xs = shape
    lab = _i32(label).copy()
    n = lab.size
    cost = np.empty(n, np.int32); pred = np.empty(n, np.int8); val = np.empty(n, np.uint32)
    order = np.empty(n, np.int32) if want_order else None
    npop = np.zeros(1, np.int64)
    S = None if seeds is None else _gft_S(seeds)
    ps = None if pol_sign is None else np.ascontiguousarray(pol_sign, np.int8)
    rc = lib().orc_oift_edges(_p(wa, C.c_uint32), _p(wb, C.c_uint32), xs, ys, zs, adj_r, _p(ps, C.c_int8),
                              -1 if ps is None else ps.size - 1, _p(S, C.c_int32), _p(lab, C.c_int32),
                              _p(cost, C.c_int32), _p(pred, C.c_int8), _p(val, C.c_uint32),
                              _p(order, C.c_int32), _p(npop, C.c_int64))
    if rc != 0:
        raise ValueError(f"orc_oift_edges failed rc={rc}")
    out = dict(label=lab.reshape(shape), cost=cost.reshape(shape), pred=pred.reshape(shape),
               value=val.reshape(shape))
    if want_order:
        out["order"] = order[: int(npop[0])]
    return out


def relax(I, seeds, label, niter, adj_r=1.0, per=0.0, per_class=None, fp64=False, dense=False, spacing=None):
    """ORelax_1_Multi[_PerClass] restated.  Returns dict(label, memb[nc, ...], classes).
    spacing = (dx, dy, dz) reaches Dpq (None: unit)."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = _gft_S(seeds)
    pc = None if per_class is None else np.ascontiguousarray(per_class, np.float32)
    nc = np.zeros(1, np.int32); cls = np.zeros(64, np.int32)
    # memberships for at most 64 classes; sized after counting classes below
    present = {0} | {int(lab.ravel()[s]) for s in np.asarray(seeds).ravel() if lab.ravel()[s] >= 0}
    ncl = len(present)
    memb = np.empty((ncl,) + I.shape, np.float64 if fp64 else np.float32)
    fn = lib().orc_relax_f64 if fp64 else lib().orc_relax_f32
    with _spacing(spacing):
        rc = fn(_p(I, C.c_int32), xs, ys, zs, adj_r, int(pc is not None), per, _p(pc, C.c_float),
                -1 if pc is None else pc.size - 1, _p(S, C.c_int32), _p(lab, C.c_int32), int(niter), int(dense),
                _p(memb, C.c_double if fp64 else C.c_float), _p(cls, C.c_int32), _p(nc, C.c_int32))
    if rc != 0:
        raise ValueError(f"orc_relax failed rc={rc}")
    assert nc[0] == ncl
    return dict(label=lab.reshape(I.shape), memb=memb, classes=cls[: nc[0]].copy())


def relax_binary(I, seeds, label, niter, adj_r=1.0, per=0.0, spacing=None):
    I = _i32(I)
    xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = None if seeds is None else _gft_S(seeds)
    memb = np.empty(I.shape, np.float32)
    with _spacing(spacing):
        rc = lib().orc_relax_binary_f32(_p(I, C.c_int32), xs, ys, zs, adj_r, per, _p(S, C.c_int32),
                                        _p(lab, C.c_int32), int(niter), _p(memb, C.c_float))
    assert rc == 0
    return dict(label=lab.reshape(I.shape), memb=memb)


RELAX_MODELS = {"intensity": 0, "grad": 1, "rbf": 2}


def relax_model(I, seeds, label, niter, model, G=None, alpha=0.5, sigma=1.0, adj_r=1.0, per=0.0, fp64=False,
                dense=False, spacing=None):
    """Multi-class relaxation with the GradWeight or Gaussian-RBF weight model restated
    (see orc_relax_model_f32).  Returns dict(label, memb[nc, ...], classes)."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = _gft_S(seeds)
    Gi = None if G is None else _i32(G)
    nc = np.zeros(1, np.int32); cls = np.zeros(64, np.int32)
    present = {0} | {int(lab.ravel()[s]) for s in np.asarray(seeds).ravel() if lab.ravel()[s] >= 0}
    memb = np.empty((len(present),) + I.shape, np.float64 if fp64 else np.float32)
    fn = lib().orc_relax_model_f64 if fp64 else lib().orc_relax_model_f32
    with _spacing(spacing):
        rc = fn(_p(I, C.c_int32), _p(Gi, C.c_int32), xs, ys, zs, adj_r, RELAX_MODELS[model], alpha, sigma, 0, per,
                None, -1, _p(S, C.c_int32), _p(lab, C.c_int32), int(niter), int(dense),
                _p(memb, C.c_double if fp64 else C.c_float), _p(cls, C.c_int32), _p(nc, C.c_int32))
    if rc != 0:
        raise ValueError(f"orc_relax_model failed rc={rc}")
    return dict(label=lab.reshape(I.shape), memb=memb[: nc[0]], classes=cls[: nc[0]].copy())


def relax_binary_w(I, seeds, label, niter, adj_r=1.0, per=0.0, W=None, Wxyz=None, dense=False, spacing=None):
    """ORelax_1 in its image, per-voxel-weight (W) or per-axis (Wxyz) form."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = None if seeds is None else _gft_S(seeds)
    Wi = None if W is None else _i32(W)
    Wx, Wy, Wz = (None, None, None) if Wxyz is None else tuple(_i32(w) for w in Wxyz)
    memb = np.empty(I.shape, np.float32)
    with _spacing(spacing):
        rc = lib().orc_relax_binary_w_f32(_p(I, C.c_int32), _p(Wi, C.c_int32), _p(Wx, C.c_int32),
                                          _p(Wy, C.c_int32), _p(Wz, C.c_int32), xs, ys, zs, adj_r, per,
                                          _p(S, C.c_int32), _p(lab, C.c_int32), int(niter), int(dense),
                                          _p(memb, C.c_float))
    assert rc == 0
    return dict(label=lab.reshape(I.shape), memb=memb)


def gradweight_edges(I, G, adj_r=1.0, alpha=0.5, per=0.0):
    """(wa, wb) of OIFT_Multi_GradWeight in the orc_edge_weights layout."""
    I = _i32(I); Gi = _i32(G)
    xs, ys, zs = _dims(I)
    E = (spheric(adj_r).shape[0] - 1) // 2
    wa = np.empty((E,) + I.shape, np.uint32); wb = np.empty((E,) + I.shape, np.uint32)
    rc = lib().orc_gradweight_edges(_p(I, C.c_int32), _p(Gi, C.c_int32), xs, ys, zs, adj_r, alpha, per,
                                    _p(wa, C.c_uint32), _p(wb, C.c_uint32))
    assert rc == 0
    return wa, wb


def w_edges(I, W=None, Wxyz=None, adj_r=1.0, per=0.0):
    """(wa, wb) of OIFT(W, ...) (base W_p + W_q) or OIFT(Wx, Wy, Wz, ...) (per-axis base,
    Spheric(1.0)) in the orc_edge_weights layout (orc_w_edges)."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    Wi = None if W is None else _i32(W)
    Wx, Wy, Wz = (None, None, None) if Wxyz is None else tuple(_i32(w) for w in Wxyz)
    E = (spheric(1.0 if Wxyz is not None else adj_r).shape[0] - 1) // 2
    wa = np.empty((E,) + I.shape, np.uint32); wb = np.empty((E,) + I.shape, np.uint32)
    rc = lib().orc_w_edges(_p(I, C.c_int32), _p(Wi, C.c_int32), _p(Wx, C.c_int32), _p(Wy, C.c_int32),
                           _p(Wz, C.c_int32), xs, ys, zs, adj_r, per, _p(wa, C.c_uint32), _p(wb, C.c_uint32))
    assert rc == 0
    return wa, wb


def oift_w(I, seeds, label, adj_r=1.0, per=0.0, W=None, Wxyz=None):
    """OIFT(W, A, scn, ...) / OIFT(Wx, Wy, Wz, scn, ...) restated: W-form edge weights + the
    keyed-heap OIFT (label 0 on the transposed weights, label > 0 on the stored ones)."""
    wa, wb = w_edges(I, W, Wxyz, adj_r, per)
    return oift_edges(wa, wb, seeds, label, 1.0 if Wxyz is not None else adj_r)


def oift_gradweight(I, G, seeds, label, adj_r=1.0, alpha=0.5, per=0.0):
    """OIFT_Multi_GradWeight restated: grad-model edge weights + the keyed-heap OIFT."""
    wa, wb = gradweight_edges(I, G, adj_r, alpha, per)
    return oift_edges(wa, wb, seeds, label, adj_r, pol_sign=np.array([-1], np.int8))


def estimate_sigma(I):
    I = _i32(I)
    return float(lib().orc_estimate_sigma(_p(I, C.c_int32), I.size))


def relax_diff(I, seeds, label, lab_obj, niter, scale, sigma, alpha, target, adj_r=1.0, eps=1e-6):
    """FP64 differentiable two-class relaxation (orc_relax_diff_f64): returns dict(loss,
    g_scale, g_sigma, g_alpha, f)."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    lab = _i32(label)
    S = _gft_S(seeds)
    sc = np.ascontiguousarray(scale, np.float64)
    tg = np.ascontiguousarray(target, np.float64)
    gsc = np.empty(I.shape, np.float64)
    f = np.empty(I.shape, np.float64)
    loss, gs, ga = np.zeros(1), np.zeros(1), np.zeros(1)
    rc = lib().orc_relax_diff_f64(_p(I, C.c_int32), xs, ys, zs, adj_r, _p(S, C.c_int32), _p(lab, C.c_int32),
                                  int(lab_obj), int(niter), _p(sc, C.c_double), float(sigma), float(alpha),
                                  _p(tg, C.c_double), float(eps), _p(loss, C.c_double), _p(gsc, C.c_double),
                                  _p(gs, C.c_double), _p(ga, C.c_double), _p(f, C.c_double))
    if rc != 0:
        raise ValueError(f"orc_relax_diff_f64 failed rc={rc}")
    return dict(loss=float(loss[0]), g_scale=gsc, g_sigma=float(gs[0]), g_alpha=float(ga[0]), f=f)


def gauss_kernel(spacing=None):
    v = np.zeros(7, np.float32)
    with _spacing(spacing):
        n = lib().orc_gauss_kernel(_p(v, C.c_float))
    return v[:n]


def gaussian_blur(I, passes=1, spacing=None):
    """gft::Scene32::GaussianBlur applied `passes` times (orc_gaussian_blur)."""
    I = _i32(I)
    xs, ys, zs = _dims(I)
    out = np.empty_like(I)
    with _spacing(spacing):
        rc = lib().orc_gaussian_blur(_p(I, C.c_int32), xs, ys, zs, int(passes), _p(out, C.c_int32))
    assert rc == 0
    return out


def sobel(I):
    """gft::Scene32::SobelFilter (orc_sobel)."""
    I = _i32(I)
    out = np.empty_like(I)
    lib().orc_sobel(_p(I, C.c_int32), *_dims(I), _p(out, C.c_int32))
    return out


def grad_normalize(G, image_max):
    """The experiments' gradient normalisation to [0, image_max] (orc_grad_normalize)."""
    G = _i32(G).copy()
    lib().orc_grad_normalize(_p(G, C.c_int32), G.size, int(image_max))
    return G


def bfs_edt(label, seeds):
    """compute_bfs_edt: 6-connected BFS distance from the seeds whose label is > 0."""
    label, S = _i32(label), _i32(seeds).ravel()
    dist = np.empty_like(label)
    lib().orc_bfs_edt(_p(label, C.c_int32), *_dims(label), _p(S, C.c_int32), S.size, _p(dist, C.c_int32))
    return dist


def boundary_seeds(label, stride):
    """inject_boundary_seeds: returns (new label, appended seed indices)."""
    label = _i32(label).copy()
    out = np.empty(max(label.size, 1), np.int32)
    k = lib().orc_boundary_seeds(_p(label, C.c_int32), *_dims(label), int(stride), _p(out, C.c_int32))
    return label, out[:k].copy()


def boundary_seeds_geodesic(label, stride, edt):
    """inject_boundary_seeds_geodesic with the given BFS distances."""
    label, edt = _i32(label).copy(), _i32(edt)
    out = np.empty(max(label.size, 1), np.int32)
    k = lib().orc_boundary_seeds_geodesic(_p(label, C.c_int32), *_dims(label), int(stride), _p(edt, C.c_int32),
                                          _p(out, C.c_int32))
    return label, out[:k].copy()


def dilation_conditional(I, label, radius=1.0, percentile=50.0):
    """The CLI's dilation_conditional: returns the new label map."""
    I, label = _i32(I), _i32(label).copy()
    n = lib().orc_dilation_conditional(_p(I, C.c_int32), _p(label, C.c_int32), *_dims(I), float(radius),
                                       float(percentile))
    assert n >= 0
    return label


# ---- the compiled reference (only where oracle/_ref was built) -------------------

def ref_oift(I, seeds, label, adj_r=1.0, per=0.0, per_class=None):
    L = ref_lib()
    I = _i32(I); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = None if seeds is None else _gft_S(seeds)
    pc = None if per_class is None else np.ascontiguousarray(per_class, np.float32)
    L.ref_oift(_p(I, C.c_int32), xs, ys, zs, adj_r, int(pc is not None), per, _p(pc, C.c_float),
               -1 if pc is None else pc.size - 1, _p(S, C.c_int32), _p(lab, C.c_int32))
    return lab


def ref_relax(I, seeds, label, niter, adj_r=1.0, per=0.0, per_class=None, spacing=None):
    L = ref_lib()
    I = _i32(I); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = _gft_S(seeds)
    pc = None if per_class is None else np.ascontiguousarray(per_class, np.float32)
    with _spacing(spacing, ref=True):
        L.ref_relax(_p(I, C.c_int32), xs, ys, zs, adj_r, int(pc is not None), per, _p(pc, C.c_float),
                    -1 if pc is None else pc.size - 1, _p(S, C.c_int32), _p(lab, C.c_int32), int(niter))
    return lab


def ref_oift_w(I, seeds, label, adj_r=1.0, per=0.0, W=None, Wxyz=None):
    """The reference's gft::ift::OIFT: (A, scn, ...) with W = Wxyz = None, (W, A, scn, ...)
    or (Wx, Wy, Wz, scn, ...)."""
    L = ref_lib()
    I = _i32(I); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = None if seeds is None else _gft_S(seeds)
    Wi = None if W is None else _i32(W)
    Wx, Wy, Wz = (None, None, None) if Wxyz is None else tuple(_i32(w) for w in Wxyz)
    L.ref_oift_w(_p(I, C.c_int32), _p(Wi, C.c_int32), _p(Wx, C.c_int32), _p(Wy, C.c_int32), _p(Wz, C.c_int32),
                 xs, ys, zs, adj_r, per, _p(S, C.c_int32), _p(lab, C.c_int32))
    return lab


def ref_pipeline(I, seeds, label, niter, adj_r=1.0, per=0.0):
    """Reference OIFT_Multi + ORelax_1_Multi; returns (labels, t_oift_s, t_relax_s)."""
    L = ref_lib()
    I = _i32(I); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = _gft_S(seeds)
    t1 = np.zeros(1); t2 = np.zeros(1)
    L.ref_pipeline(_p(I, C.c_int32), xs, ys, zs, adj_r, per, _p(S, C.c_int32), _p(lab, C.c_int32),
                   int(niter), _p(t1, C.c_double), _p(t2, C.c_double))
    return lab, float(t1[0]), float(t2[0])


def ref_relax_binary(I, seeds, label, niter, adj_r=1.0, per=0.0, W=None, Wxyz=None, spacing=None):
    L = ref_lib()
    I = _i32(I); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    S = _gft_S(np.zeros(0, np.int32) if seeds is None else seeds)
    Wi = None if W is None else _i32(W)
    Wx, Wy, Wz = (None, None, None) if Wxyz is None else tuple(_i32(w) for w in Wxyz)
    with _spacing(spacing, ref=True):
        L.ref_relax_binary(_p(I, C.c_int32), _p(Wi, C.c_int32), _p(Wx, C.c_int32), _p(Wy, C.c_int32),
                           _p(Wz, C.c_int32), xs, ys, zs, adj_r, per, _p(S, C.c_int32), _p(lab, C.c_int32),
                           int(niter))
    return lab


def ref_oift_gradweight(I, G, seeds, label, adj_r=1.0, alpha=0.5, per=0.0):
    L = ref_lib()
    I = _i32(I); Gi = _i32(G); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    L.ref_oift_gradweight(_p(I, C.c_int32), _p(Gi, C.c_int32), xs, ys, zs, adj_r, per, alpha,
                          _p(_gft_S(seeds), C.c_int32), _p(lab, C.c_int32))
    return lab


def ref_relax_gradweight(I, G, seeds, label, niter, adj_r=1.0, alpha=0.5, per=0.0, spacing=None):
    L = ref_lib()
    I = _i32(I); Gi = _i32(G); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    with _spacing(spacing, ref=True):
        L.ref_relax_gradweight(_p(I, C.c_int32), _p(Gi, C.c_int32), xs, ys, zs, adj_r, per, alpha,
                               _p(_gft_S(seeds), C.c_int32), _p(lab, C.c_int32), int(niter))
    return lab


def ref_relax_rbf(I, seeds, label, niter, sigma, adj_r=1.0, per=0.0, spacing=None):
    L = ref_lib()
    I = _i32(I); xs, ys, zs = _dims(I)
    lab = _i32(label).copy()
    with _spacing(spacing, ref=True):
        L.ref_relax_rbf(_p(I, C.c_int32), xs, ys, zs, adj_r, per, sigma, _p(_gft_S(seeds), C.c_int32),
                        _p(lab, C.c_int32), int(niter))
    return lab


def ref_estimate_sigma(I):
    I = _i32(I); xs, ys, zs = _dims(I)
    return float(ref_lib().ref_estimate_sigma(_p(I, C.c_int32), xs, ys, zs))


def ref_gaussian_blur(I, passes=1, spacing=None):
    I = _i32(I); xs, ys, zs = _dims(I)
    out = np.empty_like(I)
    with _spacing(spacing, ref=True):
        ref_lib().ref_gaussian_blur(_p(I, C.c_int32), xs, ys, zs, int(passes), _p(out, C.c_int32))
    return out


def ref_gauss_kernel(spacing=None):
    v = np.zeros(8, np.float32)
    with _spacing(spacing, ref=True):
        n = ref_lib().ref_gauss_kernel(_p(v, C.c_float), 8)
    return v[:n]


def ref_sobel(I):
    I = _i32(I)
    out = np.empty_like(I)
    ref_lib().ref_sobel(_p(I, C.c_int32), *_dims(I), _p(out, C.c_int32))
    return out


def ref_bfs_edt(label, seeds):
    label, S = _i32(label), _gft_S(seeds)
    dist = np.empty_like(label)
    ref_lib().ref_bfs_edt(_p(label, C.c_int32), *_dims(label), _p(S, C.c_int32), _p(dist, C.c_int32))
    return dist


def _ref_inject(fn, label, seeds, stride):
    label, S = _i32(label).copy(), _gft_S(seeds)
    cap = S.size + label.size + 1
    out = np.zeros(cap, np.int32)
    added = fn(_p(label, C.c_int32), *_dims(label), _p(S, C.c_int32), int(stride), _p(out, C.c_int32), cap)
    return label, out[1:1 + out[0]].copy(), added


def ref_boundary_seeds(label, seeds, stride):
    """The CLI's inject_boundary_seeds: (label, full seed list, count added)."""
    return _ref_inject(ref_lib().ref_boundary_seeds, label, seeds, stride)


def ref_boundary_seeds_geodesic(label, seeds, stride):
    """BFS distances from the seeds, then inject_boundary_seeds_geodesic."""
    return _ref_inject(ref_lib().ref_boundary_seeds_geodesic, label, seeds, stride)


def ref_dilation_conditional(I, label, radius=1.0, percentile=50.0):
    I, label = _i32(I), _i32(label).copy()
    ref_lib().ref_dilation_conditional(_p(I, C.c_int32), _p(label, C.c_int32), *_dims(I), float(radius),
                                       float(percentile))
    return label


def ref_cli(args):
    """Runs the reference CLI (oiftrelax.cpp's main) in this process with argv = args."""
    argv = ["oiftrelax"] + [str(a) for a in args]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    return ref_lib().ref_cli_main(len(argv), arr)


def ref_scene_read(path):
    """gft::Scene32::Read -> int32 array (zs, ys, xs)."""
    d = (C.c_int * 3)()
    ref_lib().ref_scene_dims(str(path).encode(), d)
    out = np.empty((d[2], d[1], d[0]), np.int32)
    assert ref_lib().ref_scene_read(str(path).encode(), _p(out, C.c_int32), out.size) == 0
    return out


def ref_scene_write(path, vol):
    vol = _i32(vol)
    ref_lib().ref_scene_write(str(path).encode(), _p(vol, C.c_int32), *_dims(vol))
