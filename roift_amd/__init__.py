"""roift_amd -- MI355X (gfx950) OIFT under f_max + fuzzy relaxation.

Python host side of libroift.so (include/roift.h).  `Context` drives the
C-ABI on torch device tensors (data resident in HBM) or on numpy host arrays;
`roift_amd.gft` mirrors the reference's gft API (Scene32, AdjRel3.Spheric,
ift.OIFT_Multi, ift.ORelax_1_Multi, ...) on top of it.

The library must be built (`python -c "import __graft_entry__ as g; g.build()"`
or `make -C roift_amd/csrc`); nothing here falls back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libroift.so")

OK, EINVAL, ENOMEM, EHIP, EUNSUPPORTED = 0, -1, -2, -3, -4
I16, I32, U16, U32, F32 = 1, 2, 3, 4, 5
NIL = -1

EXPORTS = (
    "roift_create", "roift_destroy", "roift_last_error", "roift_set_stream", "roift_last_timings",
    "roift_kernel_timings", "roift_oift_profile", "roift_set_param",
    "roift_weights_dev", "roift_oift_fmax_dev", "roift_oift_edges_dev", "roift_relax_dev",
    "roift_pipeline_dev", "roift_oift_fmax", "roift_relax", "oift_fmax_gpu_run", "oift_gpu_run",
    "roift_relax_binary_dev", "roift_oift_gradweight_dev", "roift_relax_gradweight_dev", "roift_relax_rbf_dev",
    "roift_relax_binary", "roift_batch_pipeline_dev", "roift_relax_adjoint_dev",
    "roift_gaussian_blur_dev", "roift_gaussian_blur",
    "roift_sobel_dev", "roift_bfs_edt_dev", "roift_boundary_seeds_dev", "roift_dilation_conditional_dev",
    "roift_io_last_error", "roift_volume_info_read", "roift_volume_read", "roift_volume_write", "roift_seeds_read",
    "roift_polarity_read", "roift_volume_read_dev", "roift_cli_main", "roift_relax_adjoint64_dev",
    "roift_weights_w_dev", "roift_oift_w_dev", "roift_oift_w", "roift_weights_noalloc_dev",
    "roift_batch_graph_info",
)


class RoiftError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"roift error {code}: {msg}")
        self.code = code


class VolumeInfo(C.Structure):
    _fields_ = [("xsize", C.c_int), ("ysize", C.c_int), ("zsize", C.c_int), ("dx", C.c_float), ("dy", C.c_float),
                ("dz", C.c_float), ("stored_dtype", C.c_int), ("format", C.c_int)]


U8, I8 = 6, 7
FMT_SCN, FMT_NIFTI = 1, 2
READ_SHIFT_MIN = 1


class Problem(C.Structure):
    _fields_ = [("xsize", C.c_int), ("ysize", C.c_int), ("zsize", C.c_int), ("adj_radius", C.c_float),
                ("use_per_class", C.c_int), ("per", C.c_float), ("h_per_class", C.POINTER(C.c_float)),
                ("max_label", C.c_int), ("dx", C.c_float), ("dy", C.c_float), ("dz", C.c_float)]


_lib = None


def lib():
    """Load libroift.so (raises if it was not built -- there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with make -C roift_amd/csrc")
        # torch bundles its own HIP runtime (soname libamdhip64.so.7, as ROCm's).  Loaded
        # first, it is the one libroift.so binds to; loaded after libroift.so, it would
        # come in as a second runtime beside ROCm's, and the two do not share devices.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = C.CDLL(LIB_PATH)
        vp, ip, i32p, i8p, dp = C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.POINTER(C.c_int8), C.POINTER(C.c_double)
        pp = C.POINTER(Problem)
        L.roift_create.argtypes = [C.POINTER(vp), ip]
        L.roift_destroy.argtypes = [vp]
        L.roift_destroy.restype = None
        L.roift_last_error.argtypes = [vp]
        L.roift_last_error.restype = C.c_char_p
        L.roift_set_stream.argtypes = [vp, vp]
        L.roift_last_timings.argtypes = [vp, dp]
        L.roift_kernel_timings.argtypes = [vp, dp, C.c_int]
        L.roift_oift_profile.argtypes = [vp, dp, ip]
        L.roift_set_param.argtypes = [vp, ip, ip]
        L.roift_weights_dev.argtypes = [vp, pp, vp, ip, vp, vp, ip, i32p, i32p]
        L.roift_weights_noalloc_dev.argtypes = [vp, pp, vp, ip, vp, vp, ip, i32p, i32p]
        L.roift_oift_fmax_dev.argtypes = [vp, pp, vp, ip, vp, ip, vp, vp]
        L.roift_oift_edges_dev.argtypes = [vp, pp, vp, vp, ip, i8p, vp, ip, vp, vp]
        L.roift_weights_w_dev.argtypes = [vp, pp, vp, ip, vp, vp, vp, vp, vp, vp, ip]
        L.roift_oift_w_dev.argtypes = [vp, pp, vp, ip, vp, vp, vp, vp, vp, ip, vp, vp]
        L.roift_oift_w.argtypes = [vp, pp, vp, vp, vp, vp, vp, vp, ip, vp, vp]
        L.roift_relax_dev.argtypes = [vp, pp, vp, ip, vp, ip, vp, ip, vp, i32p]
        L.roift_pipeline_dev.argtypes = [vp, pp, vp, ip, vp, ip, vp, ip, vp, vp]
        L.roift_oift_fmax.argtypes = [vp, pp, vp, vp, ip, vp, vp]
        L.roift_relax.argtypes = [vp, pp, vp, vp, ip, vp, ip, vp]
        L.oift_fmax_gpu_run.argtypes = [vp, vp, ip, ip, ip, ip, vp, ip, vp, ip]
        L.oift_gpu_run.argtypes = [ip, vp, vp, ip, ip, ip, ip, vp, ip, vp, ip]
        fl, fp = C.c_float, C.POINTER(C.c_float)
        L.roift_relax_binary_dev.argtypes = [vp, pp, vp, ip, vp, vp, vp, vp, vp, ip, vp, ip, vp]
        L.roift_oift_gradweight_dev.argtypes = [vp, pp, vp, ip, vp, fl, vp, ip, vp, vp]
        L.roift_relax_gradweight_dev.argtypes = [vp, pp, vp, ip, vp, fl, vp, ip, vp, ip, vp]
        L.roift_relax_rbf_dev.argtypes = [vp, pp, vp, ip, fl, vp, ip, vp, ip, vp, fp]
        L.roift_relax_binary.argtypes = [vp, pp, vp, vp, vp, vp, vp, vp, ip, vp, ip, vp]
        L.roift_batch_pipeline_dev.argtypes = [vp, ip, i32p, fl, fl, vp, ip, vp, i32p, vp, ip, vp, ip]
        L.roift_batch_graph_info.argtypes = [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(ip), C.c_char_p, ip]
        L.roift_relax_adjoint_dev.argtypes = [vp, pp, vp, ip, vp, ip, vp, ip, ip, ip, vp, fl, fl, vp, fl, dp, vp, dp,
                                              dp, vp]
        dbl = C.c_double
        L.roift_relax_adjoint64_dev.argtypes = [vp, pp, vp, ip, vp, ip, vp, ip, ip, ip, vp, dbl, dbl, vp, dbl, dp, vp,
                                                dp, dp, vp]
        L.roift_gaussian_blur_dev.argtypes = [vp, pp, vp, ip, vp, ip]
        L.roift_gaussian_blur.argtypes = [vp, pp, vp, vp, ip]
        L.roift_io_last_error.argtypes = []
        L.roift_io_last_error.restype = C.c_char_p
        vip = C.POINTER(VolumeInfo)
        L.roift_volume_info_read.argtypes = [C.c_char_p, vip]
        L.roift_volume_read.argtypes = [C.c_char_p, vp, C.c_longlong]
        L.roift_volume_write.argtypes = [C.c_char_p, vp, vip, C.c_char_p]
        L.roift_seeds_read.argtypes = [C.c_char_p, ip, ip, ip, vp, vp, ip, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.roift_polarity_read.argtypes = [C.c_char_p, vp, ip, C.POINTER(C.c_int)]
        L.roift_volume_read_dev.argtypes = [vp, C.c_char_p, vp, ip, ip, C.POINTER(C.c_int), vip]
        L.roift_cli_main.argtypes = [ip, C.POINTER(C.c_char_p)]
        L.roift_sobel_dev.argtypes = [vp, pp, vp, ip, vp, ip]
        L.roift_bfs_edt_dev.argtypes = [vp, pp, vp, vp, ip, vp]
        L.roift_boundary_seeds_dev.argtypes = [vp, pp, vp, ip, ip, vp, ip, vp, ip, C.POINTER(C.c_int)]
        L.roift_dilation_conditional_dev.argtypes = [vp, pp, vp, ip, vp, fl, C.POINTER(C.c_int)]
        _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return {name: hasattr(L, name) for name in EXPORTS}


def _io_check(rc):
    if rc != OK:
        raise RoiftError(rc, lib().roift_io_last_error().decode())


def volume_info(path):
    """Header of a .scn / .nii / .nii.gz / .hdr volume."""
    info = VolumeInfo()
    _io_check(lib().roift_volume_info_read(str(path).encode(), C.byref(info)))
    return info


def read_volume(path):
    """gft::Scene32::Read: int32 array of shape (zs, ys, xs) and the header."""
    info = volume_info(path)
    out = np.empty((info.zsize, info.ysize, info.xsize), np.int32)
    _io_check(lib().roift_volume_read(str(path).encode(), out.ctypes.data, out.size))
    return out, info


def write_volume(path, vol, spacing=(1.0, 1.0, 1.0), like=None):
    """gft::Scene32::Write of an int32 label volume (.scn, .nii, .nii.gz)."""
    vol = np.ascontiguousarray(vol, np.int32)
    zs, ys, xs = vol.shape
    info = VolumeInfo(xs, ys, zs, spacing[0], spacing[1], spacing[2], I32, 0)
    _io_check(lib().roift_volume_write(str(path).encode(), vol.ctypes.data, C.byref(info),
                                       None if like is None else str(like).encode()))


def read_seeds(path, shape):
    """The CLI's seeds file: (voxel indices, labels) in file order."""
    zs, ys, xs = shape
    n, clamped = C.c_int(0), C.c_int(0)
    _io_check(lib().roift_seeds_read(str(path).encode(), xs, ys, zs, None, None, 0, C.byref(n), C.byref(clamped)))
    seeds = np.empty(max(n.value, 1), np.int32)
    labels = np.empty_like(seeds)
    _io_check(lib().roift_seeds_read(str(path).encode(), xs, ys, zs, seeds.ctypes.data, labels.ctypes.data, n.value,
                                     C.byref(n), C.byref(clamped)))
    return seeds[: n.value], labels[: n.value], bool(clamped.value)


def read_polarity(path):
    """The CLI's per-class polarity file, times 100."""
    buf = np.empty(4096, np.float32)
    n = C.c_int(0)
    _io_check(lib().roift_polarity_read(str(path).encode(), buf.ctypes.data, buf.size, C.byref(n)))
    return buf[: n.value].copy()


def cli_main(args):
    """The GPU oiftrelax in this process: roift_cli_main(["oiftrelax", *args]); 0 = success."""
    argv = ["roift_oiftrelax"] + [str(a) for a in args]
    arr = (C.c_char_p * len(argv))(*[a.encode() for a in argv])
    return lib().roift_cli_main(len(argv), arr)


def make_problem(shape, adj_r=1.0, per=0.0, per_class=None, spacing=None):
    """shape = (zs, ys, xs) of a C-order volume; spacing = (dx, dy, dz) (None: unit spacing)."""
    zs, ys, xs = (int(v) for v in shape)
    dx, dy, dz = (1.0, 1.0, 1.0) if spacing is None else (float(v) for v in spacing)
    pb = Problem(xs, ys, zs, float(adj_r), 0, float(per), None, -1, dx, dy, dz)
    keep = None
    if per_class is not None:
        keep = np.ascontiguousarray(per_class, np.float32)
        pb.use_per_class = 1
        pb.h_per_class = keep.ctypes.data_as(C.POINTER(C.c_float))
        pb.max_label = keep.size - 1
    return pb, keep


def _dtype_code(t):
    import torch
    return {torch.int16: I16, torch.int32: I32, torch.uint16: U16 if hasattr(torch, "uint16") else -9,
            torch.float32: F32}.get(t.dtype, None)


class Context:
    """One roift_ctx on one HIP device; not thread-safe (as the C-ABI)."""

    def __init__(self, device: int = 0):
        self._L = lib()
        h = C.c_void_p()
        rc = self._L.roift_create(C.byref(h), int(device))
        if rc != OK:
            raise RoiftError(rc, "roift_create failed")
        self._h = h
        self.device = device

    def close(self):
        """Release the context's device memory and stream (roift_destroy).  Call it, or use
        the context as a `with` block, before the program ends."""
        if getattr(self, "_h", None):
            self._L.roift_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False

    def __del__(self):
        # During interpreter shutdown the HIP runtime (and a profiler's library wrapped
        # around it) may already be torn down: a HIP call from here would run after
        # them.  The process exit releases the device memory anyway.
        if sys is None or sys.is_finalizing():
            self._h = None
            return
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != OK:
            raise RoiftError(rc, self._L.roift_last_error(self._h).decode())

    def timings(self):
        out = (C.c_double * 6)()
        self._check(self._L.roift_last_timings(self._h, out))
        keys = ("weights_ms", "oift_ms", "relax_ms", "labels_ms", "oift_steps", "flood_rounds")
        return dict(zip(keys, list(out)))

    def kernel_timings(self):
        """Device-event ms of the last call's main kernels (weights, OIFT, relaxation loop, blur)."""
        out = (C.c_double * 6)()
        self._check(self._L.roift_kernel_timings(self._h, out, 6))
        return {"weights_ms": out[0], "oift_kernel_ms": out[1], "relax_loop_ms": out[2],
                "relax_launches": int(out[3]), "blur_ms": out[4], "stage_ms": out[5]}

    PARAMS = {"oift_blocks": 1, "flood_budget": 2, "relax_depth": 3, "flood_mode": 4, "relax_tile": 5,
              "blur_planes": 6, "oift_mode": 7, "local_candidates": 8, "local_check_steps": 9, "oift_profile": 10}

    def set_param(self, name, value):
        self._check(self._L.roift_set_param(self._h, self.PARAMS[name], int(value)))

    PHASES = ("level_scan", "collect", "key_collect", "key_rank", "flood_seed", "flood_round", "rsafe",
              "claim", "offers")

    def local_profile(self):
        """The last OIFT in local-step mode (ROIFT_PARAM_OIFT_MODE 1): grid rounds, levels,
        candidate checks, voxels visited by the checks, grid-wide flood rounds, and the
        context's count of runs that met a tie and reran level-synchronously, and the ms the
        lead workgroup spent in each phase (frontier compaction, candidates, checks, levels,
        grid-wide floods; each up to the exit of the phase's grid barrier)."""
        out = (C.c_double * 167)()
        self._check(self._L.roift_oift_profile(self._h, out, 167))
        keys = ("rounds", "levels", "checks", "check_visits", "grid_flood_rounds")
        res = {k: int(out[i]) for i, k in enumerate(keys)}
        for i, k in enumerate(("frontier_ms", "candidates_ms", "checks_ms", "levels_ms", "grid_flood_ms")):
            res[k] = out[5 + i]
        res["max_level_voxels"] = int(out[10])
        res["check_steps"] = int(out[11])      # summed over rounds: the round's longest check
        res["level_steps"] = int(out[12])      # summed over rounds: the round's longest level
        res["tie_fallbacks"] = int(out[166])
        return res

    def oift_profile(self, by_size=False):
        """Per-phase device wall-clock ms and counts of the last OIFT (and, with
        by_size, each phase's ms split by ceil(log2(entries of the step)))."""
        out = (C.c_double * 181)()
        self._check(self._L.roift_oift_profile(self._h, out, 181))
        res = {ph: (round(out[i], 3), int(out[9 + i])) for i, ph in enumerate(self.PHASES)}
        res["flood_items"] = (int(out[162]), 0)
        res["touched_slots"] = (int(out[163]), 0)
        res["flood_batches"] = (int(out[164]), int(out[165]))
        res["flood_critical_batches"] = (int(out[167]), int(out[178]))
        res["flood_batch_hist"] = tuple(int(out[168 + b]) for b in range(10))
        res["flood_wave_hops"] = (int(out[179]), round(out[180], 3))
        if by_size:
            res = {ph: [round(out[18 + 16 * i + b], 2) for b in range(16)] for i, ph in enumerate(self.PHASES)}
        return res

    # ---- device (torch) forms -------------------------------------------------------
    def weights(self, image, adj_r=1.0, per=0.0, wdtype="u16"):
        """image: torch int16/int32 (zs,ys,xs) on the device -> (wa, wb, maxI, minI)."""
        import torch
        pb, _ = make_problem(image.shape, adj_r, per)
        E = {1.0: 3, 1.75: 13}.get(float(adj_r))
        wt = torch.int32 if wdtype == "u32" else torch.int16
        wa = torch.empty((E,) + tuple(image.shape), dtype=wt, device=image.device)
        wb = torch.empty_like(wa)
        mx, mn = C.c_int32(), C.c_int32()
        self._check(self._L.roift_weights_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                              wa.data_ptr(), wb.data_ptr(), U32 if wdtype == "u32" else U16,
                                              C.byref(mx), C.byref(mn)))
        return wa, wb, mx.value, mn.value

    def oift_fmax(self, image, seeds, label, adj_r=1.0, per=0.0, per_class=None, want_cost=True):
        """In place on `label` (torch int32 on the device); returns the cost map or None."""
        import torch
        assert label.dtype == torch.int32 and label.is_contiguous() and image.is_contiguous()
        pb, _keep = make_problem(image.shape, adj_r, per, per_class)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        cost = torch.empty_like(label) if want_cost else None
        self._check(self._L.roift_oift_fmax_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                                seeds.data_ptr(), int(seeds.numel()), label.data_ptr(),
                                                cost.data_ptr() if cost is not None else None))
        return cost

    def oift_w(self, image, seeds, label, adj_r=1.0, per=0.0, W=None, Wxyz=None, want_cost=True):
        """gft::ift::OIFT(W, A, scn, ...) (W: int32 device tensor) or OIFT(Wx, Wy, Wz, scn, ...)
        (Wxyz: three int32 device tensors, Spheric(1.0)) in place on `label`; returns the cost map."""
        import torch
        assert label.dtype == torch.int32 and label.is_contiguous() and image.is_contiguous()
        pb, _ = make_problem(image.shape, 1.0 if Wxyz is not None else adj_r, per)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        cost = torch.empty_like(label) if want_cost else None
        ws = [None if W is None else W.to(torch.int32).contiguous()]
        ws += [None] * 3 if Wxyz is None else [w.to(torch.int32).contiguous() for w in Wxyz]
        ptr = [None if w is None else w.data_ptr() for w in ws]
        self._check(self._L.roift_oift_w_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image), *ptr,
                                             seeds.data_ptr(), int(seeds.numel()), label.data_ptr(),
                                             cost.data_ptr() if cost is not None else None))
        return cost

    def oift_edges(self, wa, wb, seeds, label, adj_r=1.0, pol_sign=None, want_cost=True):
        import torch
        pb, _ = make_problem(label.shape, adj_r)
        ps = None
        if pol_sign is not None:
            ps = np.ascontiguousarray(pol_sign, np.int8)
            pb.max_label = ps.size - 1
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        cost = torch.empty_like(label) if want_cost else None
        wcode = U16 if wa.element_size() == 2 else U32
        self._check(self._L.roift_oift_edges_dev(
            self._h, C.byref(pb), wa.data_ptr(), wb.data_ptr(), wcode,
            None if ps is None else ps.ctypes.data_as(C.POINTER(C.c_int8)), seeds.data_ptr(), int(seeds.numel()),
            label.data_ptr(), cost.data_ptr() if cost is not None else None))
        return cost

    def relax(self, image, seeds, label, niter, adj_r=1.0, per=0.0, per_class=None, want_memb=True, spacing=None):
        """In place on `label`; returns memberships [nclasses, zs, ys, xs] (float32) or None."""
        import torch
        pb, _keep = make_problem(image.shape, adj_r, per, per_class, spacing)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        memb = torch.empty((8,) + tuple(label.shape), dtype=torch.float32, device=label.device) if want_memb else None
        nc = C.c_int32(0)
        self._check(self._L.roift_relax_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                            seeds.data_ptr(), int(seeds.numel()), label.data_ptr(), int(niter),
                                            memb.data_ptr() if memb is not None else None, C.byref(nc)))
        return memb[: nc.value] if memb is not None else None

    def relax_binary(self, image, seeds, label, niter, adj_r=1.0, per=0.0, W=None, Wxyz=None, want_memb=True,
                     spacing=None):
        """ORelax_1 (image, W or Wx/Wy/Wz form), in place on `label` (0/1 out); returns f or None."""
        import torch
        pb, _ = make_problem(image.shape, adj_r, per, spacing=spacing)
        seeds = (torch.zeros(0, dtype=torch.int32, device=label.device) if seeds is None
                 else seeds.to(device=label.device, dtype=torch.int32).contiguous())
        memb = torch.empty(tuple(label.shape), dtype=torch.float32, device=label.device) if want_memb else None
        ws = [None, None, None, None]
        if W is not None:
            ws[0] = W.to(torch.int32).contiguous()
        if Wxyz is not None:
            ws[1:] = [w.to(torch.int32).contiguous() for w in Wxyz]
        ptr = [None if w is None else w.data_ptr() for w in ws]
        self._check(self._L.roift_relax_binary_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                                   ptr[0], ptr[1], ptr[2], ptr[3], seeds.data_ptr(),
                                                   int(seeds.numel()), label.data_ptr(), int(niter),
                                                   memb.data_ptr() if memb is not None else None))
        return memb

    def gaussian_blur(self, image, passes=2, out=None, spacing=None):
        """gft::Scene32::GaussianBlur applied `passes` times to an int16/int32 device volume
        of voxel spacing (dx, dy, dz) (None: unit).  Returns `out` (a new tensor of the input's
        dtype when None; may be `image` itself)."""
        import torch
        image = image.contiguous()
        if out is None:
            out = torch.empty_like(image)
        if out.dtype != image.dtype or out.shape != image.shape or not out.is_contiguous():
            raise ValueError("out must be a contiguous tensor of the image's dtype and shape")
        pb, _ = make_problem(image.shape, spacing=spacing)
        self._check(self._L.roift_gaussian_blur_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                                    out.data_ptr(), int(passes)))
        return out

    def read_volume_dev(self, path, dtype=None, shift_min=False):
        """Reads a volume into a new device tensor (int16 or int32) through pinned staging;
        returns (tensor, info, shift)."""
        import torch
        info = volume_info(path)
        dt = torch.int16 if dtype in (None, torch.int16) else torch.int32
        out = torch.empty((info.zsize, info.ysize, info.xsize), dtype=dt, device=f"cuda:{self.device}")
        shift = C.c_int(0)
        self._check(self._L.roift_volume_read_dev(self._h, str(path).encode(), out.data_ptr(), _dtype_code(out),
                                                  READ_SHIFT_MIN if shift_min else 0, C.byref(shift),
                                                  C.byref(info)))
        return out, info, shift.value

    def gaussian_blur_host(self, image, passes=2, spacing=None):
        img = np.ascontiguousarray(image, np.int32)
        out = np.empty_like(img)
        pb, _ = make_problem(img.shape, spacing=spacing)
        self._check(self._L.roift_gaussian_blur(self._h, C.byref(pb), img.ctypes.data, out.ctypes.data, int(passes)))
        return out

    def sobel(self, image, normalize=True, out=None):
        """gft::Scene32::SobelFilter of an int16/int32 device volume into int32; with
        normalize, scaled to [0, max(image)] as the gradient-weight experiment does."""
        import torch
        image = image.contiguous()
        if out is None:
            out = torch.empty(tuple(image.shape), dtype=torch.int32, device=image.device)
        pb, _ = make_problem(image.shape)
        self._check(self._L.roift_sobel_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                            out.data_ptr(), int(bool(normalize))))
        return out

    def bfs_edt(self, label, seeds):
        """6-connected BFS distance from the seeds whose label is > 0 (INT_MAX unreached)."""
        import torch
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        out = torch.empty_like(label)
        pb, _ = make_problem(label.shape)
        self._check(self._L.roift_bfs_edt_dev(self._h, C.byref(pb), label.data_ptr(), seeds.data_ptr(),
                                              int(seeds.numel()), out.data_ptr()))
        return out

    def boundary_seeds(self, label, stride, geodesic=False, seeds=None, capacity=None):
        """Background seeds on the volume faces (label modified in place); returns the new
        seed indices in raster order, to be appended after the caller's seeds."""
        import torch
        seeds = (torch.zeros(0, dtype=torch.int32, device=label.device) if seeds is None
                 else seeds.to(device=label.device, dtype=torch.int32).contiguous())
        zs, ys, xs = label.shape
        if capacity is None:
            capacity = 2 * (xs * ys + ys * zs + xs * zs) + 8
        out = torch.empty(max(int(capacity), 1), dtype=torch.int32, device=label.device)
        added = C.c_int(0)
        pb, _ = make_problem(label.shape)
        self._check(self._L.roift_boundary_seeds_dev(self._h, C.byref(pb), label.data_ptr(), int(stride),
                                                     int(bool(geodesic)), seeds.data_ptr(), int(seeds.numel()),
                                                     out.data_ptr(), int(capacity), C.byref(added)))
        return out[: added.value]

    def dilation_conditional(self, image, label, percentile=50.0, radius=1.0):
        """The CLI's binary post-process, in place on `label`; returns the number of voxels set."""
        pb, _ = make_problem(image.shape, radius)
        n = C.c_int(0)
        self._check(self._L.roift_dilation_conditional_dev(self._h, C.byref(pb), image.contiguous().data_ptr(),
                                                           _dtype_code(image), label.data_ptr(), float(percentile),
                                                           C.byref(n)))
        return n.value

    def oift_gradweight(self, image, grad, seeds, label, alpha=0.5, adj_r=1.0, per=0.0, want_cost=True):
        """OIFT_Multi_GradWeight in place on `label`; grad: int32 device tensor; returns the cost map."""
        import torch
        pb, _ = make_problem(image.shape, adj_r, per)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        g = grad.to(torch.int32).contiguous()
        cost = torch.empty_like(label) if want_cost else None
        self._check(self._L.roift_oift_gradweight_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                                      g.data_ptr(), float(alpha), seeds.data_ptr(),
                                                      int(seeds.numel()), label.data_ptr(),
                                                      cost.data_ptr() if cost is not None else None))
        return cost

    def relax_gradweight(self, image, grad, seeds, label, niter, alpha=0.5, adj_r=1.0, per=0.0, want_memb=True,
                         spacing=None):
        """ORelax_1_Multi_GradWeight in place on `label`; returns memberships [nclasses, ...] or None."""
        import torch
        pb, _ = make_problem(image.shape, adj_r, per, spacing=spacing)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        g = grad.to(torch.int32).contiguous()
        nc = _nclasses(label, seeds)
        memb = torch.empty((nc,) + tuple(label.shape), dtype=torch.float32, device=label.device) if want_memb else None
        self._check(self._L.roift_relax_gradweight_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                                       g.data_ptr(), float(alpha), seeds.data_ptr(),
                                                       int(seeds.numel()), label.data_ptr(), int(niter),
                                                       memb.data_ptr() if memb is not None else None))
        return memb

    def relax_rbf(self, image, seeds, label, niter, sigma=0.0, adj_r=1.0, want_memb=True, spacing=None):
        """ORelax_1_Multi_GaussRBF in place on `label`; sigma <= 0 uses the image's global std.
        Returns (memberships or None, sigma used)."""
        import torch
        pb, _ = make_problem(image.shape, adj_r, 0.0, spacing=spacing)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        nc = _nclasses(label, seeds)
        memb = torch.empty((nc,) + tuple(label.shape), dtype=torch.float32, device=label.device) if want_memb else None
        used = C.c_float(0.0)
        self._check(self._L.roift_relax_rbf_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                                float(sigma), seeds.data_ptr(), int(seeds.numel()),
                                                label.data_ptr(), int(niter),
                                                memb.data_ptr() if memb is not None else None, C.byref(used)))
        return memb, used.value

    def batch_pipeline(self, images, seeds, labels, niter, per=0.0, use_graph=True, want_cost=True):
        """OIFT_Multi + ORelax_1_Multi over a list of volumes in one batched call.

        images: list of device int16/int32 tensors (zs, ys, xs); seeds: list of int32 device
        tensors (volume-local indices); labels: list of int32 device tensors (in/out).
        Returns the list of cost maps (or None)."""
        import torch
        dev = labels[0].device
        dims = np.array([[t.shape[2], t.shape[1], t.shape[0]] for t in images], np.int32).ravel()
        nseeds = np.array([int(s.numel()) for s in seeds], np.int32)
        # With use_graph the concatenated inputs live in buffers kept for the next call of the
        # same shapes: the library's graph is keyed on their addresses and replays while they hold.
        key = (tuple(dims.tolist()), tuple(nseeds.tolist()), images[0].dtype, bool(want_cost), str(dev))
        bb = getattr(self, "_batch_bufs", None)
        if not use_graph or bb is None or bb[0] != key:
            nvox, nsd = int(sum(t.numel() for t in images)), int(nseeds.sum())
            bb = (key, torch.empty(nvox, dtype=images[0].dtype, device=dev),
                  torch.empty(max(nsd, 1), dtype=torch.int32, device=dev), torch.empty(nvox, dtype=torch.int32, device=dev),
                  torch.empty(nvox, dtype=torch.int32, device=dev) if want_cost else None)
            if use_graph:
                self._batch_bufs = bb
        _, img, S, lab, cost = bb
        torch.cat([t.reshape(-1) for t in images], out=img)
        if int(nseeds.sum()) > 0:
            torch.cat([s.to(device=dev, dtype=torch.int32).reshape(-1) for s in seeds], out=S[:int(nseeds.sum())])
        torch.cat([t.reshape(-1) for t in labels], out=lab)
        torch.cuda.current_stream(dev).synchronize()   # the library works on its own stream
        self._check(self._L.roift_batch_pipeline_dev(
            self._h, len(images), dims.ctypes.data_as(C.POINTER(C.c_int32)), 1.0, float(per), img.data_ptr(),
            _dtype_code(images[0]), S.data_ptr(), nseeds.ctypes.data_as(C.POINTER(C.c_int32)), lab.data_ptr(),
            int(niter), cost.data_ptr() if cost is not None else None, int(bool(use_graph))))
        out, o = [], 0
        for t in labels:
            k = t.numel()
            t.copy_(lab[o:o + k].view(t.shape))
            out.append(None if cost is None else cost[o:o + k].view(t.shape))
            o += k
        return out

    def batch_graph_info(self):
        """The batched pipeline's HIP graph in this context: (captures, replays, eager reruns,
        the last capture error or "")."""
        cap, rep, rer = C.c_int(), C.c_int(), C.c_int()
        err = C.create_string_buffer(512)
        self._check(self._L.roift_batch_graph_info(self._h, C.byref(cap), C.byref(rep), C.byref(rer), err, 512))
        return cap.value, rep.value, rer.value, err.value.decode()

    def relax_adjoint(self, image, seeds, label, lab_obj, niter, scale, sigma, alpha, target, ckpt_every=10,
                      adj_r=1.0, eps=1e-6, want_memb=False, fp64=False):
        """Differentiable relaxation + adjoint: returns (loss, grad_scale, grad_sigma, grad_alpha, f or None);
        fp64 runs states and gradients in double (grad_scale and f are float64 then)."""
        import torch
        pb, _ = make_problem(image.shape, adj_r, 0.0)
        seeds = seeds.to(device=label.device, dtype=torch.int32).contiguous()
        sc = scale.to(torch.float32).contiguous()
        tg = target.to(torch.float32).contiguous()
        rt = torch.float64 if fp64 else torch.float32
        gsc = torch.empty(tuple(sc.shape), dtype=rt, device=sc.device)
        memb = torch.empty(tuple(sc.shape), dtype=rt, device=sc.device) if want_memb else None
        loss, gs, ga = C.c_double(), C.c_double(), C.c_double()
        fn = self._L.roift_relax_adjoint64_dev if fp64 else self._L.roift_relax_adjoint_dev
        self._check(fn(
            self._h, C.byref(pb), image.data_ptr(), _dtype_code(image), seeds.data_ptr(), int(seeds.numel()),
            label.data_ptr(), int(lab_obj), int(niter), int(ckpt_every), sc.data_ptr(), float(sigma), float(alpha),
            tg.data_ptr(), float(eps), C.byref(loss), gsc.data_ptr(), C.byref(gs), C.byref(ga),
            memb.data_ptr() if memb is not None else None))
        return loss.value, gsc, gs.value, ga.value, memb

    def pipeline(self, image, seeds, label, niter, adj_r=1.0, per=0.0, per_class=None, cost=None, memb=None,
                 spacing=None):
        """weights + OIFT + relax + labels on device tensors (label in/out)."""
        pb, _keep = make_problem(image.shape, adj_r, per, per_class, spacing)
        self._check(self._L.roift_pipeline_dev(self._h, C.byref(pb), image.data_ptr(), _dtype_code(image),
                                               seeds.data_ptr(), int(seeds.numel()), label.data_ptr(), int(niter),
                                               cost.data_ptr() if cost is not None else None,
                                               memb.data_ptr() if memb is not None else None))

    # ---- host (numpy) forms ------------------------------------------------------------
    def oift_fmax_host(self, image, seeds, label, adj_r=1.0, per=0.0, per_class=None, want_cost=True):
        img = np.ascontiguousarray(image, np.int32)
        lab = np.ascontiguousarray(label, np.int32).copy()
        S = np.ascontiguousarray(seeds, np.int32)
        cost = np.empty_like(lab) if want_cost else None
        pb, _keep = make_problem(img.shape, adj_r, per, per_class)
        self._check(self._L.roift_oift_fmax(self._h, C.byref(pb), img.ctypes.data, S.ctypes.data, S.size,
                                            lab.ctypes.data, None if cost is None else cost.ctypes.data))
        return lab, cost

    def relax_binary_host(self, image, seeds, label, niter, adj_r=1.0, per=0.0, W=None, Wxyz=None, spacing=None):
        img = np.ascontiguousarray(image, np.int32)
        lab = np.ascontiguousarray(label, np.int32).copy()
        S = np.ascontiguousarray(np.zeros(0) if seeds is None else seeds, np.int32)
        ws = [None if W is None else np.ascontiguousarray(W, np.int32)]
        ws += [None] * 3 if Wxyz is None else [np.ascontiguousarray(w, np.int32) for w in Wxyz]
        memb = np.empty(img.shape, np.float32)
        pb, _ = make_problem(img.shape, adj_r, per, spacing=spacing)
        self._check(self._L.roift_relax_binary(self._h, C.byref(pb), img.ctypes.data,
                                               *[None if w is None else w.ctypes.data for w in ws],
                                               S.ctypes.data, S.size, lab.ctypes.data, int(niter), memb.ctypes.data))
        return lab, memb

    def relax_host(self, image, seeds, label, niter, adj_r=1.0, per=0.0, per_class=None, spacing=None):
        img = np.ascontiguousarray(image, np.int32)
        lab = np.ascontiguousarray(label, np.int32).copy()
        S = np.ascontiguousarray(seeds, np.int32)
        memb = np.empty((8,) + img.shape, np.float32)
        pb, _keep = make_problem(img.shape, adj_r, per, per_class, spacing)
        self._check(self._L.roift_relax(self._h, C.byref(pb), img.ctypes.data, S.ctypes.data, S.size,
                                        lab.ctypes.data, int(niter), memb.ctypes.data))
        return lab, memb


def _nclasses(label, seeds):
    """Number of relaxation classes: distinct seed labels plus 0 (gft_ift.cpp:8684-8720)."""
    import torch
    if seeds.numel() == 0:
        return 1
    lab = label.reshape(-1)[seeds.long()]
    lab = lab[lab >= 0]
    return int(torch.unique(torch.cat([lab, torch.zeros(1, dtype=lab.dtype, device=lab.device)])).numel())


def oift_fmax_gpu_run(h_image, h_labels, xsize, ysize, zsize, per_class, seeds):
    """The reference's legacy C entry point (gpu/oift_fmax_gpu.cu:123-147), h_labels in/out."""
    img = np.ascontiguousarray(h_image, np.int32)
    assert h_labels.dtype == np.int32 and h_labels.flags.c_contiguous
    pc = np.ascontiguousarray(per_class, np.float32)
    S = np.ascontiguousarray(seeds, np.int32)
    return lib().oift_fmax_gpu_run(img.ctypes.data, h_labels.ctypes.data, img.size, xsize, ysize, zsize,
                                   pc.ctypes.data, pc.size - 1, S.ctypes.data, S.size)


def oift_gpu_run(cost_mode, h_image, h_labels, xsize, ysize, zsize, per_class, seeds):
    """The reference's second legacy C entry point (gpu/oift_gpu.cu:149-158), h_labels in/out;
    cost_mode 0 (f_max) only, any other mode returns -1."""
    img = np.ascontiguousarray(h_image, np.int32)
    assert h_labels.dtype == np.int32 and h_labels.flags.c_contiguous
    pc = np.ascontiguousarray(per_class, np.float32)
    S = np.ascontiguousarray(seeds, np.int32)
    return lib().oift_gpu_run(int(cost_mode), img.ctypes.data, h_labels.ctypes.data, img.size, xsize, ysize, zsize,
                              pc.ctypes.data, pc.size - 1, S.ctypes.data, S.size)
