"""GPU: the gft-signature drop-in libroift_gft.so, executed.

tests/dropin/libdropin_driver.so is a caller written against the reference's gft C++
signatures (it builds gft's sScene32 / sAdjRel3 records and supplies a test-side
Scene32::Create); it runs in this process through ctypes, so the drop-in's context shares
the GPU with the other tests.  Every result is compared with the fixtures the compiled
reference wrote: OIFT_Multi[_PerClass] and ORelax_1_Multi[_PerClass] labels (tests/golden),
the two-label OIFT and the W / Wx,Wy,Wz OIFT forms (golden/oiftw), ORelax_1 (golden/variants)
and Scene32::GaussianBlur, also at anisotropic spacing (golden/blur, golden/spacing).
Bars: OIFT and blur bit-exact; relaxation labels equal except where the FP64 oracle's top
two memberships are within 2e-5.
"""
import ctypes as C
import glob
import os

import numpy as np
import pytest

import oracle as O
from conftest import golden_cases
from test_oracle_forms import load

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "dropin", "libdropin_driver.so")

_drv = None


def drv():
    global _drv
    if _drv is None:
        assert os.path.exists(DRIVER), "tests/dropin/libdropin_driver.so is not built (make -C tests/dropin)"
        import roift_amd as R
        R.lib()   # torch's HIP runtime first, as for every library here
        L = C.CDLL(DRIVER)
        ip, fp, fl, vp = C.POINTER(C.c_int), C.POINTER(C.c_float), C.c_float, C.c_void_p
        i = C.c_int
        L.drv_oift_multi.argtypes = [ip, i, i, i, fl, fl, fp, i, ip, ip]
        L.drv_relax_multi.argtypes = [ip, i, i, i, fp, fl, fl, fp, i, ip, ip, i]
        L.drv_oift.argtypes = [i, ip, ip, ip, ip, ip, i, i, i, fl, fl, ip, ip]
        L.drv_orelax_1.argtypes = [ip, ip, ip, ip, ip, i, i, i, fp, fl, fl, ip, ip, i]
        L.drv_gaussian_blur.argtypes = [ip, i, i, i, fp, i, ip]
        _drv = L
    return _drv


def _i(a):
    return None if a is None else np.ascontiguousarray(a, np.int32)


def _p(a, t=C.c_int):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def _S(seeds):
    s = np.asarray(seeds, np.int32).ravel()
    return np.concatenate([[s.size], s]).astype(np.int32)


def _near_tie(m64, tol=2e-5):
    s = np.sort(m64, axis=0)
    return (s[-1] - s[-2]) < tol


@pytest.mark.parametrize("case", golden_cases(), ids=lambda c: c["name"])
def test_dropin_oift_and_relax_match_reference(case):
    I = _i(case["image"])
    zs, ys, xs = I.shape
    pc = case["per_class"]
    pcf = None if pc is None else np.ascontiguousarray(pc, np.float32)
    S = _S(case["seeds"])
    lab = _i(case["label0"]).copy()
    print("case", case["name"], I.shape, flush=True)
    assert drv().drv_oift_multi(_p(I), xs, ys, zs, case["adj_r"], case["per"], _p(pcf, C.c_float),
                                -1 if pcf is None else pcf.size - 1, _p(S), _p(lab)) == 0
    assert np.array_equal(lab, case["oift_label"].astype(np.int32))
    oift = lab.copy()
    assert drv().drv_relax_multi(_p(I), xs, ys, zs, None, case["adj_r"], case["per"], _p(pcf, C.c_float),
                                 -1 if pcf is None else pcf.size - 1, _p(S), _p(lab), case["niter"]) == 0
    r64 = O.relax(I, case["seeds"], oift, case["niter"], case["adj_r"], case["per"], pc, fp64=True)
    diff = lab != case["relax_label"].astype(np.int32)
    assert not diff[~_near_tie(r64["memb"])].any()


def test_dropin_seeds_null_uses_every_labelled_voxel():
    """S == NULL: every labelled voxel, in raster order, is a seed (gft_ift.cpp:944-948)."""
    case = golden_cases()[1]
    I = _i(case["image"])
    zs, ys, xs = I.shape
    lab = _i(case["label0"]).copy()
    drv().drv_oift_multi(_p(I), xs, ys, zs, case["adj_r"], case["per"], None, -1, None, _p(lab))
    want = O.oift_image(I, None, case["label0"], case["adj_r"], case["per"])["label"]
    assert np.array_equal(lab, want)


FORMS = {"A": 0, "W": 1, "Wxyz": 2}


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "oiftw", "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_dropin_oift_forms_match_reference(path):
    c = load(path)
    I = _i(c["image"])
    zs, ys, xs = I.shape
    lab = _i(c["label0"]).copy()
    form = str(c["form"])
    W = _i(c.get("W"))
    Wx, Wy, Wz = (_i(c[k]) for k in ("Wx", "Wy", "Wz")) if form == "Wxyz" else (None, None, None)
    print("case", c["name"], form, flush=True)
    assert drv().drv_oift(FORMS[form], _p(I), _p(W), _p(Wx), _p(Wy), _p(Wz), xs, ys, zs, float(c["adj_r"]),
                          float(c["per"]), _p(_S(c["seeds"])), _p(lab)) == 0
    assert np.array_equal(lab, c["oift_label"].astype(np.int32))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "variants", "binary_*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_dropin_orelax_1_matches_reference(path):
    z = np.load(path, allow_pickle=False)
    I = _i(z["image"])
    zs, ys, xs = I.shape
    form = str(z["form"])
    lab = _i(z["oift_label"]).copy()
    W = _i(z["G"]) if form == "W" else None
    Wx, Wy, Wz = (_i(z[k]) for k in ("Wx", "Wy", "Wz")) if form == "Wxyz" else (None, None, None)
    S = _S(np.zeros(0, np.int32) if form == "noseeds" else z["seeds"])
    assert drv().drv_orelax_1(_p(I), _p(W), _p(Wx), _p(Wy), _p(Wz), xs, ys, zs, None, float(z["adj_r"]),
                              float(z["per"]), _p(S), _p(lab), int(z["niter"])) == 0
    assert np.array_equal(lab, z["relax_label"].astype(np.int32))


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "blur", "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_dropin_gaussian_blur_matches_reference(path):
    z = np.load(path, allow_pickle=False)
    I = _i(z["image"])
    zs, ys, xs = I.shape
    for k in (1, 2, 3):
        out = np.empty_like(I)
        assert drv().drv_gaussian_blur(_p(I), xs, ys, zs, None, k, _p(out)) == 0
        assert np.array_equal(out, z[f"out{k}"])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(HERE, "golden", "spacing", "*.npz"))),
                         ids=lambda p: os.path.basename(p)[:-4])
def test_dropin_carries_spacing(path):
    """The scene's dx/dy/dz reach the blur taps and the relaxation's Dpq through the drop-in."""
    c = load(path)
    sp = np.ascontiguousarray(c["spacing"], np.float32)
    J = _i(c["blur_in"])
    zs, ys, xs = J.shape
    out = np.empty_like(J)
    assert drv().drv_gaussian_blur(_p(J), xs, ys, zs, _p(sp, C.c_float), 2, _p(out)) == 0
    assert np.array_equal(out, c["blur2"])
    kind = str(c["kind"])
    if kind not in ("relax", "perclass", "binary"):
        return
    I = _i(c["image"])
    lab = _i(c["oift_label"]).copy()
    S = _S(c["seeds"])
    if kind == "binary":
        assert drv().drv_orelax_1(_p(I), None, None, None, None, xs, ys, zs, _p(sp, C.c_float), float(c["adj_r"]),
                                  float(c["per"]), _p(S), _p(lab), int(c["niter"])) == 0
        assert np.array_equal(lab, c["relax_label"].astype(np.int32))
        return
    pcf = np.ascontiguousarray(c["per_class"], np.float32) if c["per_class"].size else None
    assert drv().drv_relax_multi(_p(I), xs, ys, zs, _p(sp, C.c_float), float(c["adj_r"]), float(c["per"]),
                                 _p(pcf, C.c_float), -1 if pcf is None else pcf.size - 1, _p(S), _p(lab),
                                 int(c["niter"])) == 0
    from test_oracle_forms import oracle_spacing_relax
    m64 = oracle_spacing_relax(c, fp64=True)["memb"]
    diff = lab != c["relax_label"].astype(np.int32)
    assert not diff[~_near_tie(m64)].any()
