"""Host-side checks that need no GPU: the C-ABI library loads and exports every
symbol include/roift.h declares; the gft mirror's adjacency; input handling."""
import os
import re

import numpy as np
import pytest

import roift_amd as R
from roift_amd import gft, phantom as P
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = []
    for h in os.listdir(os.path.join(ROOT, "include")):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            names += re.findall(r"ROIFT_API\s+[\w\s\*]+?\b(\w+)\s*\(", text)
    return sorted(set(names))


def test_library_exports_every_declared_symbol():
    names = declared_symbols()
    assert "roift_oift_fmax_dev" in names and "oift_fmax_gpu_run" in names and "oift_gpu_run" in names
    L = R.lib()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(R.EXPORTS) == set(names)


GFT_SIGNATURES = {
    "gft::ift::OIFT_Multi(gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float, int*, gft::Scene32::sScene32*)",
    "gft::ift::OIFT_Multi_PerClass(gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float const*, int, int*, "
    "gft::Scene32::sScene32*)",
    "gft::ift::ORelax_1_Multi(gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float, int*, gft::Scene32::sScene32*, "
    "int)",
    "gft::ift::ORelax_1_Multi_PerClass(gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float const*, int, int*, "
    "gft::Scene32::sScene32*, int)",
    "gft::ift::ORelax_1(gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float, int*, gft::Scene32::sScene32*, int)",
    "gft::ift::ORelax_1(gft::Scene32::sScene32*, gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float, int*, "
    "gft::Scene32::sScene32*, int)",
    "gft::ift::ORelax_1(gft::Scene32::sScene32*, gft::Scene32::sScene32*, gft::Scene32::sScene32*, "
    "gft::Scene32::sScene32*, float, int*, gft::Scene32::sScene32*, int)",
    "gft::Scene32::GaussianBlur(gft::Scene32::sScene32*)",
    "gft::ift::OIFT(gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float, int*, gft::Scene32::sScene32*)",
    "gft::ift::OIFT(gft::Scene32::sScene32*, gft::AdjRel3::sAdjRel3*, gft::Scene32::sScene32*, float, int*, "
    "gft::Scene32::sScene32*)",
    "gft::ift::OIFT(gft::Scene32::sScene32*, gft::Scene32::sScene32*, gft::Scene32::sScene32*, "
    "gft::Scene32::sScene32*, float, int*, gft::Scene32::sScene32*)",
}


def _defined(path, dynamic=True):
    import subprocess
    out = subprocess.run(["nm", "-C"] + (["-D"] if dynamic else []) + [path], capture_output=True, text=True).stdout
    return {line.split(" T ", 1)[1] for line in out.splitlines() if " T " in line}


def test_gft_dropin_exports_reference_signatures():
    """libroift_gft.so defines the hot-path gft functions (and the ORelax_1 forms) with the reference's mangled
    signatures, so a caller links it ahead of -lgft unchanged."""
    syms = _defined(os.path.join(ROOT, "roift_amd", "libroift_gft.so"))
    assert GFT_SIGNATURES <= syms
    ref_dir = os.path.join(ROOT, "oracle", "_ref", "obj")
    if os.path.exists(os.path.join(ref_dir, "gft_ift.cpp.o")):
        ref = _defined(os.path.join(ref_dir, "gft_ift.cpp.o"), dynamic=False)
        ref |= _defined(os.path.join(ref_dir, "gft_filtering3.cpp.o"), dynamic=False)
        assert GFT_SIGNATURES <= ref


def test_create_without_device_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    import ctypes as C
    h = C.c_void_p()
    assert R.lib().roift_create(C.byref(h), 0) == R.EHIP


def test_gft_mirror_spheric_matches_oracle():
    for r in (1.0, 1.5, 1.75):
        assert np.array_equal(gft.AdjRel3.Spheric(r).d, O.spheric(r))


def test_phantom_is_deterministic_and_seeded_in_gft_order():
    a = P.small_case((8, 9, 10), 3)
    b = P.small_case((8, 9, 10), 3)
    assert np.array_equal(a.image, b.image) and np.array_equal(a.seeds, b.seeds)
    assert (a.seed_labels[:27] == 1).all() and (a.seed_labels[27:] == 0).all()
    assert len(np.unique(a.seeds)) == len(a.seeds)
    c2 = P.plumbing_case_2d()
    assert c2.image.shape == (1, 512, 512) and len(c2.seeds) == 49 + 4 * 511
