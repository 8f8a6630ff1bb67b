"""Volume, seed and polarity files (no GPU): libroift.so's readers and writers against
files written by the compiled reference's gft::Scene32::Write, against the reference's
own reader where it is built, and against NIfTI files written here with `struct` from
the published header layout (every datatype gft accepts, both byte orders, .hdr/.img)."""
import glob
import gzip
import os
import struct

import numpy as np
import pytest

import oracle as O
import roift_amd as R
from roift_amd import phantom as P

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(HERE, "golden", "cli")
HAVE_REF = O.ref_available()


def _phantom():
    return P.small_case((24, 28, 32), 5, per=50.0).image.astype(np.int32)


def test_reads_files_written_by_the_reference():
    img = _phantom()
    v, info = R.read_volume(os.path.join(CLI, "binary_scn", "vol.scn"))
    assert np.array_equal(v, img) and info.format == R.FMT_SCN and info.stored_dtype == R.U16
    v, info = R.read_volume(os.path.join(CLI, "negative_nii", "vol.nii.gz"))
    assert np.array_equal(v, img - 500) and info.format == R.FMT_NIFTI and info.stored_dtype == R.I32
    assert (info.xsize, info.ysize, info.zsize) == (32, 28, 24)


@pytest.mark.parametrize("ext", [".scn", ".nii", ".nii.gz"])
@pytest.mark.parametrize("top", [200, 60000, 3_000_000])
def test_write_read_round_trip(tmp_path, ext, top):
    rng = np.random.default_rng(top)
    vol = rng.integers(0, top, size=(5, 6, 7)).astype(np.int32)
    path = tmp_path / ("v" + ext)
    R.write_volume(path, vol, spacing=(0.5, 1.0, 2.0))
    got, info = R.read_volume(path)
    assert np.array_equal(got, vol)
    assert (info.dx, info.dy, info.dz) == (0.5, 1.0, 2.0)
    if ext == ".scn":
        assert info.stored_dtype == (R.U8 if top < 256 else (R.U16 if top < 65536 else R.I32))
    if HAVE_REF:
        assert np.array_equal(O.ref_scene_read(path), vol)


@pytest.mark.skipif(not HAVE_REF, reason="oracle/_ref not built")
def test_scn_narrow_samples_wrap_like_gft(tmp_path):
    """gft writes 8/16-bit SCN samples by casting (negative values wrap); so do we."""
    vol = np.array([[[-3, 5, 250], [7, 1, 0]]], np.int32)
    ours, theirs = tmp_path / "a.scn", tmp_path / "b.scn"
    R.write_volume(ours, vol)
    O.ref_scene_write(theirs, vol)
    assert open(ours, "rb").read() == open(theirs, "rb").read()


NII_CODES = {np.uint8: 2, np.int16: 4, np.int32: 8, np.float32: 16, np.int8: 256, np.uint16: 512}


def _nifti_bytes(vol, dtype, big=False, vox_offset=352, ndim=3):
    e = ">" if big else "<"
    zs, ys, xs = vol.shape
    h = bytearray(348)
    struct.pack_into(e + "i", h, 0, 348)
    struct.pack_into(e + "8h", h, 40, ndim, xs, ys, zs, 1, 1, 1, 1)
    struct.pack_into(e + "h", h, 70, NII_CODES[dtype])
    struct.pack_into(e + "h", h, 72, np.dtype(dtype).itemsize * 8)
    struct.pack_into(e + "8f", h, 76, 1.0, 0.8, 0.9, 1.1, 1, 1, 1, 1)
    struct.pack_into(e + "f", h, 108, float(vox_offset))
    struct.pack_into(e + "f", h, 112, 1.0)
    h[344:348] = b"n+1\0"
    data = np.ascontiguousarray(vol.astype(np.dtype(dtype).newbyteorder(e))).tobytes()
    return bytes(h) + b"\0" * (vox_offset - 348) + data


def _gft_float_read(v):
    v = v.astype(np.float32)
    mx = np.float32(max(0.0, float(v.max())))
    f = np.float32(10000.0 / float(mx))
    x = (v * f).astype(np.float32).astype(np.float64)
    return np.where(x < 0, np.trunc(x - 0.5), np.trunc(x + 0.5)).astype(np.int32)


@pytest.mark.parametrize("dtype", list(NII_CODES), ids=lambda t: np.dtype(t).name)
@pytest.mark.parametrize("big", [False, True])
@pytest.mark.parametrize("gz", [False, True])
def test_nifti_types_and_byte_orders(tmp_path, dtype, big, gz):
    rng = np.random.default_rng(3)
    info = np.iinfo(dtype) if np.dtype(dtype).kind in "iu" else None
    if info is not None:
        vol = rng.integers(max(info.min, -30000), min(info.max, 60000), size=(4, 5, 6)).astype(dtype)
    else:
        vol = (rng.random((4, 5, 6)) * 900 - 100).astype(np.float32)
    raw = _nifti_bytes(vol, dtype, big=big)
    path = tmp_path / ("v.nii.gz" if gz else "v.nii")
    if gz:
        with gzip.open(path, "wb") as f:
            f.write(raw)
    else:
        path.write_bytes(raw)
    got, hdr = R.read_volume(path)
    want = _gft_float_read(vol) if info is None else vol.astype(np.int32)
    assert np.array_equal(got, want)
    assert abs(hdr.dx - 0.8) < 1e-6 and abs(hdr.dz - 1.1) < 1e-6
    if HAVE_REF and not big:
        assert np.array_equal(O.ref_scene_read(path), got)


def test_nifti_pair_and_4d_first_volume(tmp_path):
    vol = np.arange(60, dtype=np.int16).reshape(3, 4, 5)
    raw = _nifti_bytes(vol, np.int16, vox_offset=352)
    hdr = bytearray(raw[:348])
    hdr[344:348] = b"ni1\0"
    struct.pack_into("<f", hdr, 108, 0.0)
    (tmp_path / "p.hdr").write_bytes(bytes(hdr))
    (tmp_path / "p.img").write_bytes(raw[352:])
    assert np.array_equal(R.read_volume(tmp_path / "p.hdr")[0], vol)
    assert np.array_equal(R.read_volume(tmp_path / "p.img")[0], vol)
    four = _nifti_bytes(np.concatenate([vol, vol + 100]), np.int16, ndim=4)
    four = bytearray(four)
    struct.pack_into("<8h", four, 40, 4, 5, 4, 3, 2, 1, 1, 1)
    (tmp_path / "f.nii").write_bytes(bytes(four))
    assert np.array_equal(R.read_volume(tmp_path / "f.nii")[0], vol)


def test_rejects_unsupported_and_truncated(tmp_path):
    vol = np.zeros((2, 3, 4), np.int16)
    raw = bytearray(_nifti_bytes(vol, np.int16))
    struct.pack_into("<h", raw, 70, 64)   # float64
    (tmp_path / "d.nii").write_bytes(bytes(raw))
    with pytest.raises(R.RoiftError):
        R.read_volume(tmp_path / "d.nii")
    (tmp_path / "t.nii").write_bytes(_nifti_bytes(vol, np.int16)[:360])
    with pytest.raises(R.RoiftError):
        R.read_volume(tmp_path / "t.nii")
    with pytest.raises(R.RoiftError):
        R.read_volume(tmp_path / "missing.scn")
    with pytest.raises(R.RoiftError):
        R.read_volume(tmp_path / "x.dcm")


def test_label_nifti_keeps_the_input_header(tmp_path):
    vol = np.zeros((3, 4, 5), np.int16)
    raw = bytearray(_nifti_bytes(vol, np.int16))
    raw[148:160] = b"scanner-xyz\0"   # descrip
    (tmp_path / "in.nii").write_bytes(bytes(raw))
    lab = np.arange(60, dtype=np.int32).reshape(3, 4, 5)
    R.write_volume(tmp_path / "out.nii.gz", lab, spacing=(0.8, 0.9, 1.1), like=tmp_path / "in.nii")
    out = gzip.open(tmp_path / "out.nii.gz").read()
    assert out[148:159] == b"scanner-xyz" and struct.unpack_from("<h", out, 70)[0] == 8
    assert np.array_equal(R.read_volume(tmp_path / "out.nii.gz")[0], lab)


def test_seeds_and_polarity_files(tmp_path):
    (tmp_path / "s.txt").write_text("5\n1 2 3 0 1\n9 9 9 0 1\n0 0 0 7 -2\n3 1 0 0 2\n")
    seeds, labels, clamped = R.read_seeds(tmp_path / "s.txt", (4, 5, 6))
    assert seeds.tolist() == [1 + 2 * 6 + 3 * 30, 0, 3 + 6] and labels.tolist() == [1, 0, 2] and clamped
    (tmp_path / "p.txt").write_text("3\n-0.5 0.25\n1.0\n")
    assert R.read_polarity(tmp_path / "p.txt").tolist() == [np.float32(-0.5) * np.float32(100), 25.0, 100.0]
    (tmp_path / "bad.txt").write_text("3\n0.1 x\n")
    with pytest.raises(R.RoiftError):
        R.read_polarity(tmp_path / "bad.txt")


def test_cli_fixtures_present():
    cases = sorted(glob.glob(os.path.join(CLI, "*", "argv.txt")))
    assert len(cases) == 6


@pytest.mark.parametrize("name,spacing", [("aniso_nii", (0.7, 0.9, 2.0)), ("aniso_scn", (1.5, 1.0, 0.6))])
def test_cli_anisotropic_fixtures_carry_their_spacing(name, spacing):
    """The two anisotropic CLI volumes hold the spacing they were written with (the CLI's
    blur taps and relaxation Dpq read it from the file)."""
    vol = [f for f in os.listdir(os.path.join(CLI, name)) if f.startswith("vol.")][0]
    info = R.volume_info(os.path.join(CLI, name, vol))
    assert np.allclose((info.dx, info.dy, info.dz), spacing, atol=1e-6)
