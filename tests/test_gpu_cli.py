"""GPU: roift_oiftrelax (the reference CLI's argv contract, every stage on the GPU)
produces the reference CLI's label volume on the same files, and the device volume
read (pinned staging + on-device conversion) equals the host read."""
import glob
import os
import shutil
import subprocess

import numpy as np
import pytest

import roift_amd as R

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CLI = os.path.join(HERE, "golden", "cli")
EXE = os.path.join(ROOT, "roift_amd", "roift_oiftrelax")
CASES = sorted(os.path.dirname(p) for p in glob.glob(os.path.join(CLI, "*", "argv.txt")))


@pytest.mark.parametrize("case", CASES, ids=os.path.basename)
def test_cli_matches_reference_cli(case, tmp_path):
    """roift_cli_main, the body of roift_oiftrelax, run in this process (one process keeps
    the GPU's cooperative launches; the executable is the same function)."""
    for f in os.listdir(case):
        shutil.copy(os.path.join(case, f), tmp_path)
    argv = open(os.path.join(case, "argv.txt")).read().split()
    print("case", os.path.basename(case), argv, flush=True)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        rc = R.cli_main(argv)
    finally:
        os.chdir(cwd)
    assert rc == 0
    got, _ = R.read_volume(tmp_path / argv[5])
    want = np.load(os.path.join(case, "expected.npz"))["label"]
    assert np.array_equal(got, want)


def test_cli_errors_return_one(tmp_path):
    shutil.copy(os.path.join(CLI, "binary_scn", "vol.scn"), tmp_path)
    assert R.cli_main([str(tmp_path / "vol.scn"), str(tmp_path / "missing.txt"), 0.5, 1, 50,
                       str(tmp_path / "out.scn")]) == 1
    assert R.cli_main([str(tmp_path / "missing.scn"), "s.txt", 0.5, 1, 50, str(tmp_path / "out.scn")]) == 1


def test_cli_executable_usage():
    assert os.path.exists(EXE), "roift_oiftrelax is not built (make -C roift_amd/csrc)"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "usage" in r.stdout


@pytest.mark.parametrize("case", CASES, ids=os.path.basename)
def test_device_read_equals_host_read(ctx, case):
    import torch
    vol = [f for f in os.listdir(case) if f.startswith("vol.")][0]
    path = os.path.join(case, vol)
    host, _ = R.read_volume(path)
    for dt in (torch.int16, torch.int32):
        for shift in (False, True):
            d, info, s = ctx.read_volume_dev(path, dtype=dt, shift_min=shift)
            want = host - min(0, int(host.min())) if shift else host
            assert s == (max(0, -int(host.min())) if shift else 0)
            assert np.array_equal(d.cpu().numpy().astype(np.int32), want)


def test_device_read_large_multi_chunk(ctx, tmp_path):
    """A volume of four 8 MiB staging chunks (both pinned buffers reused)."""
    rng = np.random.default_rng(9)
    vol = rng.integers(-2000, 30000, size=(160, 192, 256)).astype(np.int32)   # 30 MiB of int32 samples
    R.write_volume(tmp_path / "big.nii.gz", vol)
    d, _, s = ctx.read_volume_dev(tmp_path / "big.nii.gz", dtype=None, shift_min=True)
    assert s == -int(vol.min())
    assert np.array_equal(d.cpu().numpy().astype(np.int32), vol + s)
    import torch
    d32, _, _ = ctx.read_volume_dev(tmp_path / "big.nii.gz", dtype=torch.int32)
    assert np.array_equal(d32.cpu().numpy(), vol)
