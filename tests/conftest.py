import glob
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device); run with -m gpu")


def golden_cases():
    out = []
    for path in sorted(glob.glob(os.path.join(GOLDEN, "*.npz"))):
        z = np.load(path, allow_pickle=False)
        c = {k: z[k] for k in z.files}
        c["name"] = os.path.splitext(os.path.basename(path))[0]
        pc = c["per_class"]
        c["per_class"] = None if pc.size == 0 else pc.astype(np.float32)
        c["adj_r"] = float(c["adj_r"])
        c["per"] = float(c["per"])
        c["niter"] = int(c["niter"])
        lab0 = np.full(c["image"].size, -1, np.int32)
        lab0[c["seeds"]] = c["seed_labels"]
        c["label0"] = lab0.reshape(c["image"].shape)
        out.append(c)
    return out


@pytest.fixture(scope="session")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import roift_amd as R
    return R.Context(0)
