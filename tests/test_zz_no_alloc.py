"""GPU: roift_weights_noalloc_dev, the allocation-free form of the weight builder, called
through the C-ABI, equals roift_weights_dev bit for bit (weights, image extrema) on two calls
in a row, the second on a smaller volume (the static scratch keeps the first call's partials
and flag, which the second must not inherit); sizes beyond the scratch are rejected before any
launch."""
import ctypes as C

import numpy as np
import pytest

import roift_amd as R
from roift_amd import phantom as P

pytestmark = pytest.mark.gpu


def _weights(entry, h, shape, img, adj, per, wdtype):
    import torch
    L = R.lib()
    pb, _ = R.make_problem(shape, adj, per)
    E = 3 if adj == 1.0 else 13
    wt = torch.int32 if wdtype == R.U32 else torch.int16
    wa = torch.full((E,) + tuple(shape), 7, dtype=wt, device="cuda")
    wb = torch.full_like(wa, 7)
    mx, mn = C.c_int32(0), C.c_int32(0)
    fn = L.roift_weights_noalloc_dev if entry == "noalloc" else L.roift_weights_dev
    dt = R.I16 if img.dtype == torch.int16 else R.I32
    torch.cuda.synchronize()
    rc = fn(h, C.byref(pb), img.data_ptr(), dt, wa.data_ptr(), wb.data_ptr(), wdtype, C.byref(mx), C.byref(mn))
    torch.cuda.synchronize()
    return rc, wa.cpu().numpy(), wb.cpu().numpy(), mx.value, mn.value


@pytest.fixture(scope="module")
def handle():
    h = C.c_void_p()
    assert R.lib().roift_create(C.byref(h), 0) == R.OK
    yield h
    R.lib().roift_destroy(h)


CASES = [  # (first shape, second smaller shape, adjacency, per, image dtype, weight dtype)
    ((40, 48, 56), (17, 9, 30), 1.0, 50.0, "i16", R.U16),
    ((33, 35, 37), (5, 6, 7), 1.75, -30.0, "i32", R.U32),
    ((1, 300, 280), (1, 70, 3), 1.0, 0.0, "i32", R.U16),
]


@pytest.mark.parametrize("first,second,adj,per,idt,wdt", CASES)
def test_noalloc_equals_weights_dev_twice(handle, first, second, adj, per, idt, wdt):
    import torch
    tdt = torch.int16 if idt == "i16" else torch.int32
    for k, shape in enumerate((first, second)):
        c = P.small_case(shape, 70 + k, per=per, adj_r=adj, top=4095)
        img = torch.from_numpy(c.image.astype(np.int16 if idt == "i16" else np.int32)).cuda()
        print(f"case shape={shape} adj={adj} per={per} {idt} wdtype={wdt} call={k}", flush=True)
        want = _weights("dev", handle, shape, img, adj, per, wdt)
        got = _weights("noalloc", handle, shape, img, adj, per, wdt)
        assert want[0] == R.OK and got[0] == R.OK
        assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
        assert (got[3], got[4]) == (want[3], want[4]) == (int(c.image.max()), int(c.image.min()))
        assert tdt == img.dtype


def test_noalloc_overflow_flag_is_cleared_between_calls(handle):
    """A call whose weights overflow u16 sets the flag in the shared scratch; the next call,
    which fits, must not report it."""
    import torch
    big = torch.from_numpy(np.array([[[0, 60000, 0, 60000]]], np.int32)).cuda()
    rc = _weights("noalloc", handle, (1, 1, 4), big, 1.0, 50.0, R.U16)[0]
    assert rc == R.EUNSUPPORTED
    c = P.small_case((6, 7, 8), 3)
    img = torch.from_numpy(c.image.astype(np.int16)).cuda()
    assert _weights("noalloc", handle, (6, 7, 8), img, 1.0, 50.0, R.U16)[0] == R.OK


def test_noalloc_rejects_volumes_beyond_its_scratch(handle):
    """ceil(1/64) * ceil(8192/4) * ceil(8320/16) = 1,064,960 workgroups > 2^20: rejected before
    any launch (the buffers are real, in case the check did not hold)."""
    import torch
    shape = (8320, 8192, 1)
    img = torch.zeros(shape, dtype=torch.int16, device="cuda")
    rc = _weights("noalloc", handle, shape, img, 1.0, 50.0, R.U16)[0]
    assert rc == R.EUNSUPPORTED
    assert b"2^20" in R.lib().roift_last_error(handle) or b"1048576" in R.lib().roift_last_error(handle)
