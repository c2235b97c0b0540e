"""The streaming-bandwidth probe that bench.py reports next to the roofline
(amr_dev_copy_bandwidth): it runs, and its figures are bandwidths an MI355X
can reach (above 1 TB/s, below the 8 TB/s HBM peak for the copy's read +
write traffic of a buffer far larger than the 256 MB last-level cache)."""
import ctypes

import pytest

import miniamr_amd as M
from conftest import require

pytestmark = pytest.mark.gpu


def test_copy_bandwidth_is_plausible():
    require(M.LIB_PATH, "run __graft_entry__.build()")
    dev = M.load_library()
    c, r = ctypes.c_double(), ctypes.c_double()
    assert dev.amr_dev_copy_bandwidth(1 << 30, 2, ctypes.byref(c), ctypes.byref(r)) == 0
    print("copy %.1f GB/s, read %.1f GB/s" % (c.value, r.value), flush=True)
    assert 1000.0 < c.value < 8000.0
    assert 1000.0 < r.value < 9000.0


def test_copy_bandwidth_rejects_bad_sizes():
    require(M.LIB_PATH, "run __graft_entry__.build()")
    dev = M.load_library()
    c, r = ctypes.c_double(), ctypes.c_double()
    assert dev.amr_dev_copy_bandwidth(1000, 1, ctypes.byref(c), ctypes.byref(r)) == -1
    assert dev.amr_dev_copy_bandwidth(1 << 20, 0, ctypes.byref(c), ctypes.byref(r)) == -1
