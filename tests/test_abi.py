"""The C ABI boundary (CPU): the built libraries load and export exactly the
entry points include/amr_dev.h declares.  No compute call is made here."""
import ctypes
import os
import re
import subprocess

import pytest

import amrcheck as A
import miniamr_amd as M
from conftest import require


def header_symbols():
    with open(M.INCLUDE) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(amr_dev_\w+)\s*\(", text, re.M)))


def test_header_matches_python_binding():
    assert header_symbols() == sorted(M.ABI_SYMBOLS)


@pytest.mark.parametrize("which", ["product", "oracle"])
def test_library_exports_every_symbol(which):
    path = M.LIB_PATH if which == "product" else A.ORACLE_SO
    require(path, "run __graft_entry__.build()")
    lib = ctypes.CDLL(path)
    for name in header_symbols():
        assert hasattr(lib, name), "%s does not export %s" % (path, name)
    M.load_library(path)   # argtypes for every symbol bind


def test_config_struct_layout():
    assert ctypes.sizeof(M.Config) == 8 * 4
    assert M.FACE_JOB.itemsize == 24 and M.SPLIT_JOB.itemsize == 16


def test_product_fails_loudly_without_a_gpu():
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU driver is present; covered by the GPU tests")
    require(M.LIB_PATH, "run __graft_entry__.build()")
    with pytest.raises(M.AmrError):
        M.Device(4, 4, 4, 1, 8)


def test_driver_binary_links_the_hip_library():
    require(M.EXE_PATH, "run __graft_entry__.build()")
    out = subprocess.run(["ldd", M.EXE_PATH], stdout=subprocess.PIPE).stdout.decode()
    assert "libminiamr_dev.so" in out and "oracle" not in out
